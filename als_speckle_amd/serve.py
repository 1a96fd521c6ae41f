"""Continuous-batching reconstruction server on one GPU.

Requests of different frame sizes are admitted into resident slots as slots
free up, their iterates stay on the device, and every slot of a size class
advances in the same launches (``spk_iterate`` with per-frame schedules: a
request admitted mid-way runs its own position in its own schedule, idle
slots are skipped).  Each size class has its own stream so classes overlap on
the GPU.  Admission is bounded by an HBM budget: the slot pools are sized
from it up front and a request waits for a free slot of its size.

A request's result does not depend on what else is resident: frames never
share a reduction, and a slot's kernels are the same code with the same
inputs whether its neighbours are busy or idle (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np
import torch

from . import _capi
from .device import Context, Plan, plan_bytes, require_gpu
from .phasing import expand_order, shrinkwrap_schedule

__all__ = ["Request", "Result", "Server", "block_table", "admission_order"]


@dataclasses.dataclass
class Request:
    """One reconstruction.  Arrays may be numpy or device tensors."""
    rid: int
    modulus: object            # (n, n) float32, corner-centred
    support: object            # (n, n) float32, max 1
    guess: object              # (n, n) complex64 start (multiplied by the support)
    iterations: int = 500
    order: tuple = (("hio", 90), ("er", 10))
    beta: float = 0.8
    semantics: str = "numpy"
    shrinkwrap: object = None  # None, True or dict(every, sigma0, sigma1, tau)
    arrival: float = 0.0       # seconds on the server clock

    @property
    def n(self):
        return int(self.modulus.shape[0])


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    psi: torch.Tensor          # (n, n) complex64, device
    support: torch.Tensor      # (n, n) float32, device
    trace: np.ndarray          # (iterations,) FP64 modulus error
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


def _steps_for(req: Request) -> np.ndarray:
    """The request's whole schedule as a step table (one entry per iteration)."""
    sem = _capi.SEM_CODES[req.semantics]
    sw = {}
    if req.shrinkwrap:
        kw = dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)
        if isinstance(req.shrinkwrap, dict):
            kw.update(req.shrinkwrap)
        sw = shrinkwrap_schedule(req.iterations, **kw)
    return _capi.step_table([(_capi.ALG_CODES[a], sem, req.beta, *sw.get(k, (0.0, 0.0)))
                             for k, a in enumerate(expand_order(req.order, req.iterations))])


def block_table(active, block: int, slots: int) -> tuple[np.ndarray, dict]:
    """The (block, slots) step table of one launch round and the steps each
    slot takes: active maps slot -> (schedule, position); each active slot's
    column is the next stretch of its own schedule, idle entries are NOOP."""
    table = np.zeros((block, slots), _capi.STEP_DTYPE)
    table["alg"] = _capi.NOOP
    takes = {}
    for f, (sched, pos) in active.items():
        take = min(block, len(sched) - pos)
        table[:take, f] = sched[pos:pos + take]
        takes[f] = take
    return table, takes


def admission_order(queue, free, now: float):
    """Split the waiting requests (in submission order) into those admitted
    now and those still waiting: a request is admitted when it has arrived and
    its size class has a free slot; free maps size -> free slot count and is
    consumed.  Classes never block each other; within a class the order is
    first come, first served."""
    admitted, waiting = [], collections.deque()
    for r in queue:
        if r.arrival <= now and free.get(r.n, 0) > 0:
            free[r.n] -= 1
            admitted.append(r)
        else:
            waiting.append(r)
    return admitted, waiting


class _Pool:
    """Resident slots of one frame size on their own stream."""

    def __init__(self, device_index, n, slots, autotune=False):
        self.n, self.slots = n, slots
        self.stream = torch.cuda.Stream(device=device_index)
        self.ctx = Context(device_index, stream=self.stream)
        tile = ept = 0
        persist = False
        if autotune:
            from .tune import autotune as tune_fn
            with torch.cuda.stream(self.stream):
                r = tune_fn(self.ctx, n, slots, shared=False)
            tile, ept, persist = r.tile, r.ept, r.persist
        self.plan = Plan(self.ctx, n, batch=slots, shared=False, tile=tile, ept=ept,
                         persist=persist)
        self.free = list(range(slots))
        self.active = {}          # slot -> [request, steps, pos, trace chunks, admitted]
        self.event = None

    def admit(self, req: Request, now: float):
        slot = self.free.pop(0)
        dev = self.ctx.device
        with torch.cuda.stream(self.stream):
            m = _as_dev(req.modulus, dev, torch.float32)
            s = _as_dev(req.support, dev, torch.float32)
            g = _as_dev(req.guess, dev, torch.complex64)
            self.plan.set_modulus(m, slot)
            self.plan.set_support(s, slot)
            self.plan.set_psi(g, slot, mult_support=True)
        self.active[slot] = [req, _steps_for(req), 0, [], now]

    def launch(self, block: int):
        """Queue the next `block` iterations of every active slot."""
        if not self.active:
            return False
        table, takes = block_table({f: (st[1], st[2]) for f, st in self.active.items()},
                                   block, self.slots)
        with torch.cuda.stream(self.stream):
            err = self.plan.iterate(table, per_frame=True)
        for f, st in self.active.items():
            st[3].append((err, f, takes[f]))
            st[2] += takes[f]
        self.event = torch.cuda.Event()
        self.event.record(self.stream)
        return True

    def retire(self, now: float):
        """Results of slots whose schedule is complete (after the event)."""
        done = []
        for f in [f for f, st in self.active.items() if st[2] >= len(st[1])]:
            req, _, _, chunks, admitted = self.active.pop(f)
            with torch.cuda.stream(self.stream):
                psi = self.plan.get_psi(f)
                sup = self.plan.get_support(f)
            trace = torch.cat([e[:t, fr] for e, fr, t in chunks]).cpu().numpy() \
                if chunks else np.zeros(0)
            done.append(Result(req.rid, req.n, psi, sup, trace, req.arrival, admitted, now))
            self.free.append(f)
            self.free.sort()
        return done


def _as_dev(x, dev, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=dev, dtype=dtype)


class Server:
    """HBM-budgeted continuous batching over size classes.

    slots: {n: slot count}; or budget_bytes plus mix {n: request fraction},
    which sizes each pool in proportion to its expected work (fraction x n^2)
    within the budget.  block: iterations per launch round (also the
    admission granularity); the default 21 is prime to the shrinkwrap
    interval of 20, so that a request's shrinkwraps mostly fall inside a
    launch, where they run fused into the passes (a shrinkwrap on the last
    step of a launch runs as stand-alone kernels).  autotune: pick each
    pool's FFT layout, radix schedule and launch mode by measurement
    (tune.autotune).
    """

    def __init__(self, device=0, slots=None, budget_bytes=None, mix=None, block=21,
                 autotune=False):
        require_gpu()
        self.device_index = int(device)
        self.block = int(block)
        if slots is None:
            assert budget_bytes and mix, "give slots or budget_bytes and mix"
            slots = self.size_slots(budget_bytes, mix)
        self.pools = {int(n): _Pool(self.device_index, int(n), int(k), autotune)
                      for n, k in sorted(slots.items()) if k > 0}
        self.queue = collections.deque()
        self.results = []
        self.t0 = time.perf_counter()

    @staticmethod
    def size_slots(budget_bytes, mix, max_slots=256):
        """Slots per size within the HBM budget, proportional to expected work."""
        work = {n: f * n * n for n, f in mix.items()}
        tot = sum(work.values())
        out = {}
        for n, w in work.items():
            share = budget_bytes * w / tot
            per = plan_bytes(n, 1, False)
            out[n] = int(max(1, min(max_slots, share // per)))
        return out

    def bytes_reserved(self):
        return sum(plan_bytes(p.n, p.slots, False) for p in self.pools.values())

    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, req: Request):
        assert req.n in self.pools, f"no slots for frame size {req.n}"
        self.queue.append(req)

    def cancel(self, rid: int) -> bool:
        """Drop a waiting or running request (no result is produced).  A
        running request's slot is freed at once: its frame is not in any
        later launch and the next admission overwrites it.  Returns whether
        the request was found."""
        for r in list(self.queue):
            if r.rid == rid:
                self.queue.remove(r)
                return True
        for p in self.pools.values():
            for f, st in list(p.active.items()):
                if st[0].rid == rid:
                    if p.event is not None:   # its last launch may still run
                        p.event.synchronize()
                    del p.active[f]
                    p.free.append(f)
                    p.free.sort()
                    return True
        return False

    def _admit(self, now):
        free = {n: len(p.free) for n, p in self.pools.items()}
        admitted, self.queue = admission_order(self.queue, free, now)
        for r in admitted:
            self.pools[r.n].admit(r, now)

    def tick(self):
        """Admit, launch one block on every busy pool, wait, retire (lockstep)."""
        now = self.now()
        self._admit(now)
        launched = [p for p in self.pools.values() if p.launch(self.block)]
        for p in launched:
            p.event.synchronize()
        now = self.now()
        out = []
        for p in launched:
            out.extend(p.retire(now))
            p.event = None
        self.results.extend(out)
        return out

    def busy(self):
        return bool(self.queue) or any(p.active for p in self.pools.values())

    def poll(self):
        """One pass of the asynchronous loop: each pool whose block finished is
        retired and relaunched on its own, so a size class never waits for
        another.  Returns True if anything progressed."""
        progressed = False
        now = self.now()
        self._admit(now)
        for p in self.pools.values():
            if p.event is not None:
                if not p.event.query():
                    continue
                now = self.now()
                self.results.extend(p.retire(now))
                p.event = None
                self._admit(now)
                progressed = True
            if p.active and p.launch(self.block):
                progressed = True
        return progressed

    def run(self, requests, clock_start=None, lockstep=False):
        """Serve requests (arrival times relative to the start); returns results."""
        self.t0 = time.perf_counter() if clock_start is None else clock_start
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        while self.busy() or any(p.event is not None for p in self.pools.values()):
            if lockstep:
                if not any(p.active for p in self.pools.values()) and self.queue:
                    wait = min(r.arrival for r in self.queue) - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                self.tick()
                continue
            if not self.poll():
                idle = all(p.event is None and not p.active for p in self.pools.values())
                if idle and self.queue:
                    wait = min(r.arrival for r in self.queue) - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 0.005))
                else:
                    time.sleep(20e-6)
        return self.results
