"""Continuous-batching reconstruction server on one GPU.

Requests of different frame sizes are admitted into resident slots as slots
and memory free up, their iterates stay on the device, and every active slot
of a chunk advances in the same launches (``spk_iterate`` with per-frame
schedules: a request admitted mid-way runs its own position in its own
schedule, idle slots are skipped).  Each size class has its own stream, so
classes overlap on the GPU.

Memory: slots live in *chunks* (one plan of ``chunk_slots[n]`` frames of n x n
each).  With a byte budget, the size classes share it: a class allocates a
chunk when a request finds no free slot and the budget has room; when it has
none, chunks of other classes that hold no running request are released
(oldest idle first) to make room.  With fixed ``slots`` (the static form)
each class gets one chunk of that many slots up front and never grows.

A request's result does not depend on what else is resident: frames never
share a reduction, every chunk of a class has the same shape (slot count,
hence the same kernels), and a slot's kernels are the same code with the same
inputs whether its neighbours are busy or idle (tests/test_gpu_serve.py).

The reference has no server: it runs trials one after another from Python
(``examples/basic_phasing_example.py:27-30``).
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np
import torch

from . import _capi
from .device import Context, Group, Plan, group_shape, plan_bytes, require_gpu
from .phasing import expand_order, shrinkwrap_schedule

__all__ = ["Request", "Result", "Server", "block_table", "plan_admission", "chunk_shapes"]

# plan options a serving chunk is sized for: in-loop shrinkwrap work maps and
# the error partials of one launch round (DM adds its modified estimate)
CHUNK_OPTIONS = _capi.PB_SHRINKWRAP | _capi.PB_TRACE

# frame sizes the mixed-size launch takes (spk_group_shape)
MERGED_SIZES = (256, 512, 1024)


def group_tile_ok(n: int) -> bool:
    return int(n) in MERGED_SIZES


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


def chunk_shapes(sizes, budget_bytes, target_bytes=1 << 30, max_slots=64, options=CHUNK_OPTIONS):
    """Slots per chunk and bytes per chunk for each frame size: about
    ``target_bytes`` of slots (at most ``max_slots``, at least one, and no
    more than an eighth of the budget where one slot allows)."""
    target = max(1, min(int(target_bytes), int(budget_bytes) // 8))
    slots, byts = {}, {}
    for n in sizes:
        one = plan_bytes(n, 1, False, options)
        k = int(max(1, min(max_slots, target // one)))
        slots[n] = k
        byts[n] = plan_bytes(n, k, False, options)
    return slots, byts


@dataclasses.dataclass
class Admission:
    """Decisions of one admission round (plan_admission)."""
    admitted: list            # requests to place now, in order
    waiting: collections.deque
    allocate: dict            # size -> chunks to allocate
    release: list             # chunk keys to release first


def plan_admission(queue, now, free, idle, chunk_slots, chunk_bytes, budget, in_use,
                   growable=True) -> Admission:
    """Which waiting requests are admitted now, and the chunk changes that needs.

    queue: requests in submission order.  free: size -> free slots in the
    allocated chunks.  idle: keys (size, id) of allocated chunks that hold no
    running request, oldest idle first.  A request is admitted when it has
    arrived and (1) its size class has a free slot, or (2) a new chunk of its
    size fits in the budget, or (3) releasing idle chunks of *other* sizes,
    oldest first, makes it fit.  Classes never block each other: a request
    that cannot be placed waits while later requests of other sizes go
    ahead; within a size the order is first come, first served.  Pure
    function of its arguments (free is copied).
    """
    free = dict(free)
    idle = list(idle)
    admitted, waiting = [], collections.deque()
    allocate = collections.Counter()
    release = []
    blocked = set()                  # sizes that already failed this round
    for r in queue:
        n = r.n
        if r.arrival > now or n in blocked:
            waiting.append(r)
            continue
        if free.get(n, 0) > 0:
            free[n] -= 1
            admitted.append(r)
            continue
        need = chunk_bytes[n]
        if not growable or need > budget:
            blocked.add(n)
            waiting.append(r)
            continue
        if in_use + need > budget:
            picks, freed = [], 0
            for key in idle:
                if key[0] == n:
                    continue
                picks.append(key)
                freed += chunk_bytes[key[0]]
                if in_use - freed + need <= budget:
                    break
            if in_use - freed + need > budget:
                blocked.add(n)
                waiting.append(r)
                continue
            for key in picks:
                idle.remove(key)
                release.append(key)
                free[key[0]] = free.get(key[0], 0) - chunk_slots[key[0]]
            in_use -= freed
        allocate[n] += 1
        in_use += need
        free[n] = free.get(n, 0) + chunk_slots[n] - 1
        admitted.append(r)
    return Admission(admitted, waiting, dict(allocate), release)


class _Chunk:
    """Resident slots of one frame size (one plan) on its class's stream."""

    def __init__(self, cls, cid, slots):
        self.cls, self.cid, self.slots = cls, cid, slots
        self.n = cls.n
        # per-pass launches: the size classes run on concurrent streams, and a
        # resident launch (cooperative, the whole device) must not overlap another
        self.plan = Plan(cls.ctx, cls.n, batch=slots, shared=False, tile=cls.tile, ept=cls.ept,
                         resident=2)
        self.free = list(range(slots))
        self.active = {}          # slot -> [request, steps, pos, trace chunks, admitted]
        self.event = None
        self.idle_since = 0.0

    @property
    def key(self):
        return (self.n, self.cid)

    def admit(self, req: Request, now: float):
        slot = self.free.pop(0)
        dev = self.cls.ctx.device
        with torch.cuda.stream(self.cls.stream):
            self.plan.set_modulus(_as_dev(req.modulus, dev, torch.float32), slot)
            self.plan.set_support(_as_dev(req.support, dev, torch.float32), slot)
            self.plan.set_psi(_as_dev(req.guess, dev, torch.complex64), slot, mult_support=True)
        self.active[slot] = [req, _steps_for(req), 0, [], now]

    def table(self, block: int):
        """(per-frame step table, steps each slot takes) of the next block."""
        return block_table({f: (st[1], st[2]) for f, st in self.active.items()}, block,
                           self.slots)

    def launched(self, err, takes, event):
        """Account a queued block: its error trace and the steps each slot took."""
        for f, st in self.active.items():
            st[3].append((err, f, takes[f]))
            st[2] += takes[f]
        self.event = event

    def launch(self, block: int):
        """Queue the next `block` iterations of every active slot."""
        if not self.active:
            return False
        table, takes = self.table(block)
        with torch.cuda.stream(self.cls.stream):
            err = self.plan.iterate(table, per_frame=True)
        event = torch.cuda.Event()
        event.record(self.cls.stream)
        self.launched(err, takes, event)
        return True

    def retire(self, now: float):
        """Results of slots whose schedule is complete (after the event)."""
        done = []
        for f in [f for f, st in self.active.items() if st[2] >= len(st[1])]:
            req, _, _, chunks, admitted = self.active.pop(f)
            with torch.cuda.stream(self.cls.stream):
                psi = self.plan.get_psi(f)
                sup = self.plan.get_support(f)
            trace = torch.cat([e[:t, fr] for e, fr, t in chunks]).cpu().numpy() \
                if chunks else np.zeros(0)
            done.append(Result(req.rid, req.n, psi, sup, trace, req.arrival, admitted, now))
            self._free_slot(f, now)
        return done

    def _free_slot(self, f, now):
        self.free.append(f)
        self.free.sort()
        if not self.active:
            self.idle_since = now

    def close(self):
        if self.event is not None:
            self.event.synchronize()
        self.plan.close()


class _Class:
    """One frame size: its stream, context, kernel shape and chunks.  A class
    of the merged set shares the merged stream and context and takes the
    mixed-size launch shape."""

    def __init__(self, device_index, n, chunk_slots, autotune=False, merged=None):
        self.n, self.chunk_slots = n, chunk_slots
        self.merged = merged is not None
        self.tile = self.ept = 0
        if self.merged:
            self.stream, self.ctx = merged
            self.tile, self.ept = group_shape(n)
            self.chunks = []
            self.next_id = 0
            return
        self.stream = torch.cuda.Stream(device=device_index)
        self.ctx = Context(device_index, stream=self.stream)
        if autotune:
            from .tune import autotune as tune_fn
            with torch.cuda.stream(self.stream):
                r = tune_fn(self.ctx, n, chunk_slots, shared=False)
            self.tile, self.ept = r.tile, r.ept
        self.chunks = []
        self.next_id = 0

    def new_chunk(self):
        c = _Chunk(self, self.next_id, self.chunk_slots)
        self.next_id += 1
        self.chunks.append(c)
        return c


def _as_dev(x, dev, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=dev, dtype=dtype)


class Server:
    """HBM-budgeted continuous batching over size classes.

    Two forms:
      - budget: ``budget_bytes`` and ``sizes`` (or ``mix``, whose keys are the
        sizes): slot chunks of each size are allocated on demand from the one
        budget and idle chunks are released for other sizes (plan_admission);
        ``chunk_slots`` overrides the slots per chunk (chunk_shapes);
      - static: ``slots`` {n: count}: one chunk per size, allocated up front,
        never grown or released.
    block: iterations per launch round (also the admission granularity); the
    default 21 is prime to the shrinkwrap interval of 20, so that a request's
    shrinkwraps mostly fall inside a launch, where they run fused into the
    passes (a shrinkwrap on the last step of a launch runs as stand-alone
    kernels).  autotune: pick each size's FFT layout and radix schedule by
    measurement (tune.autotune).  merge_sizes: the classes of 256^2, 512^2
    and 1024^2 share one stream and run each block of all their chunks as
    merged launches (spk_group_iterate: one launch per pass for every size);
    other sizes keep their own streams.  Results are the same either way.
    """

    def __init__(self, device=0, slots=None, budget_bytes=None, mix=None, sizes=None,
                 block=21, autotune=False, chunk_slots=None, merge_sizes=False):
        require_gpu()
        self.device_index = int(device)
        self.block = int(block)
        if slots is not None:
            slots = {int(n): int(k) for n, k in slots.items() if k > 0}
            self.chunk_slots = dict(slots)
            self.chunk_bytes = {n: plan_bytes(n, k, False, CHUNK_OPTIONS) for n, k in slots.items()}
            self.budget = sum(self.chunk_bytes.values())
            self.growable = False
        else:
            assert budget_bytes, "give slots, or budget_bytes and sizes (or mix)"
            sz = sorted(int(n) for n in (sizes if sizes is not None else mix))
            self.budget = int(budget_bytes)
            cs, cb = chunk_shapes(sz, self.budget)
            if chunk_slots:
                cs.update({int(n): int(k) for n, k in chunk_slots.items()})
                cb = {n: plan_bytes(n, cs[n], False, CHUNK_OPTIONS) for n in cs}
            self.chunk_slots, self.chunk_bytes = cs, cb
            self.growable = True
        # merge_sizes: the classes of the mixed-size launch sizes share one
        # stream, and every block of their chunks runs as one
        # spk_group_iterate call (one launch per pass for all of them)
        self.group = None
        merged = None
        if merge_sizes and sum(group_tile_ok(n) for n in self.chunk_slots) >= 2:
            st = torch.cuda.Stream(device=self.device_index)
            self._mctx = Context(self.device_index, stream=st)
            self.group = Group(self._mctx)
            merged = (st, self._mctx)
        self.classes = {n: _Class(self.device_index, n, k, autotune,
                                  merged if (merged and group_tile_ok(n)) else None)
                        for n, k in sorted(self.chunk_slots.items())}
        if not self.growable:
            for c in self.classes.values():
                c.new_chunk()
        self.queue = collections.deque()
        self.results = []
        self.peak_bytes = self.bytes_reserved()
        self.chunks_allocated = self.chunks_released = 0
        self.t0 = time.perf_counter()

    # -- accounting ------------------------------------------------------
    @staticmethod
    def size_slots(budget_bytes, mix, max_slots=256):
        """Static slots per size within the HBM budget, proportional to
        expected work (fraction x n^2), for the static form."""
        work = {n: f * n * n for n, f in mix.items()}
        tot = sum(work.values())
        out = {}
        for n, w in work.items():
            share = budget_bytes * w / tot
            per = plan_bytes(n, 1, False, CHUNK_OPTIONS)
            out[n] = int(max(1, min(max_slots, share // per)))
        return out

    def chunks(self):
        return [ch for c in self.classes.values() for ch in c.chunks]

    def bytes_reserved(self):
        return sum(self.chunk_bytes[ch.n] for ch in self.chunks())

    def now(self):
        return time.perf_counter() - self.t0

    # -- requests ----------------------------------------------------------
    def submit(self, req: Request):
        assert req.n in self.classes, f"no slots for frame size {req.n}"
        assert self.chunk_bytes[req.n] <= self.budget, "one chunk exceeds the HBM budget"
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
        for ch in self.chunks():
            for f, st in list(ch.active.items()):
                if st[0].rid == rid:
                    if ch.event is not None:   # its last launch may still run
                        ch.event.synchronize()
                    del ch.active[f]
                    ch._free_slot(f, self.now())
                    return True
        return False

    def peek(self, rid: int):
        """(iterations done, device copy of the current iterate) of a running
        request, after its queued launches; None if it is not running."""
        for ch in self.chunks():
            for f, st in ch.active.items():
                if st[0].rid == rid:
                    if ch.event is not None:
                        ch.event.synchronize()
                    with torch.cuda.stream(ch.cls.stream):
                        psi = ch.plan.get_psi(f)
                    ch.cls.stream.synchronize()
                    return st[2], psi
        return None

    def _admit(self, now):
        chunks = self.chunks()
        free = collections.Counter()
        for ch in chunks:
            free[ch.n] += len(ch.free)
        idle = [ch.key for ch in sorted((c for c in chunks if not c.active and c.event is None),
                                        key=lambda c: c.idle_since)]
        adm = plan_admission(self.queue, now, free, idle, self.chunk_slots, self.chunk_bytes,
                             self.budget, self.bytes_reserved(), self.growable)
        self.queue = adm.waiting
        for key in adm.release:
            cls = self.classes[key[0]]
            ch = next(c for c in cls.chunks if c.key == key)
            cls.chunks.remove(ch)
            ch.close()
            self.chunks_released += 1
        for n, cnt in adm.allocate.items():
            for _ in range(cnt):
                self.classes[n].new_chunk()
                self.chunks_allocated += 1
        self.peak_bytes = max(self.peak_bytes, self.bytes_reserved())
        for r in adm.admitted:
            ch = next(c for c in self.classes[r.n].chunks if c.free)
            ch.admit(r, now)

    # -- the loop ------------------------------------------------------------
    def _merged_chunks(self):
        return [ch for c in self.classes.values() if c.merged for ch in c.chunks]

    def _own_chunks(self):
        return [ch for c in self.classes.values() if not c.merged for ch in c.chunks]

    def _launch_merged(self):
        """One block of every busy chunk of the merged classes, as
        spk_group_iterate calls of at most 16 plans; the chunks share the
        completion event."""
        chunks = [ch for ch in self._merged_chunks() if ch.active]
        if not chunks:
            return []
        tabs = [ch.table(self.block) for ch in chunks]
        errs = []
        with torch.cuda.stream(self._mctx.stream):
            for i in range(0, len(chunks), 16):
                part = chunks[i:i + 16]
                errs += self.group.iterate([ch.plan for ch in part],
                                           [t[0] for t in tabs[i:i + 16]],
                                           per_frame=[True] * len(part))
        event = torch.cuda.Event()
        event.record(self._mctx.stream)
        for ch, (_, takes), err in zip(chunks, tabs, errs):
            ch.launched(err, takes, event)
        return chunks

    def tick(self):
        """Admit, launch one block on every busy chunk, wait, retire (lockstep)."""
        now = self.now()
        self._admit(now)
        launched = [ch for ch in self._own_chunks() if ch.launch(self.block)]
        launched += self._launch_merged()
        for ch in launched:
            ch.event.synchronize()
        now = self.now()
        out = []
        for ch in launched:
            out.extend(ch.retire(now))
            ch.event = None
        self.results.extend(out)
        return out

    def busy(self):
        return bool(self.queue) or any(ch.active for ch in self.chunks())

    def poll(self):
        """One pass of the asynchronous loop: each chunk whose block finished is
        retired and relaunched on its own, so a size class never waits for
        another.  Returns True if anything progressed."""
        progressed = False
        self._admit(self.now())
        for ch in self._own_chunks():
            if ch.event is not None:
                if not ch.event.query():
                    continue
                now = self.now()
                self.results.extend(ch.retire(now))
                ch.event = None
                self._admit(now)
                progressed = True
            if ch.active and ch.launch(self.block):
                progressed = True
        if self.group is not None:
            merged = self._merged_chunks()
            pending = [ch for ch in merged if ch.event is not None]
            if pending and not pending[0].event.query():
                return progressed
            if pending:
                now = self.now()
                for ch in pending:
                    self.results.extend(ch.retire(now))
                    ch.event = None
                self._admit(now)
                progressed = True
            if self._launch_merged():
                progressed = True
        return progressed

    def run(self, requests, clock_start=None, lockstep=False):
        """Serve requests (arrival times relative to the start); returns results."""
        self.t0 = time.perf_counter() if clock_start is None else clock_start
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        while self.busy() or any(ch.event is not None for ch in self.chunks()):
            if lockstep:
                if not any(ch.active for ch in self.chunks()) and self.queue:
                    wait = min(r.arrival for r in self.queue) - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                self.tick()
                continue
            if not self.poll():
                idle = all(ch.event is None and not ch.active for ch in self.chunks())
                if idle and self.queue:
                    wait = min(r.arrival for r in self.queue) - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 0.005))
                else:
                    time.sleep(20e-6)
        return self.results

    def close(self):
        for ch in self.chunks():
            ch.close()
        for c in self.classes.values():
            c.chunks = []
        if self.group is not None:
            self.group.close()
            self._mctx.close()
            self.group = None
