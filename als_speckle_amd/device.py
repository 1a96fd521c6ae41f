"""Device plumbing: a context (device + stream) and plans over torch tensors.

PyTorch provides device memory and the stream; every computation is a call
into ``libspeckle_hip.so``.  Nothing here computes on the host.
"""
from __future__ import annotations

import atexit
import ctypes
import sys
import weakref
from ctypes import byref, c_float, c_size_t, c_void_p

import numpy as np
import torch

from . import _capi
from ._capi import check


_live_plans: "weakref.WeakSet[Plan]" = weakref.WeakSet()
_live_groups: "weakref.WeakSet[Group]" = weakref.WeakSet()
_live_ctxs: "weakref.WeakSet[Context]" = weakref.WeakSet()


@atexit.register
def _close_all() -> None:
    # release device objects while the HIP runtime is still up
    for g in list(_live_groups):
        g.close()
    for p in list(_live_plans):
        p.close()
    for c in list(_live_ctxs):
        c.close()


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "als_speckle_amd runs on a ROCm GPU only and there is none visible "
            "(torch.cuda.is_available() is False); it has no CPU path")


def _ptr(t: torch.Tensor) -> c_void_p:
    return c_void_p(t.data_ptr())


def _expect(t: torch.Tensor, dtype, shape, device, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.device != device:
        raise ValueError(f"{name} must be on {device}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class Context:
    """One device and one HIP stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, stream: torch.cuda.Stream | None = None):
        require_gpu()
        self.lib = _capi.load()
        self.index = int(device)
        self.device = torch.device("cuda", self.index)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = c_void_p()
        check(self.lib.spk_ctx_create(self.index, c_void_p(self.stream.cuda_stream), byref(h)))
        self.handle = h
        _live_ctxs.add(self)

    def sync(self) -> None:
        check(self.lib.spk_sync(self.handle))

    def align_global_phase(self, data: torch.Tensor) -> torch.Tensor:
        """In place on a (frames, rows, cols) complex64 device tensor."""
        if data.dim() == 2:
            data = data.unsqueeze(0)
        f, r, c = data.shape
        _expect(data, torch.complex64, (f, r, c), self.device, "data")
        check(self.lib.spk_align_global_phase(self.handle, _ptr(data), f, r, c))
        return data

    def stack_mean(self, data: torch.Tensor) -> torch.Tensor:
        """Mean over the first axis of a (frames, rows, cols) complex64 tensor."""
        f, r, c = data.shape
        _expect(data, torch.complex64, (f, r, c), self.device, "data")
        out = torch.empty((r, c), dtype=torch.complex64, device=self.device)
        check(self.lib.spk_stack_mean(self.handle, _ptr(data), f, r, c, _ptr(out)))
        return out

    def blur2d(self, img: torch.Tensor, sigma: float) -> torch.Tensor:
        """refine_support's periodic Gaussian blur of a (rows, cols) float32 image."""
        r, c = img.shape
        _expect(img, torch.float32, (r, c), self.device, "image")
        out = torch.empty_like(img)
        check(self.lib.spk_blur2d(self.handle, _ptr(img), _ptr(out), r, c, float(sigma)))
        return out

    def com_average(self, stack: torch.Tensor) -> torch.Tensor:
        """Sum of the frames rolled onto frame 0's centre of mass (complex64)."""
        f, r, c = stack.shape
        _expect(stack, torch.complex64, (f, r, c), self.device, "stack")
        out = torch.empty((r, c), dtype=torch.complex64, device=self.device)
        check(self.lib.spk_com_average(self.handle, _ptr(stack), f, r, c, _ptr(out)))
        return out

    def copy_bytes(self, dst: torch.Tensor, src: torch.Tensor) -> None:
        nb = src.numel() * src.element_size()
        check(self.lib.spk_copy_bytes(self.handle, _ptr(dst), _ptr(src), nb))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.spk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def plan_bytes(n: int, batch: int, shared: bool, options: int = 0) -> int:
    """Device bytes of a plan once the buffers of ``options`` (``_capi.PB_*``
    bits: DM, SHRINKWRAP, IPSF, SPECTRUM, TRACE) are allocated."""
    out = c_size_t()
    check(_capi.load().spk_plan_bytes(n, batch, int(shared), int(options), byref(out)))
    return out.value


class Plan:
    """``batch`` resident frames of n x n (see spk_plan_create)."""

    def __init__(self, ctx: Context, n: int, batch: int = 1, shared: bool = True,
                 tile: int = 0, ept: int = 0, persist: bool = False,
                 sw_separate: bool = False, resident: int = 0, exchange: int = 0,
                 barrier_polls: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        self.n, self.batch, self.shared = int(n), int(batch), bool(shared)
        self.persist = bool(persist)
        tune = _capi.SpkTune()
        tune.tile = int(tile)
        tune.ept = int(ept)
        tune.persist = int(self.persist)
        tune.sw_separate = int(bool(sw_separate))
        tune.resident = int(resident)      # 1 on, 2 off, 0 library default
        tune.exchange = int(exchange)
        tune.barrier_polls = int(barrier_polls)   # 0: library default
        h = c_void_p()
        check(self.lib.spk_plan_create(ctx.handle, self.n, self.batch, int(self.shared),
                                       byref(tune), byref(h)))
        self.handle = h
        _live_plans.add(self)
        t, e, pe = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.spk_plan_info(h, byref(t), byref(e), byref(pe)))
        self.tile, self.ept = t.value, e.value       # the resolved shape

    # -- helpers ---------------------------------------------------------
    def _frames_shape(self, frame, per_frame):
        n = self.n
        if frame == -1 and per_frame and self.batch > 1:
            return (self.batch, n, n)
        return (n, n)

    # -- data in / out ---------------------------------------------------
    def set_modulus(self, m: torch.Tensor, frame: int = -1) -> None:
        _expect(m, torch.float32, self._frames_shape(frame, not self.shared), self.ctx.device, "modulus")
        check(self.lib.spk_set_modulus(self.handle, frame, _ptr(m)))

    def set_support(self, s: torch.Tensor, frame: int = -1) -> None:
        _expect(s, torch.float32, self._frames_shape(frame, not self.shared), self.ctx.device, "support")
        check(self.lib.spk_set_support(self.handle, frame, _ptr(s)))

    def get_support(self, frame: int = -1) -> torch.Tensor:
        out = torch.empty(self._frames_shape(frame, not self.shared), dtype=torch.float32,
                          device=self.ctx.device)
        check(self.lib.spk_get_support(self.handle, frame, _ptr(out)))
        return out

    def set_psi(self, guess: torch.Tensor, frame: int = -1, mult_support: bool = True) -> None:
        _expect(guess, torch.complex64, self._frames_shape(frame, True), self.ctx.device, "guess")
        check(self.lib.spk_set_psi(self.handle, frame, _ptr(guess), int(mult_support)))

    def get_psi(self, frame: int = -1) -> torch.Tensor:
        out = torch.empty(self._frames_shape(frame, True), dtype=torch.complex64,
                          device=self.ctx.device)
        check(self.lib.spk_get_psi(self.handle, frame, _ptr(out)))
        return out

    # -- the loop --------------------------------------------------------
    def iterate(self, steps, per_frame: bool = False, error_trace: bool = True,
                err_out: torch.Tensor | None = None):
        """steps: list of (alg, sem, beta, sw_sigma, sw_tau) codes, or a
        ``_capi.STEP_DTYPE`` array (step-major, frames fastest with per_frame).

        Returns a (nsteps, batch) float64 device tensor of modulus errors, or
        None.  With per_frame, ``steps`` has nsteps * batch entries.
        """
        tab = _capi.step_table(steps)
        nsteps = tab.size // (self.batch if per_frame else 1)
        err = None
        if error_trace:
            err = err_out if err_out is not None else torch.empty(
                (nsteps, self.batch), dtype=torch.float64, device=self.ctx.device)
        # the library copies the table into its staged upload before returning
        check(self.lib.spk_iterate(self.handle,
                                   tab.ctypes.data_as(ctypes.POINTER(_capi.SpkStep)), nsteps,
                                   int(per_frame), _ptr(err) if err is not None else None))
        return err

    def sw_counts(self) -> tuple[int, int]:
        """(fused, separate): shrinkwraps run so far inside the passes and as
        stand-alone kernels."""
        fu, se = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.spk_plan_sw_counts(self.handle, byref(fu), byref(se)))
        return fu.value, se.value

    def resident_counts(self) -> tuple[int, int]:
        """(launches, steps) of the resident single-frame launches so far."""
        la, st = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.spk_plan_resident_counts(self.handle, byref(la), byref(st)))
        return la.value, st.value

    def shrinkwrap(self, sigma: float, tau: float, frame: int = -1) -> None:
        check(self.lib.spk_shrinkwrap(self.handle, frame, float(sigma), float(tau)))

    def copy_roi(self, frame: int, dst: torch.Tensor, slot: int, r0: int, c0: int,
                 rows: int, cols: int) -> None:
        if dst.dim() != 3 or tuple(dst.shape[1:]) != (rows, cols):
            raise ValueError("dst must be (slots, rows, cols)")
        _expect(dst, torch.complex64, tuple(dst.shape), self.ctx.device, "savebuffer")
        check(self.lib.spk_copy_roi(self.handle, frame, _ptr(dst), dst.shape[0], slot,
                                    r0, c0, rows, cols))

    def fft2(self, x: torch.Tensor, inverse: bool = False, out: torch.Tensor | None = None):
        squeeze = x.dim() == 2
        xx = x.unsqueeze(0) if squeeze else x
        cnt = xx.shape[0]
        _expect(xx, torch.complex64, (cnt, self.n, self.n), self.ctx.device, "input")
        o = out if out is not None else torch.empty_like(xx)
        check(self.lib.spk_fft2(self.handle, _ptr(xx), _ptr(o), cnt, int(inverse)))
        return o[0] if squeeze else o

    def capture(self, fn) -> "Graph":
        """Record the plan calls made by fn() into a HIP graph (not run)."""
        check(self.lib.spk_graph_begin(self.handle))
        try:
            fn()
        finally:
            h = c_void_p()
            rc = self.lib.spk_graph_end(self.handle, byref(h))
        check(rc)
        return Graph(self.lib, h, self)

    def prtf(self, stack: torch.Tensor) -> torch.Tensor:
        """|mean_f F_f / |F_f|| of a (frames, rows, cols) complex64 stack, padded to n x n."""
        f, r, c = stack.shape
        _expect(stack, torch.complex64, (f, r, c), self.ctx.device, "estimates")
        out = torch.empty((self.n, self.n), dtype=torch.float32, device=self.ctx.device)
        check(self.lib.spk_prtf(self.handle, _ptr(stack), f, r, c, _ptr(out)))
        return out

    def rftf(self, estimate: torch.Tensor, goal: torch.Tensor, scale: bool = False,
             ipsf: torch.Tensor | None = None):
        r, c = estimate.shape
        _expect(estimate, torch.complex64, (r, c), self.ctx.device, "estimate")
        _expect(goal, torch.float32, (self.n, self.n), self.ctx.device, "goal_modulus")
        if ipsf is not None:
            _expect(ipsf, torch.complex64, (self.n, self.n), self.ctx.device, "ipsf")
        out = torch.empty((self.n, self.n), dtype=torch.float32, device=self.ctx.device)
        check(self.lib.spk_rftf(self.handle, _ptr(estimate), r, c, _ptr(goal),
                                _ptr(ipsf) if ipsf is not None else None, int(scale), _ptr(out)))
        return out

    def rftf_stack(self, stack: torch.Tensor, goal: torch.Tensor,
                   ipsf: torch.Tensor | None = None) -> torch.Tensor:
        f, r, c = stack.shape
        _expect(stack, torch.complex64, (f, r, c), self.ctx.device, "estimates")
        _expect(goal, torch.float32, (self.n, self.n), self.ctx.device, "goal_modulus")
        if ipsf is not None:
            _expect(ipsf, torch.complex64, (self.n, self.n), self.ctx.device, "ipsf")
        out = torch.empty((self.n, self.n), dtype=torch.float32, device=self.ctx.device)
        check(self.lib.spk_rftf_stack(self.handle, _ptr(stack), f, r, c, _ptr(goal),
                                      _ptr(ipsf) if ipsf is not None else None, _ptr(out)))
        return out

    def coherence_rl(self, estimate: torch.Tensor, modulus: torch.Tensor, iterations: int,
                     semantics: str = "numpy", is_fourier: bool = False) -> torch.Tensor:
        """Richardson-Lucy ipsf (n x n complex64, not centre-shifted) of one frame."""
        r, c = estimate.shape
        _expect(estimate, torch.complex64, (r, c), self.ctx.device, "estimate")
        _expect(modulus, torch.float32, (self.n, self.n), self.ctx.device, "modulus")
        out = torch.empty((self.n, self.n), dtype=torch.complex64, device=self.ctx.device)
        check(self.lib.spk_coherence_rl(self.handle, _ptr(estimate), r, c, _ptr(modulus),
                                        int(iterations), _capi.SEM_CODES[semantics],
                                        int(bool(is_fourier)), _ptr(out)))
        return out

    def gauss_prepare(self, estimate: torch.Tensor, modulus: torch.Tensor) -> None:
        r, c = estimate.shape
        _expect(estimate, torch.complex64, (r, c), self.ctx.device, "estimate")
        _expect(modulus, torch.float32, (self.n, self.n), self.ctx.device, "modulus")
        check(self.lib.spk_coherence_gauss_prepare(self.handle, _ptr(estimate), r, c,
                                                   _ptr(modulus)))

    def gauss_eval(self, cl0: float, cl1: float) -> float:
        v = ctypes.c_double()
        check(self.lib.spk_coherence_gauss_eval(self.handle, float(cl0), float(cl1),
                                                ctypes.byref(v)))
        return v.value

    def profile(self, enable: bool) -> None:
        check(self.lib.spk_plan_profile(self.handle, int(enable)))

    def profile_read(self, which: int):
        """(launches, total_ms) of pass 0 (column), 1 (row) or 2 (resident
        launches) since the last read."""
        n = ctypes.c_int()
        ms = ctypes.c_double()
        check(self.lib.spk_plan_profile_read(self.handle, which, byref(n), byref(ms)))
        return n.value, ms.value

    def time_passes(self, which: int, reps: int, step) -> float:
        arr = _capi.make_steps([step])
        ms = c_float()
        check(self.lib.spk_time_passes(self.handle, which, reps, arr, byref(ms)))
        return ms.value

    def set_ipsf(self, ipsf: torch.Tensor | None) -> None:
        """Partial-coherence ipsf (n x n complex64, corner-centred), or None."""
        if ipsf is None:
            check(self.lib.spk_set_ipsf(self.handle, None))
            return
        _expect(ipsf, torch.complex64, (self.n, self.n), self.ctx.device, "ipsf")
        check(self.lib.spk_set_ipsf(self.handle, _ptr(ipsf)))

    def set_spectrum(self, rescales, weights) -> None:
        """Spectral blur: K rescale factors and K weights (host arrays), or None."""
        if rescales is None:
            check(self.lib.spk_set_spectrum(self.handle, None, None, 0))
            return
        r = np.ascontiguousarray(rescales, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        assert r.ndim == 1 and r.shape == w.shape, "rescales and weights: equal 1-d arrays"
        check(self.lib.spk_set_spectrum(self.handle, r.ctypes.data, w.ctypes.data, r.size))

    def check(self) -> None:
        """Wait for queued work; raise if a grid barrier of a resident or
        persistent launch gave up waiting (SpkError, the flag is cleared)."""
        check(self.lib.spk_plan_check(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.spk_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of plan calls; launch() replays it."""

    def __init__(self, lib, handle, plan):
        self.lib, self.handle, self.plan = lib, handle, plan

    def launch(self) -> None:
        check(self.lib.spk_graph_launch(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.spk_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def group_shape(n: int) -> tuple[int, int]:
    """(tile, ept) a plan of size n needs to join a mixed-size launch
    (sizes 256, 512 and 1024); raises SpkError for other sizes."""
    lib = _capi.load()
    t, e = ctypes.c_int(), ctypes.c_int()
    check(lib.spk_group_shape(int(n), byref(t), byref(e)))
    return t.value, e.value


class Group:
    """Frames of different sizes in one launch per pass (spk_group_iterate):
    ``iterate`` runs one spk_iterate call per plan with the plans'
    single-projection passes merged into one column-pass and one row-pass
    launch per step.  Plans come from ``Group.plan`` (or are any plans in
    the ``group_shape`` tile and values per thread) on this group's
    context."""

    def __init__(self, ctx: Context):
        self.ctx, self.lib = ctx, ctx.lib
        h = c_void_p()
        check(self.lib.spk_group_create(ctx.handle, byref(h)))
        self.handle = h
        _live_groups.add(self)

    def plan(self, n: int, batch: int = 1, shared: bool = True, **kw) -> Plan:
        """A plan of size n in the mixed-size shape, on the group's context."""
        tile, ept = group_shape(n)
        return Plan(self.ctx, n, batch=batch, shared=shared, tile=tile, ept=ept, **kw)

    def iterate(self, plans, steps, per_frame=None, error_trace: bool = True):
        """plans: list of Plan; steps: one step table per plan (as for
        Plan.iterate), all with the same number of steps; per_frame: one flag
        per plan (default all False).  Returns one (nsteps, batch) float64
        error tensor per plan, or Nones."""
        count = len(plans)
        if len(steps) != count:
            raise ValueError("one step table per plan")
        per_frame = [False] * count if per_frame is None else [bool(x) for x in per_frame]
        tabs = [_capi.step_table(s) for s in steps]
        nsteps = [t.size // (p.batch if pf else 1) for t, p, pf in zip(tabs, plans, per_frame)]
        if len(set(nsteps)) > 1:
            raise ValueError(f"every plan takes the same number of steps, got {nsteps}")
        ns = nsteps[0] if nsteps else 0
        errs = [torch.empty((ns, p.batch), dtype=torch.float64, device=self.ctx.device)
                if error_trace else None for p in plans]
        hs = (c_void_p * count)(*[p.handle.value for p in plans])
        ts = (c_void_p * count)(*[t.ctypes.data for t in tabs])
        pf = (ctypes.c_int * count)(*[int(x) for x in per_frame])
        es = (c_void_p * count)(*[(e.data_ptr() if e is not None else None) for e in errs])
        check(self.lib.spk_group_iterate(self.handle, hs, count, ts, ns, pf, es))
        return errs

    def counts(self) -> tuple[int, int]:
        """(merged launches, member passes they covered) so far."""
        la, mp = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.spk_group_counts(self.handle, byref(la), byref(mp)))
        return la.value, mp.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.spk_group_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def to_device(a: np.ndarray, device, dtype) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)
