"""Batched independent trials of one diffraction pattern.

The reference runs trials one after another from Python
(``examples/basic_phasing_example.py:27-30``: ``for trial: r.seed();
r.iterate(n)``), then aligns and averages them on the host
(``phasing.align_global_phase`` ``phasing.py:1355-1404``, ``np.mean``
``phasing.py:960-961``).  ``phasing_batch`` keeps every trial resident in one
plan with a shared modulus and support, advances all of them in the same
launches, and aligns and averages them on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _capi
from .device import Context, Plan, require_gpu, to_device
from .masking import bounding_box
from .phasing import expand_order

__all__ = ["phasing_batch"]


class phasing_batch:
    """``trials`` random-start reconstructions of one modulus/support."""

    def __init__(self, trials, gpu_info=None, tile=0, ept=0, autotune=False):
        """tile/ept: fixed FFT layout and values per thread (0: library
        default); autotune: measure every built (tile, ept, launch mode) for
        this frame size and trial count at ``load`` and use the fastest."""
        require_gpu()
        assert int(trials) >= 1, "trials must be >= 1"
        self.trials = int(trials)
        self.ctx = gpu_info if isinstance(gpu_info, Context) else Context(
            int(gpu_info) if gpu_info is not None else 0)
        self._tune = (tile, ept, False)
        self._autotune = bool(autotune)
        self.tuning = None
        self.N = self.shape = None
        self._plan = None
        self._trace = None
        self._seeded = False

    def load(self, modulus, support):
        """Corner-centred N x N modulus and an N x N support (normalised to max 1)."""
        assert isinstance(modulus, np.ndarray) and modulus.ndim == 2
        assert modulus.shape[0] == modulus.shape[1], "modulus must be square"
        n = modulus.shape[0]
        assert n & (n - 1) == 0, "modulus size must be a power of 2"
        assert isinstance(support, np.ndarray) and support.shape == modulus.shape
        sup = support.astype(np.float32)
        sup *= 1.0 / sup.max()
        if self._plan is None or self.N != n:
            self.N, self.shape = n, modulus.shape
            if self._autotune:
                from .tune import autotune
                self.tuning = autotune(self.ctx, n, self.trials, shared=True)
                self._tune = (self.tuning.tile, self.tuning.ept, self.tuning.persist)
            self._plan = Plan(self.ctx, n, batch=self.trials, shared=True,
                              tile=self._tune[0], ept=self._tune[1], persist=self._tune[2])
        dev = self.ctx.device
        self._m = to_device(modulus.astype(np.float32), dev, torch.float32)
        self._s = to_device(sup, dev, torch.float32)
        self._plan.set_modulus(self._m)
        self._plan.set_support(self._s)
        b = bounding_box(sup, pad=10)
        self.r0, self.rows = int(b[0]), int(b[1] - b[0])
        self.c0, self.cols = int(b[2]), int(b[3] - b[2])

    def seed(self, vals=None, guesses=None):
        """Random starts as ``phasing.seed`` draws them, one per trial.

        vals: None (consecutive draws from numpy's global legacy generator) or
        a sequence of ``trials`` seeds (``np.random.seed(v)`` before each
        draw).  guesses: alternatively a (trials, N, N) complex array or a
        device tensor, used as given (times the support).
        """
        n, T = self.N, self.trials
        assert n is not None, "load data first"
        if guesses is None:
            g = np.empty((T, n, n), np.complex64)
            for t in range(T):
                if vals is not None:
                    np.random.seed(vals[t])
                g[t] = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
            guesses = to_device(g, self.ctx.device, torch.complex64)
        elif isinstance(guesses, np.ndarray):
            guesses = to_device(guesses.astype(np.complex64), self.ctx.device, torch.complex64)
        self._plan.set_psi(guesses, -1, mult_support=True)
        self._seeded = True

    def iterate(self, iterations, order=None, beta=0.8, semantics="numpy", error_trace=True):
        """All trials advance ``iterations`` iterations together."""
        assert self._seeded, "seed() before iterate()"
        sem = _capi.SEM_CODES[semantics]
        steps = [(_capi.ALG_CODES[a], sem, beta, 0.0, 0.0)
                 for a in expand_order(order, iterations)]
        self._trace = self._plan.iterate(steps, error_trace=error_trace) if steps else None
        if self._plan.persist and steps:
            self._plan.check()   # raises if a persistent block's grid barrier gave up

    def get_error_trace(self):
        """(iterations, trials) FP64 modulus errors of the last ``iterate``."""
        return None if self._trace is None else self._trace.cpu().numpy()

    def roi_stack(self) -> torch.Tensor:
        """(trials, rows, cols) device tensor of every trial's ROI (the savebuffer)."""
        buf = torch.empty((self.trials, self.rows, self.cols), dtype=torch.complex64,
                          device=self.ctx.device)
        for t in range(self.trials):
            self._plan.copy_roi(t, buf, t, self.r0, self.c0, self.rows, self.cols)
        return buf

    def aligned_average(self):
        """Globally phase-aligned trial ROIs and their mean, on the GPU.

        Returns (aligned stack, average) as device tensors.
        """
        stack = self.roi_stack()
        self.ctx.align_global_phase(stack)
        return stack, self.ctx.stack_mean(stack)

    def get(self, x):
        return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x
