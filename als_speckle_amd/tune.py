"""Autotuner for the FFT layout and radix schedule of a plan.

A plan's pass kernels are parameterised by
  - the tile edge of the internal frame layout (2, 4 or 8: how many rows or
    columns one workgroup transforms, and the size of the contiguous blocks
    its loads and stores move),
  - the values per thread of one transform (8 or 16), which fixes the radix
    schedule: radix-8 stages for 8 (1024 = 8*8*8*2), radix-16 (4x4) stages
    for 16 (1024 = 16*16*4), with one LDS exchange between stages,
  - the launch mode (one launch per pass, or a persistent launch per block of
    iterations; the persistent candidates are timed only when asked for:
    on MI355X they were 1.7-4x slower than the per-pass launches at every
    measured size and trial count, so the default search leaves them out).
The best choice depends on the frame size, the number of resident frames and
the device, so it is measured: every built candidate runs the same
HIO iterations on a synthetic frame and the fastest per iteration wins.
Results are cached per (device, n, batch, shared) in memory and, if a path
is given, in a JSON file.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import torch

from . import _capi
from .device import Context, Plan

__all__ = ["candidates", "autotune", "TuneResult"]

TILES = (2, 4, 8)
EPTS = (8, 16)

_cache: dict = {}


class TuneResult(dict):
    """dict(tile, ept, persist, us_per_iter, table) with attribute access."""

    __getattr__ = dict.__getitem__


def candidates(n: int, persist: bool = False):
    """Built (tile, ept, persist) shapes for frame size n (see size_supported)."""
    out = []
    for tile in TILES:
        for ept in EPTS:
            if n == 4096 and tile == 8:
                continue
            if ept > n or (n // ept) * tile > 1024:
                continue
            out.append((tile, ept, False))
            if persist:
                out.append((tile, ept, True))
    return out


def _key(ctx, n, batch, shared):
    name = torch.cuda.get_device_name(ctx.device)
    return f"{name}|n={n}|batch={batch}|shared={int(shared)}"


def _time_candidate(ctx, n, batch, shared, tile, ept, persist, iters, m, s, g):
    hio = (_capi.HIO, _capi.SEM_CODES["numpy"], 0.8, 0.0, 0.0)
    plan = Plan(ctx, n, batch=batch, shared=shared, tile=tile, ept=ept, persist=persist)
    try:
        plan.set_modulus(m)
        plan.set_support(s)
        plan.set_psi(g)
        plan.iterate([hio] * 4, error_trace=False)          # warm-up
        ctx.sync()
        best = float("inf")
        for _ in range(3):
            plan.set_psi(g)
            ctx.sync()
            t0 = time.perf_counter()
            plan.iterate([hio] * iters, error_trace=False)
            ctx.sync()
            best = min(best, (time.perf_counter() - t0) / iters)
        plan.check()
        return best
    finally:
        plan.close()


def autotune(ctx: Context, n: int, batch: int = 1, shared: bool = True, iters: int = 20,
             persist: bool = False, cache_path: str | None = None) -> TuneResult:
    """Fastest (tile, ept, persist) for `batch` resident n x n frames; the
    persistent launch mode is a candidate only with persist=True.

    Times `iters` HIO iterations per candidate (best of 3) on a synthetic
    frame; allocates one plan at a time.  The result is cached.
    """
    key = _key(ctx, n, batch, shared) + ("|persist" if persist else "")
    if key in _cache:
        return _cache[key]
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            disk = json.load(fh)
        if key in disk:
            _cache[key] = TuneResult(disk[key])
            return _cache[key]
    rng = np.random.default_rng(n)
    frames = batch if not shared else 1
    shape = (frames, n, n) if frames > 1 else (n, n)
    dev = ctx.device
    m = torch.from_numpy((rng.random(shape) + 0.1).astype(np.float32)).to(dev)
    s = torch.from_numpy((rng.random(shape) < 0.2).astype(np.float32)).to(dev)
    gshape = (batch, n, n) if batch > 1 else (n, n)
    g = torch.from_numpy((rng.random(gshape) + 1j * rng.random(gshape)).astype(np.complex64)).to(dev)
    table = []
    for tile, ept, pers in candidates(n, persist):
        us = _time_candidate(ctx, n, batch, shared, tile, ept, pers, iters, m, s, g) * 1e6
        table.append({"tile": tile, "ept": ept, "persist": pers, "us_per_iter": round(us, 3)})
    best = min(table, key=lambda r: r["us_per_iter"])
    res = TuneResult(tile=best["tile"], ept=best["ept"], persist=best["persist"],
                     us_per_iter=best["us_per_iter"], table=table)
    _cache[key] = res
    if cache_path:
        disk = {}
        if os.path.exists(cache_path):
            with open(cache_path) as fh:
                disk = json.load(fh)
        disk[key] = dict(res)
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as fh:
            json.dump(disk, fh, indent=1)
    return res
