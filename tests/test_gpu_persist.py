"""Persistent iteration blocks (one cooperative launch per run of steps, with
grid barriers between the passes) against the one-launch-per-pass path.

The persistent kernel runs the same per-item code on the same inputs, and
the library is built without implicit FMA contraction (fused multiply-adds
only where the source writes them), so the arithmetic does not depend on
which kernel inlines it: iterates, traces and supports agree bitwise.  DM
steps, their predecessors and shrinkwrap events fall back to ordinary
launches inside the same call.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Context, Plan

pytestmark = pytest.mark.gpu


def step(alg, beta=0.9, sem="numpy", sigma=0.0, tau=0.0):
    if alg == "noop":
        return (_capi.NOOP, 0, 0.0, 0.0, 0.0)
    return (_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, sigma, tau)


def run(ctx, n, batch, shared, persist, steps, per_frame=False, seed=0, tile=0, ept=0):
    f = synth.make_frame(n, seed, shape="polygon", fill=0.12, photons=1e7)
    rng = np.random.default_rng(seed + 1)
    g = (rng.random((batch, n, n)) + 1j * rng.random((batch, n, n))).astype(np.complex64)
    p = Plan(ctx, n, batch=batch, shared=shared, tile=tile, ept=ept, persist=persist)
    dev = ctx.device
    if shared:
        p.set_modulus(torch.from_numpy(f.modulus).to(dev))
        p.set_support(torch.from_numpy(f.support).to(dev))
    else:
        p.set_modulus(torch.from_numpy(np.stack([f.modulus] * batch)).to(dev))
        p.set_support(torch.from_numpy(np.stack([f.support] * batch)).to(dev))
    p.set_psi(torch.from_numpy(g if batch > 1 else g[0]).to(dev))
    err = p.iterate(steps, per_frame=per_frame).cpu().numpy()
    p.check()
    psi = p.get_psi().cpu().numpy()
    sup = p.get_support().cpu().numpy()
    p.close()
    return psi, err, sup


def same(a, b):
    # idle (frame, step) entries of a trace are NaN in both
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("n,batch,tile,ept", [(256, 1, 0, 0), (1024, 1, 0, 0), (1024, 1, 2, 8),
                                              (128, 6, 0, 0), (1024, 8, 0, 0), (2048, 1, 0, 0),
                                              (4096, 1, 0, 0)])
def test_persistent_block_equals_per_pass_launches(gpu_ctx, n, batch, tile, ept):
    print(f"[persist n={n} batch={batch} tile={tile} ept={ept}]", flush=True)
    iters = 12 if n == 4096 else 30
    sched = [step("hio", 0.8)] * (iters - 5) + [step("er")] * 5
    a = run(gpu_ctx, n, batch, True, False, sched, tile=tile, ept=ept)
    b = run(gpu_ctx, n, batch, True, True, sched, tile=tile, ept=ept)
    same(a, b)
    c = run(gpu_ctx, n, batch, True, True, sched, tile=tile, ept=ept)
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_persistent_blocks_with_shrinkwrap_dm_and_idle_frames(gpu_ctx):
    n, B = 128, 4
    print(f"[persist mixed n={n} batch={B}]", flush=True)
    sched = []
    for k in range(40):
        for fr in range(B):
            if fr == 3 and k < 10:
                sched.append(step("noop"))           # admitted late
            elif fr == 1 and 20 <= k < 24:
                sched.append(step("dm", 0.7))        # a DM run inside the block
            elif k % 10 == 9:
                sched.append(step("hio", 0.9, "numpy", 2.0, 0.2))   # shrinkwrap after
            else:
                sched.append(step("raar" if fr == 2 else "hio", 0.85))
    a = run(gpu_ctx, n, B, False, False, sched, per_frame=True, seed=3)
    b = run(gpu_ctx, n, B, False, True, sched, per_frame=True, seed=3)
    same(a, b)


def test_persistent_plan_on_its_own_stream(gpu_ctx):
    """A plan on a non-default stream, followed by ordinary work on it."""
    stream = torch.cuda.Stream()
    ctx = Context(0, stream=stream)
    sched = [step("hio", 0.8)] * 20
    with torch.cuda.stream(stream):
        a = run(ctx, 256, 2, True, False, sched)
        b = run(ctx, 256, 2, True, True, sched)
    same(a, b)
