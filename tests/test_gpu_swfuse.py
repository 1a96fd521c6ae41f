"""Shrinkwrap fused into the passes of its step and the next (row blur in the
row pass, column blur and band maxima in the next column pass, threshold in
the next row pass) against the stand-alone shrinkwrap kernels (k_blur_rows,
k_blur_cols, k_sw_threshold), which the parity tests pin to the oracle's
shrinkwrap (phasing.py:1787-1815 restated in oracle/phasing_c128.py).

The bar is bitwise: the fused form sums the same blur terms in the same order
and thresholds in FP64 as the stand-alone kernels, so iterates, supports and
error traces must be identical.  Each case also checks that the fusion ran
(spk_plan_sw_counts), and the fallbacks (radius > 24, an idle frame or a DM
step after the shrinkwrap, the last step of a call) run stand-alone.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan

pytestmark = pytest.mark.gpu


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def step(alg, beta=0.9, sem="numpy", sigma=0.0, tau=0.0):
    return (_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, sigma, tau)


NOOP = (_capi.NOOP, 0, 0.0, 0.0, 0.0)


def frames(n, count, seed=0):
    out = []
    for k in range(count):
        f = synth.make_frame(n, seed + k, shape="disk", radius=max(3, n * 40 // 256),
                             photons=1e6, noise=True)
        rng = np.random.default_rng(seed + 100 + k)
        s = np.zeros((n, n), np.float32)
        r = n // 4 + 2 * k                      # a loose support for the shrinkwrap to cut
        s[n // 2 - r:n // 2 + r, n // 2 - r:n // 2 + r] = 1
        g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64) * s
        out.append((f.modulus.astype(np.float32), s, g))
    return out


def run(ctx, data, sched, per_frame=False, separate=False, **tune):
    n = data[0][0].shape[0]
    b = len(data)
    p = Plan(ctx, n, batch=b, shared=False, sw_separate=separate, **tune)
    stack = (lambda k: data[0][k]) if b == 1 else (lambda k: np.stack([d[k] for d in data]))
    p.set_modulus(dev(stack(0), torch.float32))
    p.set_support(dev(stack(1), torch.float32))
    p.set_psi(dev(stack(2), torch.complex64), mult_support=False)
    tr = p.iterate(sched, per_frame=per_frame).cpu().numpy()
    out = (p.get_psi().cpu().numpy(), p.get_support().cpu().numpy(), tr, p.sw_counts())
    p.close()
    return out


def same(a, b):
    pa, sa, ta, _ = a
    pb, sb, tb, _ = b
    assert np.array_equal(sa, sb), int((sa != sb).sum())
    assert np.array_equal(pa, pb)
    assert np.array_equal(ta, tb, equal_nan=True)


def schedule(iters, every, sigmas=(3.0, 1.5), tau=0.2, alg="hio"):
    out = []
    for k in range(iters):
        if (k + 1) % every == 0:
            frac = k / max(1, iters - 1)
            sig = sigmas[0] + (sigmas[1] - sigmas[0]) * frac
            out.append(step(alg, 0.9, "numpy", sig, tau))
        else:
            out.append(step(alg, 0.9))
    return out


@pytest.mark.parametrize("n,tile,ept", [(64, 0, 0), (256, 2, 8), (256, 4, 16), (256, 8, 8),
                                        (1024, 0, 0), (1024, 4, 16), (2048, 0, 0),
                                        (4096, 0, 0), (4096, 4, 16)])
def test_fused_equals_separate_single_frame(gpu_ctx, n, tile, ept):
    print(f"[swfuse n={n} tile={tile} ept={ept}]", flush=True)
    data = frames(n, 1, seed=n)
    sched = schedule(41, 10, tau=0.5)
    fused = run(gpu_ctx, data, sched, tile=tile, ept=ept)
    sep = run(gpu_ctx, data, sched, separate=True, tile=tile, ept=ept)
    same(fused, sep)
    assert fused[3] == (4, 0) and sep[3] == (0, 4)
    assert fused[1].sum() < data[0][1].sum()          # the support shrank


def test_fused_back_to_back_and_fallbacks(gpu_ctx):
    n = 128
    data = frames(n, 2, seed=3)
    sw = lambda alg, sig: step(alg, 0.9, "numpy", sig, 0.15)   # noqa: E731
    sched = ([step("raar", 0.87)] * 3 + [sw("raar", 2.0), sw("er", 1.0), sw("hio", 2.5)] +
             [step("hio")] * 2 +
             [sw("hio", 5.0)] +                        # radius 30: stand-alone kernels
             [step("er")] * 2 +
             [sw("hio", 2.0), step("dm", 0.8)] +       # DM next: stand-alone
             [step("dm", 0.8), step("er")] * 2 +
             [sw("er", 1.5)])                          # last step of the call: stand-alone
    fused = run(gpu_ctx, data, sched)
    sep = run(gpu_ctx, data, sched, separate=True)
    same(fused, sep)
    assert fused[3] == (3, 3) and sep[3] == (0, 6)


def test_fused_per_frame_tables(gpu_ctx):
    n = 256
    data = frames(n, 3, seed=9)
    iters = 24
    steps = []
    for k in range(iters):
        row = []
        for f in range(3):
            if f == 1 and k in (7, 8):
                row.append(NOOP)                       # idle after its shrinkwrap at 6
            elif (k + 1) % (6 + f) == 0:
                row.append(step("hio", 0.9, "numpy", 1.0 + f, 0.1 + 0.05 * f))
            else:
                row.append(step("hio" if f != 2 else "raar", 0.9 if f != 2 else 0.85))
        steps += row
    fused = run(gpu_ctx, data, steps, per_frame=True)
    sep = run(gpu_ctx, data, steps, per_frame=True, separate=True)
    same(fused, sep)
    # fused after steps 5, 7, 11, 13, 15, 17, 20; stand-alone after 6 (frame 1
    # idle next) and 23 (the last step)
    assert fused[3] == (7, 2) and sep[3] == (0, 9)


def test_fused_batched_shared_coefficients(gpu_ctx):
    n = 1024
    data = frames(n, 4, seed=21)
    sched = schedule(21, 10)
    fused = run(gpu_ctx, data, sched)
    sep = run(gpu_ctx, data, sched, separate=True)
    same(fused, sep)
    assert fused[3] == (2, 0)


@pytest.mark.parametrize("mode", ["ipsf", "persist"])
def test_no_fusion_with_ipsf_or_persistent_blocks(gpu_ctx, mode):
    """The partially coherent projection and the persistent launch mode run
    every shrinkwrap with the stand-alone kernels."""
    n = 64
    m, s, g = frames(n, 1, seed=31)[0]
    p = Plan(gpu_ctx, n, persist=(mode == "persist"))
    p.set_modulus(dev(m, torch.float32))
    p.set_support(dev(s, torch.float32))
    p.set_psi(dev(g, torch.complex64), mult_support=False)
    if mode == "ipsf":
        x = np.fft.fftfreq(n) * n
        psf = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 8.0)
        p.set_ipsf(dev(np.fft.fft2(psf / psf.sum()), torch.complex64))
    tr = p.iterate(schedule(21, 10)).cpu().numpy()
    assert np.all(np.isfinite(tr))
    assert p.sw_counts() == (0, 2)
    p.close()
