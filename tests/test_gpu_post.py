"""Module functions of speckle.phasing after the iterations, on the GPU:
refine_support (phasing.py:1747-1815), prtf (phasing.py:1406-1461), rftf
(phasing.py:1521-1609) and the stack rftf of gpu_rftf (phasing.py:1611-1708),
each against the oracle on the same inputs.

Tolerances: the device transforms and blur keep FP32 values (FP64 sums), the
oracle is complex128, so values agree to ~1e-5 relative; thresholded outputs
(refine_support) agree exactly except at pixels whose blurred value lies
within 1e-5 of the threshold.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import phasing as P
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu


def crand(shape, rng):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def two_part_object(n, rng):
    s = np.zeros((n, n))
    s[6:20, 8:22] = 1
    s[30:44, 34:50] = 1
    obj = s * (1.0 + 0.3 * rng.random((n, n))) * np.exp(1j * rng.uniform(-1, 1, (n, n)))
    obj[30:44, 34:50] *= 0.4              # weaker second part
    return obj, s


@pytest.mark.parametrize("shape,sigma", [((64, 64), 3.0), ((64, 64), 1.5), ((37, 50), 2.0),
                                         ((128, 96), 20.0), ((33, 33), 0.7)])
def test_blur_matches_fft_blur(gpu_ctx, shape, sigma):
    rng = np.random.default_rng(1)
    img = np.abs(crand(shape, rng)).astype(np.float32)
    t = torch.from_numpy(img).to(gpu_ctx.device)
    got = gpu_ctx.blur2d(t, sigma).cpu().numpy()
    want = O.blur_fft(img.astype(np.float64), sigma)
    assert np.max(np.abs(got - want)) < 2e-6 * np.abs(want).max()


def _agree_away_from_threshold(got, want, blurred, thresholds):
    bad = got != want
    if not bad.any():
        return
    near = np.zeros_like(bad)
    for t in thresholds:
        near |= np.abs(blurred - t) < 1e-5 * np.abs(blurred).max()
    assert not (bad & ~near).any()


@pytest.mark.parametrize("kw", [dict(blur=3, local_threshold=0.2, global_threshold=0),
                                dict(blur=2, local_threshold=0.3, global_threshold=0.1),
                                dict(blur=1.5, local_threshold=0.2, global_threshold=0.05,
                                     kill_weakest=True)])
def test_refine_support_matches_oracle(kw):
    rng = np.random.default_rng(5)
    obj, s = two_part_object(64, rng)
    avg = obj + 0.02 * crand(obj.shape, rng)
    s_before = s.copy()
    got, gb = P.refine_support(s, avg, **kw)
    want, wb = O.refine_support(s.copy(), avg, **kw)
    assert np.array_equal(s, s_before)            # the caller's support is kept
    assert got.dtype == s.dtype and got.shape == s.shape
    assert np.max(np.abs(gb - wb)) < 2e-6 * np.abs(wb).max()
    thr = [wb.max() * kw["global_threshold"]]
    parts = O.find_all_objects(s.copy())
    thr += [(wb * p).max() * kw["local_threshold"] for p in parts]
    _agree_away_from_threshold(got, want, wb, thr)
    assert np.all(got <= s)                       # shrink only
    if kw.get("kill_weakest"):
        assert got[30:44, 34:50].sum() == 0       # the weak part is removed


def test_refine_support_odd_crop_and_int_support():
    rng = np.random.default_rng(6)
    s = np.zeros((45, 51), int)
    s[10:30, 12:40] = 1
    avg = s * (1 + rng.random(s.shape))
    got, gb = P.refine_support(s, avg, blur=3)
    want, wb = O.refine_support(s.copy(), avg, blur=3)
    assert got.dtype == s.dtype
    assert np.max(np.abs(gb - wb)) < 2e-6 * np.abs(wb).max()
    parts = O.find_all_objects(s.copy())
    thr = [0.0] + [(wb * p).max() * 0.2 for p in parts]
    _agree_away_from_threshold(got, want.astype(int), wb, thr)


@pytest.mark.parametrize("frames,r,c,N", [(5, 32, 32, 32), (7, 20, 24, 64), (3, 100, 128, 128)])
def test_prtf_matches_oracle(frames, r, c, N):
    rng = np.random.default_rng(frames)
    base = crand((r, c), rng)
    est = np.stack([base + 0.5 * crand((r, c), rng) for _ in range(frames)])
    got = P.prtf(est, N=N)
    want = O.prtf(est, N)
    assert got.shape == (N, N)
    assert np.max(np.abs(got - want)) < 2e-5


def test_prtf_identical_frames_is_one_and_checks():
    rng = np.random.default_rng(8)
    f = crand((64, 64), rng)
    assert np.allclose(P.prtf(np.stack([f, f, f])), 1.0, atol=1e-5)
    with pytest.raises(ValueError):
        P.prtf(np.stack([f, f]), N=48)
    with pytest.raises(NotImplementedError):
        P.prtf(np.stack([f, f]), prtfq=True)


@pytest.mark.parametrize("centred", [False, True])
@pytest.mark.parametrize("scale", [False, True])
def test_rftf_matches_oracle(centred, scale):
    rng = np.random.default_rng(9)
    obj, s = two_part_object(64, rng)
    est = (obj + 0.05 * crand(obj.shape, rng))[:50, :56]
    g = np.abs(np.fft.fft2(obj)) * (1 + 0.2 * rng.random((64, 64))) + 0.5
    if centred:
        g = np.fft.fftshift(g)
    got = P.rftf(est, g, scale=scale)
    want = O.rftf(est, g, scale=scale)
    assert np.max(np.abs(got - want)) < 2e-5


def test_gpu_rftf_stack_matches_oracle():
    rng = np.random.default_rng(10)
    obj, s = two_part_object(64, rng)
    est = np.stack([(obj + 0.05 * crand(obj.shape, rng))[:48, :52] for _ in range(4)])
    g = np.abs(np.fft.fft2(obj)) * (1 + 0.2 * rng.random((64, 64))) + 0.5
    got = P.gpu_rftf(None, est, g)
    want = O.rftf_stack(est, g)
    assert np.max(np.abs(got - want)) < 2e-5
    # one frame at the corner alignment equals rftf
    one = P.gpu_rftf(None, est[:1], g)
    assert np.max(np.abs(one - O.rftf(est[0], g))) < 2e-5


@pytest.mark.parametrize("stack", [False, True])
def test_rftf_with_ipsf_matches_oracle(stack):
    rng = np.random.default_rng(11)
    obj, s = two_part_object(64, rng)
    g = np.abs(np.fft.fft2(obj)) * (1 + 0.2 * rng.random((64, 64))) + 0.5
    x = np.fft.fftfreq(64) * 64
    psf = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 4.0)
    ipsf = np.fft.fft2(psf / psf.sum()) * np.exp(0.2j)     # complex, to see Re taken
    if stack:
        est = np.stack([(obj + 0.05 * crand(obj.shape, rng))[:50, :50] for _ in range(3)])
        got = P.gpu_rftf(None, est, g, ipsf=ipsf)
        want = O.rftf_stack(est, g, ipsf=ipsf)
    else:
        est = (obj + 0.05 * crand(obj.shape, rng))[:50, :56]
        got = P.rftf(est, g, ipsf=ipsf)
        want = O.rftf(est, g, ipsf=ipsf)
    assert np.max(np.abs(got - want)) < 2e-5


def test_center_of_mass_average_matches_oracle():
    rng = np.random.default_rng(12)
    obj, s = two_part_object(64, rng)
    frames = [np.roll(np.roll(obj, int(rng.integers(-6, 7)), 0), int(rng.integers(-6, 7)), 1)
              * np.exp(1j * rng.uniform(0, 6)) + 0.01 * crand(obj.shape, rng) for _ in range(5)]
    stack = np.stack(frames)[:, :60, :50]
    got = P.center_of_mass_average(stack)
    want = O.center_of_mass_average(stack)
    assert got.shape == (60, 50)
    assert np.max(np.abs(got - want)) < 1e-5 * np.abs(want).max()
