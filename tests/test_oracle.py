"""Pins the complex128 oracle (oracle/phasing_c128.py) on the CPU.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it: reference test
vectors where the reference has them, otherwise closed forms, invariants,
fixed points and brute force.
"""
import math

import numpy as np
import pytest

from oracle import phasing_c128 as O
from oracle.ref_numpy import RefNumpyPhasing

RNG = np.random.default_rng(1234)


def crand(shape, rng=RNG):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# --------------------------------------------------------------------- FFT
def dft_matrix(n):
    k = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(k, k) / n)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_fft2_matches_brute_force_dft(n):
    x = crand((n, n))
    W = dft_matrix(n)
    ref = W @ x @ W.T
    assert np.allclose(O.fft2(x), ref, atol=1e-10 * n * n)
    inv = np.conj(W) @ x @ np.conj(W).T / (n * n)
    assert np.allclose(O.ifft2(x), inv, atol=1e-12 * n * n)


def test_fft2_closed_forms():
    n = 32
    d = np.zeros((n, n), complex)
    d[0, 0] = 1
    assert np.allclose(O.fft2(d), 1)               # impulse -> constant
    assert np.allclose(O.fft2(np.ones((n, n))), n * n * d)  # constant -> delta
    k0, k1 = 3, 5
    yy, xx = np.indices((n, n))
    e = np.exp(2j * np.pi * (k0 * yy + k1 * xx) / n)
    F = O.fft2(e)
    expect = np.zeros((n, n), complex)
    expect[k0, k1] = n * n                          # exponential -> delta
    assert np.allclose(F, expect, atol=1e-9)
    x = crand((n, n))
    # shift theorem: roll by (a, b) multiplies by exp(-2 pi i (ka a + kb b)/n)
    a, b = 4, 7
    ph = np.exp(-2j * np.pi * (yy * a + xx * b) / n)
    assert np.allclose(O.fft2(np.roll(x, (a, b), (0, 1))), O.fft2(x) * ph)
    # Parseval and round trip
    assert math.isclose(np.sum(np.abs(O.fft2(x)) ** 2), n * n * np.sum(np.abs(x) ** 2), rel_tol=1e-12)
    assert np.allclose(O.ifft2(O.fft2(x)), x)


# -------------------------------------------------------------- bounding box
def test_bounding_box_reference_vectors():
    # test/test_mask.py:19-31
    box = np.zeros((256, 256))
    box[18:90, 45:63] = 1
    assert tuple(O.bounding_box(box)) == (18, 90, 45, 63)
    assert tuple(O.bounding_box(box, pad=14)) == (4, 104, 31, 77)


def test_bounding_box_brute_force():
    rng = np.random.default_rng(5)
    for _ in range(50):
        m = np.zeros((40, 50))
        r0, c0 = rng.integers(0, 30), rng.integers(0, 40)
        m[r0:r0 + rng.integers(1, 10), c0:c0 + rng.integers(1, 10)] = rng.uniform(0.1, 1)
        rr, cc = np.nonzero(m > 1e-10)
        pad = int(rng.integers(0, 6))
        exp = (max(rr.min() - pad, 0), min(rr.max() + 1 + pad, 40),
               max(cc.min() - pad, 0), min(cc.max() + 1 + pad, 50))
        assert tuple(O.bounding_box(m, pad=pad)) == exp


def test_bounding_box_square_and_p2():
    m = np.zeros((128, 128))
    m[40:61, 50:58] = 1                      # 21 x 8
    r0, r1, c0, c1 = O.bounding_box(m, force_to_square=True)
    assert r1 - r0 == c1 - c0 == 22
    r0, r1, c0, c1 = O.bounding_box(m, power_of_2=True)
    assert (r1 - r0, c1 - c0) == (32, 8)


# ----------------------------------------------------------- P_M, updates
def known_object(n=32, seed=3):
    rng = np.random.default_rng(seed)
    yy, xx = np.indices((n, n))
    s = ((yy - n / 2) ** 2 + (xx - n / 2) ** 2 < (n / 5) ** 2).astype(float)
    obj = s * rng.uniform(0.5, 1, (n, n)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n)))
    m = np.abs(np.fft.fft2(obj))
    return obj, s, m


def test_fourier_constraint_properties():
    n = 32
    x = crand((n, n))
    m = np.abs(crand((n, n)))
    out, Psi = O.fourier_constraint(x, m)
    F = O.fft2(out)
    assert np.allclose(np.abs(F), m)                              # |P_M| = m
    assert np.allclose(np.angle(F), np.angle(Psi))                 # phase kept
    out2, _ = O.fourier_constraint(out, m)
    assert np.allclose(out2, out)                                  # idempotent
    # zero guard: where Psi = 0 the replaced value is 0, not NaN
    z = np.zeros((n, n), complex)
    o, _ = O.fourier_constraint(z, m)
    assert np.all(np.isfinite(o)) and np.allclose(o, 0)


@pytest.mark.parametrize("sem", ["numpy", "opencl"])
def test_update_special_cases(sem):
    n = 8
    pin, pout, t = crand((n, n)), crand((n, n)), crand((n, n))
    one, zero = np.ones((n, n)), np.zeros((n, n))
    b = 0.7
    # inside the support every non-ER update returns psi_out (ER-numpy: psi_in)
    for alg in ("hio", "raar", "sf"):
        assert np.allclose(O.update(alg, pin, pout, one, b, sem), pout)
    assert np.allclose(O.update("er", pin, pout, one, b, sem), pin if sem == "numpy" else pout)
    # outside
    assert np.allclose(O.update("hio", pin, pout, zero, b, sem), pin - b * pout)
    assert np.allclose(O.update("raar", pin, pout, zero, b, sem), b * pin + (1 - 2 * b) * pout)
    assert np.allclose(O.update("er", pin, pout, zero, b, sem), 0)
    assert np.allclose(O.update("sf", pin, pout, zero, b, sem), -pin if sem == "numpy" else -pout)
    # DM: inside (1+b) psi_out + psi_in - b t, outside psi_in - b psi_out
    assert np.allclose(O.update("dm", pin, pout, one, b, sem, t), (1 + b) * pout + pin - b * t)
    assert np.allclose(O.update("dm", pin, pout, zero, b, sem, t), pin - b * pout)
    # HIO with beta = 0 outside keeps psi_in
    assert np.allclose(O.update("hio", pin, pout, zero, 0.0, sem), pin)
    # DM modified estimate
    assert np.allclose(O.dm_modified(pin, one, b), pin)
    assert np.allclose(O.dm_modified(pin, zero, b), -pin / b)


@pytest.mark.parametrize("sem", ["numpy", "opencl"])
@pytest.mark.parametrize("alg", ["hio", "er", "raar", "sf"])
def test_known_object_is_fixed_point(alg, sem):
    obj, s, m = known_object()
    new, pout, err = O.iteration(obj, m, s, alg, 0.9, sem)
    assert np.allclose(pout, obj, atol=1e-12)
    assert np.allclose(new, obj, atol=1e-12)
    assert err < 1e-13


def test_dm_solution_doubles_inside_support():
    obj, s, m = known_object()
    new, _, _ = O.iteration(obj, m, s, "dm", 0.7)
    assert np.allclose(new, 2 * obj, atol=1e-12)


def test_coherence_divisor_closed_forms():
    rng = np.random.default_rng(21)
    n = 16
    Psi = crand((n, n), rng)
    # ipsf = 1 everywhere: a delta point-spread function, the coherent divisor
    assert np.allclose(O.coherence_divisor(Psi, np.ones((n, n))), np.abs(Psi), rtol=1e-12)
    # a constant ipsf c scales the intensity: div = sqrt(c) |Psi|
    assert np.allclose(O.coherence_divisor(Psi, 4.0 * np.ones((n, n))), 2 * np.abs(Psi))
    # general ipsf: circular convolution of |Psi|^2 with psf = ifft2(ipsf),
    # by brute force (a dropped square, sqrt or abs fails one of these)
    psf = np.exp(-((np.arange(n)[:, None] - 3) ** 2 + (np.arange(n)[None, :] - 5) ** 2) / 4.0)
    ipsf = np.fft.fft2(psf)
    I = np.abs(Psi) ** 2
    conv = np.zeros((n, n))
    for y in range(n):
        for x in range(n):
            conv[y, x] = sum(psf[a, b] * I[(y - a) % n, (x - b) % n]
                             for a in range(n) for b in range(n))
    assert np.allclose(O.coherence_divisor(Psi, ipsf), np.sqrt(np.abs(conv)), rtol=1e-10)
    # the P_M with a coherent ipsf is the coherent P_M
    psi = crand((n, n), rng)
    m = np.abs(crand((n, n), rng))
    a, _, div = O.fourier_constraint_ipsf(psi, m, np.ones((n, n)))
    b, _ = O.fourier_constraint(psi, m)
    assert np.allclose(a, b) and np.allclose(div, np.abs(np.fft.fft2(psi)))
    # |P_M| = m * |Psi| / div
    _, Psi2, div2 = O.fourier_constraint_ipsf(psi, m, ipsf)
    assert np.allclose(np.abs(np.fft.fft2(O.fourier_constraint_ipsf(psi, m, ipsf)[0])),
                       m * np.abs(Psi2) / div2)


def test_radial_rescale_closed_forms():
    rng = np.random.default_rng(31)
    n = 16
    I = rng.random((n, n))
    # one sample, factor 1, weight 1: every node is an integer, identity
    # except the row/column at -n/2 (its floor node is dropped, .cl line 85)
    out = O.radial_rescale(I, [1.0], [1.0])
    keep = np.ones((n, n), bool)
    keep[n // 2, :] = keep[:, n // 2] = False
    assert np.allclose(out[keep], I[keep]) and np.all(out[~keep] == 0)
    # linear in the weights and additive over samples
    a = O.radial_rescale(I, [1.1], [1.0])
    b = O.radial_rescale(I, [0.9], [1.0])
    assert np.allclose(O.radial_rescale(I, [1.1, 0.9], [0.3, 0.7]), 0.3 * a + 0.7 * b)
    # bilinear interpolation reproduces a bilinear function exactly: with
    # I(y, x) = (2 + x)(3 + y) in signed coordinates and factor 2, pixels whose
    # nodes stay inside (-n/2, n/2) read I(y/2, x/2)
    c = np.where(np.arange(n) > n // 2 - 1, np.arange(n) - n, np.arange(n)).astype(float)
    J = (3 + c[:, None]) * (2 + c[None, :])
    got = O.radial_rescale(J, [2.0], [1.0])
    want = (3 + c[:, None] / 2) * (2 + c[None, :] / 2)
    inner = (np.abs(c) <= n // 2 - 2)
    sel = inner[:, None] & inner[None, :]
    assert np.allclose(got[sel], want[sel])
    # by hand: the pixel at (y, x) = (1, 3) with factor 2 reads the midpoint of
    # (0, 1), (0, 2), (1, 1), (1, 2) -> mean of those four values
    got13 = O.radial_rescale(I, [2.0], [1.0])[1, 3]
    assert np.isclose(got13, 0.25 * (I[0, 1] + I[0, 2] + I[1, 1] + I[1, 2]))
    # spectrum rows: rescales are energies over the energy of the peak weight
    r, w = O.spectrum_rescales(np.array([[90.0, 100.0, 110.0], [0.2, 1.0, 0.3]]))
    assert np.allclose(r, [0.9, 1.0, 1.1]) and np.allclose(w, [0.2, 1.0, 0.3])


def test_richardson_lucy_closed_forms():
    rng = np.random.default_rng(41)
    n = 16
    est = crand((n, n), rng)
    # the start is the fftshifted unnormalised Gaussian of width n/4
    g = O.rl_start_gaussian(n)
    assert g[0, 0] == 1.0 and np.isclose(g[0, 4], np.exp(-0.5))
    # zero iterations: the ipsf is fft2 of the start, real part, over |.[0,0]|
    want = np.fft.fft2(g).real
    assert np.allclose(O.rl_coherence(est, np.abs(crand((n, n), rng)), 0), want / abs(want[0, 0]))
    # a datum consistent with the start, m^2 proportional to u (*) p, makes
    # d / (u (*) p) a constant; with the plain reversal Ph = P[::-1, ::-1] the
    # correction ifft2(fft2(const) Ph) is const * P[n-1, n-1] (the proper
    # adjoint would give const * sum(p)), so one iteration multiplies u by a
    # complex constant whose phase survives the |u[0,0]| normalisation
    p = np.abs(np.fft.fft2(est)) ** 2
    p /= p.sum()
    P = np.fft.fft2(p)
    conv = np.fft.ifft2(np.fft.fft2(g) * P).real     # u (*) p, real and > 0
    m = np.sqrt(conv) * 3.0                           # the scale cancels in d
    one = O.rl_coherence(est, m, 1, semantics="opencl")
    want = np.fft.fft2(g * P[n - 1, n - 1])
    assert np.allclose(one, want / abs(want[0, 0]), atol=1e-9)


def test_gaussian_coherence_objective_limits():
    rng = np.random.default_rng(42)
    n = 16
    est = crand((n, n), rng)
    m = np.abs(crand((n, n), rng))
    ac = O.autocorrelation(est, n)
    # very long coherence: g = 1, sqrt|fft2(ac)| = |fft2(est)|
    far = O.gaussian_coherence_objective(ac, m, (1e9, 1e9))
    assert np.isclose(far, np.sum((np.abs(np.fft.fft2(est)) - m) ** 2))
    # very short coherence: g = delta at 0, a flat pattern of sqrt(sum|est|^2)
    near = O.gaussian_coherence_objective(ac, m, (1e-3, 1e-3))
    assert np.isclose(near, np.sum((np.sqrt(np.sum(np.abs(est) ** 2)) - m) ** 2))


def test_modulus_error_closed_form():
    Psi = np.array([[3 + 4j, 0], [1, -2j]])
    m = np.array([[4.0, 1.0], [1.0, 1.0]])
    # |Psi| = 5, 0, 1, 2 ; diffs 1, -1, 0, 1 ; sum m^2 = 19
    assert math.isclose(O.modulus_error(Psi, m), math.sqrt(3 / 19), rel_tol=1e-15)


def test_expand_order_reference_example():
    # phasing.py:203-213
    seq = O.expand_order((("hio", 5), ("er", 5), ("sf", 1)), 13)
    assert seq == ["hio"] * 5 + ["er"] * 5 + ["sf", "hio", "hio"]
    assert O.expand_order(None, 101) == ["hio"] * 99 + ["er"] + ["hio"]


def test_seed_matches_numpy_legacy_global():
    n = 16
    s = (np.arange(n * n).reshape(n, n) % 3 == 0).astype(np.float32)
    g = O.seed_guess(n, 42, s)
    np.random.seed(42)
    ref = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64) * s
    assert np.array_equal(g, ref.astype(np.complex128))


# ---------------------------------------------------------------- shrinkwrap
def test_gaussian_kernel_and_blur():
    n = 64
    k = O.gaussian_kernel((n, n), 2.0)
    assert math.isclose(k.sum(), 1.0, rel_tol=1e-14)
    assert np.unravel_index(k.argmax(), k.shape) == (0, 0)
    assert np.allclose(O.blur_fft(np.full((n, n), 3.0), 2.0), 3.0)
    d = np.zeros((n, n))
    d[0, 0] = 1
    assert np.allclose(O.blur_fft(d, 2.0), k, atol=1e-14)
    # separable direct periodic convolution gives the same blur
    x = np.abs(crand((n, n)))
    g = np.exp(-((np.arange(n) - n // 2) ** 2) / (2 * 2.0 ** 2))
    g = np.fft.ifftshift(g) / g.sum()
    rows = np.array([[np.sum(g * np.roll(x[r], -c)) for c in range(n)] for r in range(n)])
    both = np.array([[np.sum(g * np.roll(rows[:, c], -r)) for c in range(n)] for r in range(n)]).T
    assert np.allclose(O.blur_fft(x, 2.0), both.T, atol=1e-12)


def test_shrinkwrap_support_invariants():
    obj, s, m = known_object(64)
    psi = obj + 0.05 * crand(obj.shape)
    new, b = O.shrinkwrap_support(psi, s, 1.5, 0.2)
    assert np.all(new <= s)
    keep0, _ = O.shrinkwrap_support(psi, s, 1.5, 0.0)
    assert np.array_equal(keep0, s * (b > 0))
    assert np.array_equal(new, s * (b > 0.2 * b.max()))


def test_sigma_schedule():
    sig = O.sigma_schedule(50)
    assert sig[0] == 3.0 and math.isclose(sig[-1], 1.5) and len(sig) == 50
    assert all(a > b for a, b in zip(sig, sig[1:]))


def test_find_all_objects_matches_label():
    from scipy import ndimage
    rng = np.random.default_rng(9)
    for _ in range(5):
        m = (rng.random((24, 24)) > 0.8).astype(float)
        m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0   # no wrap-around contacts
        parts = O.find_all_objects(m)
        lab, n = ndimage.label(m, structure=np.ones((3, 3)))
        big = [lab == i for i in range(1, n + 1) if (lab == i).sum() > 1]
        got = [p.astype(bool) for p in parts if p.sum() > 0]
        assert len(got) == len(big)
        for p in got:
            assert any(np.array_equal(p, q) for q in big)
        # isolated pixels give empty parts (flood fill never marks the start)
        n_iso = sum(1 for i in range(1, n + 1) if (lab == i).sum() == 1)
        assert sum(1 for p in parts if p.sum() == 0) == n_iso


def test_find_all_objects_wraps():
    m = np.zeros((10, 10))
    m[0, 3:5] = 1
    m[9, 3:5] = 1                 # touches row 0 across the edge
    parts = O.find_all_objects(m)
    assert len(parts) == 1 and parts[0].sum() == 4


def test_refine_support_invariants():
    obj, s, m = known_object(64)
    avg = obj + 0.1 * crand(obj.shape)
    ref, blurred = O.refine_support(s, avg, blur=2, local_threshold=0.2, global_threshold=0.1)
    assert np.all(ref <= s)                                     # shrink only
    keep, _ = O.refine_support(s, avg, blur=2, local_threshold=0.0, global_threshold=0.0)
    assert np.array_equal(keep, s * (blurred > 0))
    const, _ = O.refine_support(s, np.ones_like(s), blur=2, local_threshold=0.5)
    assert np.array_equal(const, s)


# ------------------------------------------------------------- align phase
def test_align_closed_form_matches_fminbound():
    rng = np.random.default_rng(11)
    data = crand((6, 20, 24), rng) + (rng.standard_normal((6, 1, 1)) * 3)
    data = data * np.exp(1j * rng.uniform(0, 2 * np.pi, (6, 1, 1)))
    a = O.align_global_phase(data)
    b = O.align_global_phase_closed(data)
    assert np.allclose(a, b, atol=1e-7 * np.abs(data).max())
    for f in b:
        assert f.real.sum() >= 0
        # no rotation lowers the imaginary energy further
        for p in np.linspace(0, 2 * np.pi, 37):
            assert np.sum((f * np.exp(1j * p)).imag ** 2) >= np.sum(f.imag ** 2) - 1e-9


def test_align_is_rotation_invariant():
    rng = np.random.default_rng(12)
    f = crand((16, 16), rng) + 2
    a = O.align_global_phase_closed(f)
    b = O.align_global_phase_closed(f * np.exp(1j * 1.234))
    assert np.allclose(a, b, atol=1e-12)


# ------------------------------------------------------------- prtf / rftf
def test_prtf_rftf_closed_forms():
    obj, s, m = known_object(32)
    frames = np.stack([obj, obj, obj])
    assert np.allclose(O.prtf(frames, 32), 1.0)
    rng = np.random.default_rng(4)
    rnd = np.stack([crand((32, 32), rng) for _ in range(400)])
    assert abs(O.prtf(rnd, 32).mean() - math.sqrt(math.pi / 4 / 400)) < 0.01
    assert np.allclose(O.rftf(obj, m + 1e-30), 1.0)


def test_rftf_alignment_and_scale():
    obj, s, m = known_object(32)
    g = m + 1e-3
    fa = np.abs(np.fft.fft2(obj))
    # goal twice the modulus: err = 1/4 everywhere -> sqrt(1/1.25)
    assert np.allclose(O.rftf(obj, 2 * fa + 1e-30), math.sqrt(1 / 1.25))
    # a centre-shifted goal is detected and gives the same map
    assert np.allclose(O.rftf(obj, np.fft.fftshift(g)), O.rftf(obj, g))
    # scale maps 1/sqrt(2) (err = 1) to 0 and 1 to 1
    r = O.rftf(obj, fa + 1e-30, scale=True)
    assert np.allclose(r, 1.0)
    assert np.allclose(O.rftf(obj, 0.5 * fa + 1e-30, scale=True), 0.0, atol=1e-12)
    # one frame at the corner alignment: the stack form equals rftf
    assert np.allclose(O.rftf_stack(obj[None], g), O.rftf(obj, g))
    # a stack averages the per-frame values
    rng = np.random.default_rng(2)
    a, b = obj + 0.1 * crand(obj.shape, rng), obj + 0.1 * crand(obj.shape, rng)
    want = 0.5 * (O.rftf_stack(a[None], g) + O.rftf_stack(b[None], g))
    assert np.allclose(O.rftf_stack(np.stack([a, b]), g), want)


def test_rftf_ipsf_closed_forms():
    obj, s, m = known_object(32)
    g = m + 1e-3
    ones = np.ones((32, 32))
    # gpu_rftf blurs the intensity: a delta point-spread function changes nothing
    assert np.allclose(O.rftf_stack(obj[None], g, ipsf=ones), O.rftf_stack(obj[None], g))
    # rftf blurs the modulus and takes a square root (phasing.py:1573): with a
    # delta point-spread function the compared pattern is sqrt|F|
    fa = np.abs(np.fft.fft2(obj))
    sq = np.sqrt(fa)
    goal = sq * 2 + 1e-30                     # err = 1/4 against the sqrt|F| pattern
    assert np.allclose(O.rftf(obj, goal, ipsf=ones), np.sqrt(1 / 1.25))
    # only the real part of the ipsf enters the stack form
    ip = np.fft.fft2(np.exp(-np.arange(32)[:, None] ** 2 / 3.0) * np.ones((1, 32)))
    assert np.allclose(O.rftf_stack(obj[None], g, ipsf=ip * np.exp(0.7j)),
                       O.rftf_stack(obj[None], g, ipsf=(ip * np.exp(0.7j)).real))


def test_center_of_mass_average_closed_forms():
    rng = np.random.default_rng(51)
    a = np.zeros((20, 24), complex)
    a[5:9, 6:11] = crand((4, 5), rng)
    # copies of one frame shifted by whole pixels sum back onto frame 0
    stack = np.stack([a, np.roll(np.roll(a, 3, 0), -2, 1), np.roll(a, 5, 1)])
    assert np.allclose(O.center_of_mass_average(stack), 3 * a)
    # a single frame is returned as is
    assert np.allclose(O.center_of_mass_average(a), a)


def test_blur_kernel_odd_size_peak():
    # centre at n//2 then fftshift: the peak lands at n - 1 for odd n
    k = O.gaussian_kernel((9, 7), 1.5)
    assert np.unravel_index(k.argmax(), k.shape) == (8, 6)
    assert math.isclose(k.sum(), 1.0, rel_tol=1e-14)


# ------------------------------------------------ restated reference path
def test_ref_numpy_one_iteration_matches_oracle():
    obj, s, m = known_object(32)
    rng = np.random.default_rng(2)
    psi = (s * (rng.random((32, 32)) + 1j * rng.random((32, 32)))).astype(np.complex64)
    for alg in ("hio", "er", "raar", "sf", "dm"):
        r = RefNumpyPhasing(m, s, beta=0.8)
        r.psi_in = psi.copy()
        r.iteration(alg)
        ref, _, _ = O.iteration(psi.astype(np.complex128), m.astype(np.float32), s, alg, 0.8)
        assert np.linalg.norm(r.psi_in - ref) / np.linalg.norm(ref) < 1e-5, alg
