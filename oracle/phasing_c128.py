"""complex128 oracle of ``speckle.phasing`` (reference ``dhparks/als_speckle``).

TEST INFRASTRUCTURE ONLY -- the checker for the HIP path, never the product.

Everything here is a restatement, in numpy complex128 / float64, of the
reference's CPU (numpy) phasing path.  Each function cites the reference
``file:line`` it follows.  Where the reference computes nothing (the
modulus-error trace and the in-loop shrinkwrap), the quantity is defined here
and the definition is the contract.

Pinning (how each function is checked, see ``tests/test_oracle.py``):

========================  ===================================================
function                  pinned by
========================  ===================================================
fft2 / ifft2              brute-force matrix DFT (N <= 32), impulse/constant,
                          single exponential -> delta, shift theorem,
                          Parseval, inverse round trip
bounding_box              the reference's own vectors ``test/test_mask.py:19-31``
                          plus brute force on random masks
fourier_constraint        |P_M(psi)| == m, phase preserved, idempotence,
                          zero-guard special case
update (all algorithms)   special cases s in {0, 1}; a known object inside its
                          support with a noise-free modulus is a fixed point of
                          ER/HIO/RAAR/SF (both semantics); DM grows by 2x
                          inside the support at a solution (reference quirk)
expand_order              the worked example in ``phasing.py:203-213``
seed                      numpy legacy MT19937 draws (library-pinned)
modulus_error             zero at a noise-free solution; closed form on a
                          hand-computed 2x2 case
gaussian_kernel / blur    unit sum, blur of a constant is the constant,
                          separable-vs-FFT agreement, blur of a delta is the
                          normalised Gaussian
shrinkwrap_support        invariants: result <= support, tau=0 keeps the
                          support wherever the blur is positive
refine_support            invariants (shrink-only, tau=0 keeps parts,
                          constant image keeps everything), flood fill vs
                          scipy.ndimage.label (8-connectivity)
align_global_phase        closed-form minor-eigenvector vs the reference's
                          ``fminbound`` call as written (scipy present)
prtf                      identical frames -> 1 everywhere; conjugate-free
                          random phases -> ~1/sqrt(T)
rftf                      perfect estimate -> 1 everywhere
========================  ===================================================
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "fft2", "ifft2", "bounding_box", "expand_order", "seed_guess",
    "fourier_constraint", "modulus_error", "update", "iteration", "iterate",
    "gaussian_kernel", "blur_fft", "shrinkwrap_support", "sigma_schedule",
    "trace_object", "find_all_objects", "refine_support",
    "align_global_phase", "align_global_phase_closed", "prtf", "rftf",
    "roi_copy", "ALGORITHMS",
]

ALGORITHMS = ("hio", "er", "raar", "sf", "dm")


# ---------------------------------------------------------------------------
# FFT (reference: phasing.py:1067-1086, pyfft.cl.Plan at phasing.py:372-374)
# ---------------------------------------------------------------------------
def fft2(x):
    """Forward 2-D DFT, unnormalised, complex128.  ``phasing.py:1081-1083``."""
    return np.fft.fft2(np.asarray(x, dtype=np.complex128))


def ifft2(x):
    """Inverse 2-D DFT with 1/(rows*cols), complex128.  ``phasing.py:1078-1080``."""
    return np.fft.ifft2(np.asarray(x, dtype=np.complex128))


# ---------------------------------------------------------------------------
# masking.bounding_box (masking.py:7-159)
# ---------------------------------------------------------------------------
def bounding_box(data, threshold=1e-10, force_to_square=False, pad=0,
                 power_of_2=False):
    """Minimal bounding box (rmin, rmax, cmin, cmax), ``masking.py:7-159``.

    Pixels strictly above ``threshold`` count (``masking.py:140``); the box is
    padded and clipped (``masking.py:147-151``), optionally grown to powers of
    two (``masking.py:35-84``, ``153-155``) and squared (``masking.py:86-117``,
    ``157-159``), then shifted back inside the array if negative
    (``masking.py:161-165`` of the "panic section").
    """
    data = np.asarray(data)
    assert data.ndim == 2
    rows, cols = data.shape
    aw = np.argwhere(data > threshold)
    rmin, cmin = (int(v) for v in aw.min(0))
    rmax, cmax = (int(v) + 1 for v in aw.max(0))
    rmin = 0 if rmin < pad else rmin - pad
    rmax = rows if rmax + pad > rows else rmax + pad
    cmin = 0 if cmin < pad else cmin - pad
    cmax = cols if cmax + pad > cols else cmax + pad

    def next_p2(v):
        if v & (v - 1) == 0:
            return v
        return int(2 ** (math.floor(math.log(v, 2)) + 1))

    if power_of_2:
        br, bc = rmax - rmin, cmax - cmin
        rp, cp = next_p2(br), next_p2(bc)
        if rp >= rows:
            rmin, rmax = 0, rows
        else:
            d = rp - br
            lo = d // 2
            rmin -= lo
            rmax += d - lo
        if cp >= cols:
            cmin, cmax = 0, cols
        else:
            d = cp - bc
            lo = d // 2
            cmin -= lo
            cmax += d - lo

    if force_to_square:
        dr, dc = rmax - rmin, cmax - cmin
        if dr % 2 == 1:
            dr += 1
            if rmax < rows:
                rmax += 1
            else:
                rmin -= 1
        if dc % 2 == 1:
            dc += 1
            if cmax < cols:
                cmax += 1
            else:
                cmin -= 1
        # Python-2 integer division of ints (masking.py:108-116)
        if dr > dc:
            ac = (cmax + cmin) // 2
            cmin, cmax = ac - dr // 2, ac + dr // 2
        if dc > dr:
            ar = (rmax + rmin) // 2
            rmin, rmax = ar - dc // 2, ar + dc // 2

    if rmin < 0:
        rmin, rmax = 0, rmax - rmin
    if cmin < 0:
        cmin, cmax = 0, cmax - cmin
    return np.array([rmin, rmax, cmin, cmax]).astype(int)


# ---------------------------------------------------------------------------
# iterate(): schedule expansion (phasing.py:244-260)
# ---------------------------------------------------------------------------
def expand_order(order, iterations):
    """Algorithm per iteration, ``phasing.py:244-260``.

    ``order`` defaults to (('hio', 99), ('er', 1)); the expanded list is cycled
    with ``order_list[iteration % k]``.
    """
    if order is None:
        order = (("hio", 99), ("er", 1))
    order_list = []
    for entry in order:
        order_list.extend([str(entry[0])] * int(entry[1]))
    k = len(order_list)
    return [order_list[i % k] for i in range(int(iterations))]


# ---------------------------------------------------------------------------
# seed() (phasing.py:990-1031)
# ---------------------------------------------------------------------------
def seed_guess(n, val, support):
    """Random start, ``phasing.py:1002-1006`` and ``1026-1031``.

    ``np.random.seed(val)``; ``rand(N,N) + 1j*rand(N,N)`` (real part drawn
    first); cast to complex64 (``phasing.py:1026``); multiplied by the support.
    Returned as complex128 holding the complex64 values.
    """
    rs = np.random.RandomState(val)
    re = rs.random_sample((n, n))
    im = rs.random_sample((n, n))
    g = (re + 1j * im).astype(np.complex64)
    return g.astype(np.complex128) * np.asarray(support, np.float64)


# ---------------------------------------------------------------------------
# Fourier constraint (phasing.py:1187-1228, _build_divisor 1109-1185)
# ---------------------------------------------------------------------------
def fourier_constraint(psi, m, zero_guard=True):
    """P_M: ``psi_out = ifft2(m * Psi / |Psi|)`` with ``Psi = fft2(psi)``.

    ``phasing.py:1192`` (forward), ``1177`` (divisor ``np.abs``), ``1217``
    (``m*pf/np.abs(fd)``), ``1223`` (inverse).  The reference has no zero
    guard (0/0 gives NaN, ``phasing_fourier.cl:13``); with ``zero_guard`` the
    replaced value is 0 where |Psi| == 0, the documented deviation.
    Returns (psi_out, Psi).
    """
    Psi = fft2(psi)
    a = np.abs(Psi)
    m = np.asarray(m, np.float64)
    if zero_guard:
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(a > 0, m / np.where(a > 0, a, 1.0), 0.0)
        Pp = Psi * r
    else:
        Pp = m * Psi / a
    return ifft2(Pp), Psi


def modulus_error(Psi, m):
    """Build-defined modulus error E = sqrt(sum((|Psi|-m)^2) / sum(m^2)).

    The reference computes no error in ``iterate`` (``phasing.py:260-276``);
    its nearest precedents are ``phasing.py:686-698`` and ``1584``.  FP64.
    """
    m = np.asarray(m, np.float64)
    d = np.abs(Psi) - m
    return math.sqrt(float(np.sum(d * d)) / float(np.sum(m * m)))


# ---------------------------------------------------------------------------
# Real-space updates (phasing.py:1231-1328; kernels phasing_*.cl)
# ---------------------------------------------------------------------------
def update(alg, psi_in, psi_out, s, beta, semantics="numpy", dm_t=None):
    """One real-space update as the reference writes it.

    semantics='numpy' follows the CPU branches:
      hio  ``phasing.py:1238-1243``: (1-s)*(psi_in-beta*psi_out)+s*psi_out
      er   ``phasing.py:1256-1258``: s*psi_in
      raar ``phasing.py:1267-1273``: (1-s)*(beta*psi_in+(1-2beta)*psi_out)+s*psi_out
      sf   ``phasing.py:1281-1284``: (1-s)*(-psi_in)+s*psi_out
      dm   ``phasing.py:1302-1303``, ``1321-1325``
    semantics='opencl' follows the GPU kernels where they differ:
      er   ``phasing_er.cl:13-15``: s*psi_out
      sf   ``phasing_sf.cl:16-17``: (1-s)*(-psi_out)+s*psi_out
    (hio, raar and dm are the same formulas in both: ``phasing_hio.cl:16-17``,
    ``phasing_raar.cl:10,19-20``, ``phasing_dm1.cl:16``, ``phasing_dm2.cl:21-28``.)
    ``dm_t`` is P_M applied to ``s*psi_in - gamma_m*(1-s)*psi_in``.
    """
    s = np.asarray(s, np.float64)
    if alg == "hio":
        return (1 - s) * (psi_in - beta * psi_out) + s * psi_out
    if alg == "er":
        if semantics == "numpy":
            return s * psi_in
        return s * psi_out
    if alg == "raar":
        return (1 - s) * (beta * psi_in + (1 - 2 * beta) * psi_out) + s * psi_out
    if alg == "sf":
        if semantics == "numpy":
            return (1 - s) * (-1 * psi_in) + s * psi_out
        return (1 - s) * (-1 * psi_out) + s * psi_out
    if alg == "dm":
        gamma_m = 1.0 / beta
        gamma_s = -gamma_m
        t1 = 2 * psi_out - beta * dm_t + beta * ((1 + gamma_s) * psi_out - gamma_s * psi_in)
        t2 = psi_in - beta * psi_out
        return s * t1 + (1 - s) * t2
    raise AssertionError("real space enforcement algorithm %s is unknown" % alg)


def dm_modified(psi_in, s, beta):
    """DM modified estimate ``s*psi_in - gamma_m*(1-s)*psi_in``, gamma_m=1/beta.

    ``phasing.py:1293-1303``, ``phasing_dm1.cl:16``.
    """
    s = np.asarray(s, np.float64)
    gamma_m = 1.0 / beta
    return s * psi_in - gamma_m * (1 - s) * psi_in


def coherence_divisor(Psi, ipsf):
    """Partially coherent modulus, ``_build_divisor`` with an ipsf, CPU branch
    without numexpr (``phasing.py:1175-1183``): div = |Psi|**2, forward
    transform, times ipsf, inverse transform, sqrt(abs(.)).  ``ipsf`` is the
    transform of the point-spread function, corner-centred, as loaded
    (``phasing.py:414-425``)."""
    return np.sqrt(np.abs(ifft2(fft2(np.abs(Psi) ** 2) * ipsf)))


def radial_rescale(I, rescales, weights):
    """Spectral (longitudinal) blur of an intensity, ``phasing_radial_rescale.cl``
    (called at ``phasing.py:1140-1156``), in FP64.

    Pixel (row j, column i) of an n x n corner-centred array has signed
    coordinates x = i (i < n/2) or i - n, y likewise from j.  For every
    spectral sample k the intensity is bilinearly interpolated at
    (x / r_k, y / r_k) and summed with weight w_k.  Interpolation nodes with
    floor coordinate <= -n/2 or >= n/2 (and their +1 neighbours likewise)
    get weight 0 (``.cl`` lines 85-92); kept nodes wrap back by +n.
    """
    I = np.asarray(I, np.float64)
    n = I.shape[0]
    h = n // 2
    idx = np.arange(n)
    coord = np.where(idx > h - 1, idx - n, idx).astype(np.float64)
    X = np.broadcast_to(coord[None, :], (n, n))
    Y = np.broadcast_to(coord[:, None], (n, n))
    out = np.zeros((n, n))

    def node(c):
        keep = (c > -h) & (c < h)
        return np.where(keep, np.where(c < 0, c + n, c), 0).astype(int), keep

    for rv, sw in zip(np.asarray(rescales, np.float64), np.asarray(weights, np.float64)):
        rx, ry = X / rv, Y / rv
        x1 = np.floor(rx)
        y1 = np.floor(ry)
        x2, y2 = x1 + 1, y1 + 1
        (ix1, kx1), (ix2, kx2) = node(x1), node(x2)
        (iy1, ky1), (iy2, ky2) = node(y1), node(y2)
        wx1 = (1 - np.abs(x1 - rx)) * kx1
        wx2 = (1 - np.abs(x2 - rx)) * kx2
        wy1 = (1 - np.abs(y1 - ry)) * ky1
        wy2 = (1 - np.abs(y2 - ry)) * ky2
        iv = (I[iy1, ix1] * wx1 * wy1 + I[iy2, ix1] * wx1 * wy2 +
              I[iy1, ix2] * wx2 * wy1 + I[iy2, ix2] * wx2 * wy2)
        out += iv * sw
    return out


def spectrum_rescales(spectrum):
    """``load(spectrum=)`` (``phasing.py:430-448``): rows (energies, weights);
    rescale factors are energies / the energy of the largest weight."""
    energies, weights = np.asarray(spectrum[0], np.float64), np.asarray(spectrum[1], np.float64)
    return energies / energies[weights.argmax()], weights


def fourier_constraint_ipsf(psi, m, ipsf, zero_guard=True, spectrum=None):
    """P_M with the partial-coherence divisor: ``m * Psi / div``
    (``phasing.py:1217``, ``fd`` = ``coherence_divisor``; without an ipsf the
    coherent |Psi|).  With ``spectrum`` = (rescales, weights) the divisor is
    then squared, spectrally blurred and square-rooted (``phasing.py:1140-
    1156``).  Returns (psi_out, Psi, div); 0 where div == 0 with
    ``zero_guard``."""
    Psi = fft2(psi)
    div = coherence_divisor(Psi, ipsf) if ipsf is not None else np.abs(Psi)
    if spectrum is not None:
        div = np.sqrt(radial_rescale(div ** 2, spectrum[0], spectrum[1]))
    m = np.asarray(m, np.float64)
    if zero_guard:
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(div > 0, m / np.where(div > 0, div, 1.0), 0.0)
        Pp = Psi * r
    else:
        Pp = m * Psi / div
    return ifft2(Pp), Psi, div


def modulus_error_div(div, m):
    """Error of the modelled modulus: sqrt(sum((div - m)^2) / sum(m^2)); with
    a coherent divisor (div = |Psi|) it is ``modulus_error``."""
    m = np.asarray(m, np.float64)
    return float(np.sqrt(np.sum((div - m) ** 2) / np.sum(m ** 2)))


def iteration(psi_in, m, s, alg, beta=0.8, semantics="numpy", zero_guard=True, ipsf=None,
              spectrum=None):
    """One iteration of ``_iteration`` (``phasing.py:1088-1353``).

    Returns (psi_new, psi_out, err) where err is the modulus error of
    Psi = fft2(psi_in) before replacement (of the partially coherent divisor
    when an ipsf or a spectrum (rescales, weights) is given).
    """
    if ipsf is None and spectrum is None:
        psi_out, Psi = fourier_constraint(psi_in, m, zero_guard)
        err = modulus_error(Psi, m)
    else:
        psi_out, Psi, div = fourier_constraint_ipsf(psi_in, m, ipsf, zero_guard, spectrum)
        err = modulus_error_div(div, m)
    t = None
    if alg == "dm":
        if ipsf is None and spectrum is None:
            t, _ = fourier_constraint(dm_modified(psi_in, s, beta), m, zero_guard)
        else:
            t, _, _ = fourier_constraint_ipsf(dm_modified(psi_in, s, beta), m, ipsf, zero_guard,
                                              spectrum)
    return update(alg, psi_in, psi_out, s, beta, semantics, t), psi_out, err


# ---------------------------------------------------------------------------
# Shrinkwrap (build-defined in-loop form of refine_support, phasing.py:1747-1815)
# ---------------------------------------------------------------------------
def gaussian_kernel(shape, sigma):
    """Normalised, corner-centred Gaussian, ``phasing.py:1787-1788``.

    ``shape.gaussian(size, (blur, blur))`` centres at ``size/2``
    (``shape.py:462-465``, Python-2 integer division), evaluates the separable
    product (``shape.py:498-512``); the reference then applies ``fftshift``
    and divides by the sum.
    """
    r, c = shape
    g1 = np.exp(-(np.arange(r, dtype=float) - (r // 2)) ** 2 / (2.0 * sigma ** 2))
    g2 = np.exp(-(np.arange(c, dtype=float) - (c // 2)) ** 2 / (2.0 * sigma ** 2))
    k = np.fft.fftshift(np.outer(g1, g2))
    return k / k.sum()


def blur_fft(img, sigma):
    """``Re ifft2(fft2(K) * fft2(img))``, ``phasing.py:1789-1790``."""
    k = gaussian_kernel(img.shape, sigma)
    return np.fft.ifft2(np.fft.fft2(k) * np.fft.fft2(np.asarray(img, np.float64))).real


def sigma_schedule(events, sigma0=3.0, sigma1=1.5):
    """Linear blur anneal sigma0 -> sigma1 over ``events`` shrinkwrap events."""
    if events <= 1:
        return [float(sigma0)] * max(events, 0)
    return [sigma0 + (sigma1 - sigma0) * e / (events - 1) for e in range(events)]


def shrinkwrap_support(psi, s, sigma, tau):
    """In-loop shrinkwrap: S' = S * [G_sigma (*) |psi| > tau * max].

    Build-defined in-loop form of ``refine_support`` (``phasing.py:1792-1809``)
    with local_threshold 0, one part per binary support (``find_all_objects``
    then returns the support itself) and global_threshold tau.  Returns
    (new_support, blurred).
    """
    b = blur_fft(np.abs(psi), sigma)
    mx = float(b.max())
    keep = b > mx * tau
    return np.asarray(s, np.float64) * keep, b


# ---------------------------------------------------------------------------
# iterate() (phasing.py:181-289) with the build's additive kwargs
# ---------------------------------------------------------------------------
def iterate(psi, m, s, iterations, order=None, beta=0.8, semantics="numpy",
            shrinkwrap=None, zero_guard=True, record=None):
    """Run ``iterations`` iterations; returns (psi, trace, support).

    ``phasing.py:260-268``.  ``shrinkwrap`` is None or a dict with keys
    ``every`` (20), ``sigma0`` (3), ``sigma1`` (1.5), ``tau`` (0.2): after
    iteration k (1-based) with k % every == 0 the support is replaced by
    ``shrinkwrap_support`` of the current iterate, sigma following
    ``sigma_schedule(iterations // every)``.  ``record`` (optional list) gets
    a copy of psi before every iteration.
    """
    psi = np.asarray(psi, np.complex128)
    s = np.asarray(s, np.float64)
    algs = expand_order(order, iterations)
    trace = np.zeros(len(algs))
    sw = None
    if shrinkwrap:
        sw = dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)
        sw.update(shrinkwrap)
        sigmas = sigma_schedule(len(algs) // sw["every"], sw["sigma0"], sw["sigma1"])
    for k, alg in enumerate(algs):
        if record is not None:
            record.append(psi.copy())
        psi, _, trace[k] = iteration(psi, m, s, alg, beta, semantics, zero_guard)
        if sw and (k + 1) % sw["every"] == 0:
            e = (k + 1) // sw["every"] - 1
            s, _ = shrinkwrap_support(psi, s, sigmas[e], sw["tau"])
    return psi, trace, s


def roi_copy(psi, r0, rows, c0, cols):
    """``savebuffer[n] = psi_in[r0:r0+rows, c0:c0+cols]``, ``phasing.py:284-287``."""
    return np.asarray(psi)[r0:r0 + rows, c0:c0 + cols].copy()


# ---------------------------------------------------------------------------
# masking.trace_object / find_all_objects (masking.py:271-434)
# ---------------------------------------------------------------------------
def trace_object(data, start, detect_val):
    """8-connected flood fill with periodic wrap, ``masking.py:271-377``.

    Returns a bool mask.  Pixels with ``detect_val[0] <= v <= detect_val[1]`` are connected; the
    neighbour list is the 8-neighbourhood (``masking.py:305-309``) and indices
    wrap modulo the shape (``masking.py:356``, ``370-372``; negative indices
    wrap through numpy's own negative indexing).
    """
    data = np.asarray(data)
    n, mm = data.shape
    working = (data >= detect_val[0]) & (data <= detect_val[1])
    detected = np.zeros(data.shape, bool)
    if not working[start]:
        return detected
    # the start pixel is only marked when reached as a neighbour of a
    # neighbour (masking.py:362-373), so an isolated pixel yields an empty part
    edge = [start]
    while edge:
        new = []
        for (y, x) in edge:
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dy == 0 and dx == 0:
                        continue
                    q = ((y + dy) % n, (x + dx) % mm)
                    if working[q] and not detected[q]:
                        detected[q] = True
                        new.append(q)
        edge = new
    return detected


def find_all_objects(data, detect_val=(0.95, 1.05)):
    """All connected parts, in raster order of first pixel, ``masking.py:379-434``.

    Works on a copy (the reference zeroes its input, ``masking.py:428``).
    Returns an int array (parts, rows, cols).
    """
    data = np.array(data, dtype=np.float64, copy=True)
    parts = []
    for r in range(data.shape[0]):
        if not data[r].any():
            continue
        for c in range(data.shape[1]):
            if detect_val[0] <= data[r, c] <= detect_val[1]:
                found = trace_object(data, (r, c), detect_val)
                parts.append(found.astype(int))
                data[found] = 0
    if not parts:
        return np.zeros((0,) + data.shape, int)
    return np.array(parts).astype(int)


def refine_support(support, average_mag, blur=3, local_threshold=0.2,
                   global_threshold=0, kill_weakest=False):
    """Marchesini shrinkwrap as written, ``phasing.py:1747-1815``.

    Returns (refined, blurred).  ``refined`` is the union over parts of the
    local-threshold pass, times the global pass, optionally minus the
    weakest part.
    """
    support = np.asarray(support)
    refined = np.zeros(support.shape, np.float64)
    average_mag = np.abs(average_mag)
    blurred = blur_fft(average_mag, blur)
    global_passed = np.where(blurred > blurred.max() * global_threshold, 1, 0)
    parts = find_all_objects(support)
    part_sums = np.zeros(len(parts))
    for n, part in enumerate(parts):
        current = blurred * part
        refined += np.where(current > current.max() * local_threshold, 1, 0)
        if kill_weakest:
            part_sums[n] = np.sum(average_mag * part)
    refined *= global_passed
    if kill_weakest and len(parts):
        refined *= 1 - parts[part_sums.argmin()]
    return refined, blurred


# ---------------------------------------------------------------------------
# align_global_phase (phasing.py:1355-1404)
# ---------------------------------------------------------------------------
def align_global_phase(data_in):
    """As written: ``fminbound`` of sum|Re sin p + Im cos p|^2 on [0, 2pi].

    ``phasing.py:1382-1399``: rotate by exp(i p*), then negate the frame if
    sum(-Re) > sum(Re).
    """
    from scipy.optimize import fminbound
    data = np.array(data_in, dtype=np.complex128, copy=True)
    was2d = data.ndim == 2
    if was2d:
        data = data[None]
    for frame in data:
        fr, fi = frame.real.copy(), frame.imag.copy()
        e = lambda p: np.sum(np.abs(fr * math.sin(p) + fi * math.cos(p)) ** 2)
        p = fminbound(e, 0, 2 * np.pi, xtol=1e-9, full_output=1)[0]
        frame *= np.exp(1j * p)
        if (-1 * frame).real.sum() > frame.real.sum():
            frame *= -1
    return data[0] if was2d else data


def align_global_phase_closed(data_in):
    """Closed form of the same minimisation.

    e(p) = [sin p, cos p] C [sin p, cos p]^T with C the Gram matrix of
    (Re, Im); the minimiser is the eigenvector of C's smaller eigenvalue.
    Then the same sign rule (``phasing.py:1396-1399``).
    """
    data = np.array(data_in, dtype=np.complex128, copy=True)
    was2d = data.ndim == 2
    if was2d:
        data = data[None]
    for frame in data:
        a = float(np.sum(frame.real ** 2))
        b = float(np.sum(frame.imag ** 2))
        c = float(np.sum(frame.real * frame.imag))
        # e(p) = (a+b)/2 + (b-a)/2 cos 2p + c sin 2p ; minimum at
        # 2p = atan2(c, (b-a)/2) + pi
        p = 0.5 * (math.atan2(c, 0.5 * (b - a)) + math.pi)
        frame *= np.exp(1j * p)
        if (-1 * frame).real.sum() > frame.real.sum():
            frame *= -1
    return data[0] if was2d else data


# ---------------------------------------------------------------------------
# prtf / rftf (phasing.py:1406-1461, 1521-1609)
# ---------------------------------------------------------------------------
def prtf(estimates, N=None):
    """PRTF, ``phasing.py:1406-1461`` (without the q-average).

    mean over frames of Psi/|Psi| of the zero-padded estimates, abs, fftshift.
    """
    est = np.asarray(estimates)
    f, r, c = est.shape
    if N is None:
        N = max(r, c)
    acc = np.zeros((N, N), np.complex128)
    for n in range(f):
        pad = np.zeros((N, N), np.complex128)
        pad[:r, :c] = est[n]
        F = np.fft.fft2(pad)
        acc += F / np.abs(F)
    return np.fft.fftshift(np.abs(acc / f))


def rftf(estimate, goal_modulus, scale=False, ipsf=None):
    """RFTF, ``phasing.py:1521-1586`` (no hot-pixel removal).

    Zero-pads the estimate into the modulus shape; with an ipsf |F| becomes
    sqrt(|ifft2(fft2(|F|) * ipsf)|) (``phasing.py:1567-1573``: the modulus, not
    the intensity, is blurred); compares it with the goal at both corner and
    centre alignments (``phasing.py:1564-1573``) and returns sqrt(1/(1+err)),
    centre-shifted.
    """
    est = np.asarray(estimate)
    g = np.asarray(goal_modulus, np.float64)
    new = np.zeros(g.shape, np.complex128)
    new[:est.shape[0], :est.shape[1]] = est
    fourier = np.abs(np.fft.fft2(new))
    if ipsf is not None:
        fourier = np.sqrt(np.abs(np.fft.ifft2(np.fft.fft2(fourier) * ipsf)))
    shift = np.fft.fftshift
    d1 = np.sum(np.abs(fourier - g))
    d2 = np.sum(np.abs(shift(fourier) - g))
    if d1 > d2:
        error = (shift(fourier) - g) ** 2 / g ** 2
    else:
        error = shift((fourier - g) ** 2 / g ** 2)
    out = np.sqrt(1.0 / (1 + error))
    if scale:
        x = 1.0 / np.sqrt(2)
        out = (out - x) / (1 - x)
    return out


def rftf_stack(estimates, goal_modulus, ipsf=None):
    """Stack RFTF of ``gpu_rftf``, ``phasing.py:1611-1708`` with
    ``phasing_rftf_accumulate.cl:15-24``: per pixel, the mean over frames of
    1/sqrt(1 + (g - |F_f|)^2 / g^2), goal at the corner, then fftshift.  With
    an ipsf |F_f| becomes sqrt(|ifft2(fft2(|F_f|^2) * Re ipsf)|): the GPU
    branch blurs the intensity and uploads the ipsf as float32, i.e. its real
    part (``phasing.py:1668-1690``).
    """
    est = np.asarray(estimates)
    g = np.asarray(goal_modulus, np.float64)
    acc = np.zeros(g.shape, np.float64)
    for e in est:
        new = np.zeros(g.shape, np.complex128)
        new[:e.shape[0], :e.shape[1]] = e
        fa = np.abs(np.fft.fft2(new))
        if ipsf is not None:
            fa = np.sqrt(np.abs(np.fft.ifft2(np.fft.fft2(fa ** 2) * np.real(ipsf))))
        d = g - fa
        acc += 1.0 / np.sqrt(1.0 + d * d / (g * g))
    return np.fft.fftshift(acc / len(est))


# ---------------------------------------------------------------------------
# optimize_coherence (phasing.py:507-930), CPU branch
# ---------------------------------------------------------------------------
def rl_start_gaussian(n):
    """Initial coherence estimate of the Richardson-Lucy method
    (``phasing.py:818``): fftshift(shape.gaussian((n, n), (n/4, n/4))),
    unnormalised, centre n/2 (Python 2 integer division of the shape)."""
    x = np.arange(n, dtype=float) - n // 2
    g1 = np.exp(-x ** 2 / (2.0 * (n // 4) ** 2))
    return np.fft.fftshift(np.outer(g1, g1))


def rl_coherence(estimate, modulus, iterations=500, semantics="numpy", is_fourier=False):
    """Richardson-Lucy coherence estimate of ONE frame (``phasing.py:702-837``).

    p = |fft2(estimate embedded top-left)|^2 / sum (or |estimate|^2 when
    is_fourier, the GPU branch), P = fft2(p), Ph = P[::-1, ::-1] (a plain
    reversal, ``common_flip_f2.cl``), d = m^2 / sum(m^2), u = rl_start_gaussian;
    per iteration u *= ifft2(fft2(d / ifft2(fft2(u) P)) Ph); then u = fft2(u),
    real part with numpy semantics (the CPU branch casts to float32,
    ``phasing.py:778-783``), divided by |u[0, 0]| (``phasing.py:641-643``).
    The reference returns fftshift of this and loads the unshifted array;
    only the last frame of a stack determines it (``rl_sum`` is never read).
    Returns the unshifted ipsf.
    """
    m = np.asarray(modulus, np.float64)
    n = m.shape[0]
    est = np.asarray(estimate)
    if is_fourier:
        act = np.zeros((n, n), np.complex128)
        act[:est.shape[0], :est.shape[1]] = est
        p = np.abs(act) ** 2
    else:
        act = np.zeros((n, n), np.complex128)
        act[:est.shape[0], :est.shape[1]] = est
        p = np.abs(np.fft.fft2(act)) ** 2
    p = p / p.sum()
    P = np.fft.fft2(p)
    Ph = P[::-1, ::-1]
    d = m ** 2
    d = d / d.sum()
    u = rl_start_gaussian(n).astype(np.complex128)
    for _ in range(iterations):
        blur = np.fft.ifft2(np.fft.fft2(u) * P)
        blur = d / blur
        blur = np.fft.ifft2(np.fft.fft2(blur) * Ph)
        u = u * blur
    u = np.fft.fft2(u)
    if semantics == "numpy":
        u = u.real
    return u / abs(u[0, 0])


def gaussian_coherence_objective(ac, m, cl):
    """``_opt_iter_g`` (``phasing.py:657-698``): sum (sqrt|fft2(g * ac)| - m)^2
    with g = exp(-rows^2/(2 cl0^2)) exp(-cols^2/(2 cl1^2)) on fftshifted
    signed coordinates."""
    n = m.shape[0]
    rows, cols = np.indices((n, n)) - n // 2
    rows, cols = np.fft.fftshift(rows), np.fft.fftshift(cols)
    g = np.exp(-rows ** 2 / (2.0 * cl[0] ** 2)) * np.exp(-cols ** 2 / (2.0 * cl[1] ** 2))
    bl = np.sqrt(np.abs(np.fft.fft2(g * ac)))
    return float(np.sum(np.abs(bl - m) ** 2))


def autocorrelation(estimate, n):
    """``_preprocess_g`` (``phasing.py:785-803``): ifft2(|fft2(active)|^2)."""
    est = np.asarray(estimate)
    act = np.zeros((n, n), np.complex128)
    act[:est.shape[0], :est.shape[1]] = est
    return np.fft.ifft2(np.abs(np.fft.fft2(act)) ** 2)


def gaussian_coherence(estimates, modulus):
    """``optimize_coherence(method='gaussian')``: per frame, scipy ``fmin``
    from (n/2, n/2) with ftol 0.1, xtol 0.1 (``phasing.py:897-906``); returns
    the (2, frames) coherence lengths."""
    from scipy.optimize import fmin
    m = np.asarray(modulus, np.float64)
    n = m.shape[0]
    est = np.asarray(estimates)
    if est.ndim == 2:
        est = est[None]
    out = np.zeros((2, est.shape[0]))
    for k, fr in enumerate(est):
        ac = autocorrelation(fr, n)
        out[:, k] = fmin(lambda c: gaussian_coherence_objective(ac, m, c), (n // 2, n // 2),
                         disp=False, ftol=.1, xtol=0.1)
    return out


def center_of_mass_average(imgs):
    """``phasing.py:1710-1745``: the sum over frames of each frame rolled by
    minus its truncated centre-of-mass offset from frame 0's (|frame|
    weights)."""
    imgs = np.asarray(imgs)
    if imgs.ndim == 2:
        imgs = imgs[None]
    rows, cols = np.indices(imgs.shape[1:])
    total = np.zeros(imgs.shape[1:], np.complex128)
    for n, img in enumerate(imgs):
        a = np.abs(img)
        f = np.sum(a)
        r, c = int(np.sum(rows * a) / f), int(np.sum(cols * a) / f)
        if n == 0:
            r0, c0 = r, c
        total += np.roll(np.roll(img, -(r - r0), axis=0), -(c - c0), axis=1)
    return total
