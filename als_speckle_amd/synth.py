"""Seeded synthetic diffraction data for the phasing workloads.

Each frame: an object (disk or random star polygon) with amplitude U[a0, 1]
and phase U[-pi, pi] on its support, centred in an N x N frame; modulus
m = sqrt(Poisson(lambda |FFT(obj)|^2)) with lambda set so the expected total
equals the photon budget, corner-centred (q = 0 at [0, 0]) as phasing.load
expects (phasing.py:46, 297-298), float32; the given support is the true
support dilated by a disk of radius 3 px.  Data are drawn from
numpy.random.Generator(PCG64(seed)).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Frame:
    modulus: np.ndarray      # float32 (N, N), corner-centred
    support: np.ndarray      # float32 (N, N) in {0, 1}, dilated
    obj: np.ndarray          # complex128 (N, N), the true object
    true_support: np.ndarray  # bool (N, N)


def disk_mask(n, radius, center=None):
    """Hard disk r < radius about (n/2, n/2) (shape.circle AA=False, shape.py:248-268)."""
    c = (n // 2, n // 2) if center is None else center
    yy, xx = np.indices((n, n))
    return np.hypot(yy - c[0], xx - c[1]) < radius


def polygon_mask(n, fill, rng, vertices=None):
    """Random star polygon centred in the frame with area ~ fill * n^2.

    Vertices at sorted random angles with radii U[0.6, 1] * R, R scaled so the
    polygon area matches; rasterised by the even-odd rule at pixel centres.
    """
    k = int(vertices or rng.integers(9, 17))
    ang = np.sort(rng.uniform(0, 2 * np.pi, k))
    rad = rng.uniform(0.6, 1.0, k)
    area = 0.5 * np.sum(rad * np.roll(rad, -1) * np.sin(np.roll(ang, -1) - ang + 2 * np.pi * (np.arange(k) == k - 1)))
    R = np.sqrt(fill * n * n / area)
    R = min(R, 0.45 * n)
    vx = n / 2 + R * rad * np.cos(ang)
    vy = n / 2 + R * rad * np.sin(ang)
    r0, r1 = int(max(0, np.floor(vy.min()))), int(min(n, np.ceil(vy.max()) + 1))
    c0, c1 = int(max(0, np.floor(vx.min()))), int(min(n, np.ceil(vx.max()) + 1))
    yy, xx = np.mgrid[r0:r1, c0:c1].astype(np.float64)
    inside = np.zeros(yy.shape, bool)
    for i in range(k):
        x1, y1 = vx[i], vy[i]
        x2, y2 = vx[(i + 1) % k], vy[(i + 1) % k]
        cond = (y1 > yy) != (y2 > yy)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (x2 - x1) * (yy - y1) / (y2 - y1) + x1
        inside ^= cond & (xx < xint)
    mask = np.zeros((n, n), bool)
    mask[r0:r1, c0:c1] = inside
    return mask


def dilate(mask, r=3):
    """Binary dilation by a disk of radius r (periodic)."""
    out = mask.copy()
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            if dy * dy + dx * dx <= r * r and (dy or dx):
                out |= np.roll(np.roll(mask, dy, 0), dx, 1)
    return out


def make_frame(n, seed, shape="polygon", fill=0.15, radius=40, photons=1e8,
               amp=(0.5, 1.0), dilate_px=3, noise=True) -> Frame:
    rng = np.random.Generator(np.random.PCG64(seed))
    if shape == "disk":
        ts = disk_mask(n, radius)
    else:
        ts = polygon_mask(n, fill, rng)
    a = rng.uniform(amp[0], amp[1], (n, n))
    ph = rng.uniform(-np.pi, np.pi, (n, n))
    obj = ts * a * np.exp(1j * ph)
    inten = np.abs(np.fft.fft2(obj)) ** 2
    if noise:
        lam = photons / inten.sum()
        counts = rng.poisson(lam * inten)
        m = np.sqrt(counts).astype(np.float32)
    else:
        m = np.sqrt(inten).astype(np.float32)
    sup = dilate(ts, dilate_px).astype(np.float32)
    return Frame(m, sup, obj, ts)


# The five workloads (sizes, photon budgets and schedules of each).
CONFIGS = {
    "cpu_disk_256": dict(n=256, shape="disk", radius=40, photons=1e6, amp=(0.5, 1.0),
                         order=(("hio", 200), ("er", 20)), iterations=220, beta=0.9),
    "single_1024": dict(n=1024, shape="polygon", fill=0.15, photons=1e8, amp=(0.5, 1.0),
                        order=(("hio", 90), ("er", 10)), iterations=1000, beta=0.8,
                        shrinkwrap=dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)),
    "trials_512": dict(n=512, shape="polygon", fill=0.10, photons=1e7, amp=(0.5, 1.0),
                       trials=64, iterations=500,
                       runs=(("raar", 0.87), ("dm", 0.7))),
    "serving_mix": dict(sizes={256: 0.4, 512: 0.3, 1024: 0.2, 2048: 0.1}, requests=2000,
                        fill=(0.10, 0.15), photons=1e7, iterations=500,
                        order=(("hio", 90), ("er", 10)),
                        shrinkwrap=dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)),
    "large_4096": dict(n=4096, shape="polygon", fill=0.08, photons=1e9, amp=(0.3, 1.0),
                       order=(("hio", 90), ("er", 10)), iterations=2000, beta=0.8,
                       shrinkwrap=dict(every=20, sigma0=3.0, sigma1=1.5, tau=0.2)),
}


def config_frame(name, seed=0) -> Frame:
    c = CONFIGS[name]
    return make_frame(c["n"], seed, shape=c.get("shape", "polygon"), fill=c.get("fill", 0.15),
                      radius=c.get("radius", 40), photons=c["photons"], amp=c.get("amp", (0.5, 1.0)))
