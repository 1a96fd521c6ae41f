"""Host-side mask helpers used by the phasing API (the reference runs these on
the host too).

bounding_box      speckle/masking.py:7-159  (ROI of the support, load())
find_all_objects  speckle/masking.py:379-434 (parts of a support, refine_support)
"""
from __future__ import annotations

import math

import numpy as np


def bounding_box(data, threshold=1e-10, force_to_square=False, pad=0, power_of_2=False):
    """Minimal (rmin, rmax, cmin, cmax) of ``data > threshold``, padded.

    Same contract as ``speckle.masking.bounding_box`` (masking.py:7-159):
    pad and clip to the array (masking.py:147-151), optional power-of-two
    growth (masking.py:35-84) and squaring (masking.py:86-117), then shift a
    negative origin back to 0 keeping the extent (masking.py:161-165).
    """
    assert isinstance(data, np.ndarray) and data.ndim == 2, "data must be a 2d ndarray"
    assert isinstance(pad, (int, np.integer)), "pad must be an integer"
    assert isinstance(force_to_square, bool), "force_to_square must be a bool"
    assert isinstance(threshold, (int, float)), "threshold must be int or float"
    rows, cols = data.shape
    rr = np.flatnonzero((data > threshold).any(axis=1))
    cc = np.flatnonzero((data > threshold).any(axis=0))
    if rr.size == 0:
        raise ValueError("bounding_box: no pixel above threshold")
    lo_r, hi_r = int(rr[0]), int(rr[-1]) + 1
    lo_c, hi_c = int(cc[0]), int(cc[-1]) + 1
    lo_r, hi_r = max(lo_r - pad, 0), min(hi_r + pad, rows)
    lo_c, hi_c = max(lo_c - pad, 0), min(hi_c + pad, cols)

    if power_of_2:
        lo_r, hi_r = _grow_p2(lo_r, hi_r, rows)
        lo_c, hi_c = _grow_p2(lo_c, hi_c, cols)
    if force_to_square:
        lo_r, hi_r, lo_c, hi_c = _squarify(lo_r, hi_r, lo_c, hi_c, rows, cols)
    if lo_r < 0:
        lo_r, hi_r = 0, hi_r - lo_r
    if lo_c < 0:
        lo_c, hi_c = 0, hi_c - lo_c
    return np.array([lo_r, hi_r, lo_c, hi_c]).astype(int)


def _grow_p2(lo, hi, size):
    ext = hi - lo
    p2 = ext if ext & (ext - 1) == 0 else 1 << (int(math.floor(math.log2(ext))) + 1)
    if p2 >= size:
        return 0, size
    extra = p2 - ext
    return lo - extra // 2, hi + (extra - extra // 2)


def _squarify(lo_r, hi_r, lo_c, hi_c, rows, cols):
    er, ec = hi_r - lo_r, hi_c - lo_c
    if er % 2:
        er += 1
        if hi_r < rows:
            hi_r += 1
        else:
            lo_r -= 1
    if ec % 2:
        ec += 1
        if hi_c < cols:
            hi_c += 1
        else:
            lo_c -= 1
    if er > ec:
        mid = (hi_c + lo_c) // 2
        lo_c, hi_c = mid - er // 2, mid + er // 2
    elif ec > er:
        mid = (hi_r + lo_r) // 2
        lo_r, hi_r = mid - ec // 2, mid + ec // 2
    return lo_r, hi_r, lo_c, hi_c


def find_all_objects(data, detect_val=(0.95, 1.05)):
    """Connected parts of ``data`` as an int array (parts, rows, cols).

    Same contract as ``speckle.masking.find_all_objects`` with
    ``trace_object`` (masking.py:271-434): pixels with detect_val[0] <= v <=
    detect_val[1] are 8-connected, indices wrap periodically
    (masking.py:356, 370-372), parts are ordered by their first pixel in
    raster order, and a pixel with no in-range neighbour yields an empty part
    (the flood fill only marks pixels reached from a neighbour,
    masking.py:362-373).  Unlike the reference the input is not modified
    (masking.py:428 zeroes it).
    """
    from scipy import ndimage

    data = np.asarray(data)
    assert data.ndim == 2, "for now data must be 2d"
    work = (data >= detect_val[0]) & (data <= detect_val[1])
    rows, cols = work.shape
    labels, nlab = ndimage.label(work, structure=np.ones((3, 3), int))
    if nlab == 0:
        return np.zeros((0, rows, cols), int)
    # periodic wrap: union labels that touch across the opposite edges
    parent = np.arange(nlab + 1)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    def union_edges(e1, e2):
        # e1[i] touches e2[i-1], e2[i], e2[i+1] (8-connectivity, wrapping)
        n = e1.size
        for d in (-1, 0, 1):
            a = e1
            b = np.roll(e2, -d)
            m = (a > 0) & (b > 0)
            for x, y in zip(a[m], b[m]):
                rx, ry = find(x), find(y)
                if rx != ry:
                    parent[max(rx, ry)] = min(rx, ry)
        return n

    union_edges(labels[0, :], labels[-1, :])
    union_edges(labels[:, 0], labels[:, -1])
    # corners wrap diagonally as well (covered by the rolls above)
    roots = np.array([find(i) for i in range(nlab + 1)])
    lab = roots[labels]
    lab[~work] = 0
    # order parts by the raster index of their first pixel
    flat = lab.ravel()
    idx = np.flatnonzero(flat)
    first = {}
    for i in idx:
        v = flat[i]
        if v not in first:
            first[v] = i
    order = sorted(first, key=first.get)
    parts = []
    for v in order:
        part = (lab == v)
        if part.sum() == 1:
            part = np.zeros_like(part)  # isolated pixel: the reference's empty part
        parts.append(part.astype(int))
    return np.array(parts).astype(int)
