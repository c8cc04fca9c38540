"""CPU restatement of the rotation that cardiax's ``triangular`` stimulus
applies to a linear stripe.

TEST INFRASTRUCTURE ONLY. Only ``tests/`` may import this module, to check
results; the product path (``cardiax_amd``) never does.

Reference behaviour followed (file:line under the reference tree):
  * ``rotate``  cardiax/stimulus.py:137-139: ``scipy.ndimage.rotate(field,
    angle=angle, mode="nearest", prefilter=False, reshape=False)`` with the
    default ``order=3`` and ``axes=(1, 0)``. scipy is a third-party
    dependency (unpinned by the reference). Its ``rotate`` builds
    M = [[cos a, sin a], [-sin a, cos a]] (``special.cosdg``/``sindg``,
    exact at multiples of 90 degrees) and offset = c - M c with
    c = ((H-1)/2, (W-1)/2), then ``affine_transform``: output o samples the
    input at x = M o + offset through the cubic B-spline of the unfiltered
    samples, taps floor(x)-1 .. floor(x)+2 per axis with the B-spline weights
    of t = x - floor(x) (the last one computed as 1 minus the others), and
    ``mode="nearest"`` takes every tap outside the grid from the nearest edge
    sample (the coordinate itself is not clamped). Restated below in FP64
    with scipy's operation order (the map is built with scipy's own
    ``cosdg``/``sindg`` and NumPy's ``@``, as ``rotate`` builds it); pinned
    bit for bit against scipy itself, the library the reference calls, and by
    a closed form, in ``tests/test_oracle_stimulus_rotate.py``.
"""
from __future__ import annotations

import numpy as np


def rotation_map(shape, angle: float):
    """The affine map scipy.ndimage.rotate (axes (1, 0), reshape=False)
    passes to affine_transform, built with the same functions."""
    from scipy import special

    c, s = special.cosdg(angle), special.sindg(angle)
    M = np.array([[c, s], [-s, c]])
    centre = (np.asarray(shape) - 1) / 2
    return M, centre - M @ centre


def bspline3(t):
    """Weights of taps floor(x)-1 .. floor(x)+2 at fractional part t, in
    scipy's arithmetic: the last weight is 1 minus the other three (so at
    t = 0 it is 8.3e-17, not 0, and a cell one row past a sharp edge reads
    a tiny nonzero value)."""
    z = 1.0 - t
    w0 = z * z * z / 6.0
    w1 = (t * t * (t - 2.0) * 3.0 + 4.0) / 6.0
    w2 = (z * z * (z - 2.0) * 3.0 + 4.0) / 6.0
    w3 = 1.0 - w0 - w1 - w2
    return np.stack([w0, w1, w2, w3])


def affine_spline3(a, M, offset) -> np.ndarray:
    """scipy.ndimage.affine_transform(a, M, offset, order=3, mode="nearest",
    prefilter=False) in FP64 with scipy's operation order: coordinates
    (0 + o0 M[k,0] + o1 M[k,1]) + offset[k]; row taps outer, column taps
    inner, each term (value * w_row) * w_col."""
    a = np.asarray(a, dtype=np.float64)
    H, W = a.shape
    o0, o1 = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64),
                         indexing="ij")
    x0 = (0.0 + o0 * M[0][0] + o1 * M[0][1]) + offset[0]
    x1 = (0.0 + o0 * M[1][0] + o1 * M[1][1]) + offset[1]
    f0, f1 = np.floor(x0), np.floor(x1)
    w0, w1 = bspline3(x0 - f0), bspline3(x1 - f1)
    out = np.zeros((H, W))
    for i in range(4):
        r = np.clip(f0.astype(np.int64) - 1 + i, 0, H - 1)
        for j in range(4):
            q = np.clip(f1.astype(np.int64) - 1 + j, 0, W - 1)
            out += (a[r, q] * w0[i]) * w1[j]
    return out


def rotate(a, angle: float) -> np.ndarray:
    """scipy.ndimage.rotate(a, angle, mode="nearest", prefilter=False,
    reshape=False), order 3, in FP64."""
    M, off = rotation_map(np.shape(a), float(angle))
    return affine_spline3(a, M, off)
