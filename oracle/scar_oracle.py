"""CPU restatement of the per-pixel half of the reference's random scar
generator: rasterisation, Gaussian kernel, blur, inversion and rescale.

TEST INFRASTRUCTURE ONLY. Only ``tests/`` may import this module, to check
results; the product path (``cardiax_amd``) never does.

Reference behaviour followed (file:line under the reference tree):
  * ``polygon_mask``  deepx/utils_scars.py:79-97 (drawpolyintoemptycanvas)
    through ``skimage.draw.polygon`` (scikit-image, unpinned by the reference
    and absent here; its published algorithm is restated): every pixel (r, c)
    of the box [int(max(0, min R)), ceil(max R)] x [int(max(0, min C)),
    ceil(max C)] is tested by the crossing-number rule at (c, r): edge (i, j),
    j = i - 1 cyclically, toggles when (R_i <= r < R_j or R_j <= r < R_i) and
    c < (C_j - C_i)(r - R_i)/(R_j - R_i) + C_i. Pixels beyond the image are
    dropped (JAX's scatter drops out-of-bounds updates), and the blobs are
    OR'ed (summed and clipped to 1, :377-381). Brute force over the box, in
    FP64 with the rule's own operation order. Pinned in
    ``tests/test_scars_cpu.py`` against closed forms (the rule's half-open
    edges on integer rectangles), ``scipy.spatial.Delaunay`` on convex
    polygons and ``matplotlib.path.Path`` on random star polygons.
  * ``gauss_kernel``  deepx/utils_scars.py:64-76 (makegauss2D): exp of
    -(x^2 + y^2)/(2 sigma^2) on the centred grid, entries below
    eps(float32) * max zeroed, normalised. Pinned against the outer product
    of ``scipy.signal.windows.gaussian`` and by its threshold.
  * ``convolve_same``  deepx/utils_scars.py:402-406 (blur_scar): the
    reference calls ``jax.scipy.signal.convolve2d(..., "same")``, whose
    padding (k - 1 - (k - 1)//2, (k - 1)//2) gives scipy's "same" alignment;
    here ``scipy.signal.convolve2d`` in FP64 itself.
  * ``scar_to_diffusivity``  deepx/utils_scars.py:426, deepx/generate.py:
    79-83, cardiax/convert.py:23-26: c = 1 - scar, then
    (c - min c) (hi - lo) / (max c - min c) + lo in FP64, in that order.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import signal


def polygon_mask(rows, cols, shape) -> np.ndarray:
    """skimage.draw.polygon(rows, cols) drawn into a zero (H, W) image,
    out-of-image pixels dropped. bool."""
    H, W = shape
    r = np.asarray(rows, dtype=np.float64)
    c = np.asarray(cols, dtype=np.float64)
    out = np.zeros((H, W), dtype=bool)
    r0, r1 = int(max(0.0, r.min())), min(H - 1, int(math.ceil(r.max())))
    c0, c1 = int(max(0.0, c.min())), min(W - 1, int(math.ceil(c.max())))
    if r1 < r0 or c1 < c0:
        return out
    y = np.arange(r0, r1 + 1, dtype=np.float64)[:, None]
    x = np.arange(c0, c1 + 1, dtype=np.float64)[None, :]
    inside = np.zeros((r1 - r0 + 1, c1 - c0 + 1), dtype=bool)
    n = len(r)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(n):
            j = i - 1 if i > 0 else n - 1
            yi, yj, xi, xj = r[i], r[j], c[i], c[j]
            cross = ((yi <= y) & (y < yj)) | ((yj <= y) & (y < yi))
            if not cross.any():
                continue
            xe = (xj - xi) * (y - yi) / (yj - yi) + xi
            inside ^= cross & (x < xe)
    out[r0:r1 + 1, c0:c1 + 1] = inside
    return out


def union_mask(contours, shape) -> np.ndarray:
    """The blobs OR'ed as float32 0/1 (MakeAndSumCompositeBlob's sum clipped
    to 1)."""
    m = np.zeros(shape, dtype=bool)
    for rows, cols in contours:
        m |= polygon_mask(rows, cols, shape)
    return m.astype(np.float32)


def gauss_kernel(size: int, sigma: float) -> np.ndarray:
    m = (size - 1) / 2.0
    g = np.arange(size, dtype=np.float64) - m
    d2 = g[:, None] ** 2 + g[None, :] ** 2
    h = np.exp(-d2 / (2.0 * sigma * sigma))
    h = np.where(h < np.finfo(np.float32).eps * h.max(), 0.0, h)
    s = h.sum()
    return h / s if s != 0 else h


def convolve_same(img, kernel) -> np.ndarray:
    """scipy.signal.convolve2d(img, kernel, "same") in FP64."""
    return signal.convolve2d(np.asarray(img, np.float64), np.asarray(kernel, np.float64),
                             mode="same")


def convolve_same_fft(img, kernel) -> np.ndarray:
    """The same "same" convolution by FFT (scipy.signal.fftconvolve, whose
    "same" crop starts at (k - 1)//2 as convolve2d's does), for images where
    the direct sum would take minutes on the CPU; agrees with
    ``convolve_same`` to FP64 rounding."""
    return signal.fftconvolve(np.asarray(img, np.float64), np.asarray(kernel, np.float64),
                              mode="same")


def scar_to_diffusivity(blurred, domain=(0.0001, 0.001)) -> np.ndarray:
    """diffusivity_rescale(1 - blurred, domain) in FP64, rounded to float32."""
    c = 1.0 - np.asarray(blurred, np.float64)
    a, b = c.min(), c.max()
    lo, hi = float(domain[0]), float(domain[1])
    return ((c - a) * (hi - lo) / (b - a) + lo).astype(np.float32)
