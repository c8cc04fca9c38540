"""Mirror of deepx/utils_scars.py: the random scar behind the dataset
generator's diffusivity map (``random_diffusivity``, deepx/generate.py:79-83).

A scar is the union of 3 to 6 blobs. Each blob is a closed cubic spline
through 11 radially perturbed points (makePolyAndSplineCurve, :100-125); the
blob centres are consecutive vertices of a perturbed 12-gon
(CreateSplineCentroids, :228-243); the union is placed at a random centre of
an 8 x 8 field scaled onto the image (MakeAndSumCompositeBlob, :246-390),
rasterised (drawpolyintoemptycanvas, :79-97), blurred by a truncated Gaussian
(blur_scar and random_diffusivity_scar, :402-426) and inverted, and
``random_diffusivity`` rescales 1 - scar onto the D range
(cardiax/convert.py:23-26).

The contours (a few hundred points) are built on the host with
scipy.interpolate, as the reference builds them. The per-pixel work runs on
the GPU: the rasterisation (``fk_rasterize_polygons``: skimage.draw.polygon's
crossing-number rule at pixel centres, the blobs OR'ed), the blur
(``fk_convolve_same``: convolve2d "same", in double) and the inversion and
rescale (``fk_scar_to_diffusivity``).

Random draws come from seeded NumPy generators in place of ``jax.random``
keys (absent here), split in the reference's pattern, with its distributions:
standard-normal radius perturbations, the blob count from [3, 7), the first
centre vertex from [0, 12), a uniform global centre, a standard-normal blur
width. The geometry runs in float64 (the reference's JAX arrays are float32).
The reference's coordinates are kept on non-square images too: y maps onto
shape[1] rows and x onto shape[0] columns of the (shape[0], shape[1]) image,
so part or all of a scar can fall outside it and is dropped.
Where the reference's code cannot run, this mirror does what it means: a mask
of one value becomes a single pixel at the centre (the reference assigns into
an immutable JAX array there), and a kernel of size 0 (images under 10 rows)
raises ValueError as JAX's convolve2d does.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Sequence, Tuple

import numpy as np
import torch
from scipy import interpolate

from . import _lib
from ._device import require_gpu, stream_handle
from ._keys import _randint, _split

# deepx/utils_scars.py:45-56 (the keys MakeAndSumCompositeBlob reads;
# RequiredImageSize is set per call from the image shape)
DEF_PARAMS = {
    "maxProtrudeFactorCentroid": 0.2,
    "r0Centroid": 1.5,
    "maxProtrudeFactor": 0.2,
    "r0": 1,
    "ScarScaleFactor": 0.4,
    "CONSERVATIVE_CENTROID": True,
}

Contour = Tuple[np.ndarray, np.ndarray]


def poly_and_spline_curve(rng, r0=30.0, max_protrude=0.3, n_gon=11, n_spline_pts=100):
    """deepx/utils_scars.py:100-125. A closed ``n_gon``-point polygon of
    radius r0 + max_protrude r0 (N(0,1) + 0.5) 0.36 at angles 2 pi k/(n_gon-1)
    (the last point repeats the first), and the interpolating cubic spline
    through it (splprep, s=0) sampled at ``n_spline_pts`` + 1 parameters in
    [0, 1]. Returns ((Px, Py), [Sx, Sy])."""
    g1, _ = _split(rng, 2)
    t = np.arange(n_gon, dtype=np.float64) / (n_gon - 1)
    rpert = (g1.standard_normal(n_gon) + 0.5) * 0.6 ** 2
    rdash = r0 + max_protrude * r0 * rpert
    px = rdash * np.sin(2 * np.pi * t)
    py = rdash * np.cos(2 * np.pi * t)
    px[-1], py[-1] = px[0], py[0]
    tck, _ = interpolate.splprep([px, py], s=0)
    u = np.arange(n_spline_pts + 1, dtype=np.float64) / n_spline_pts
    sx, sy = interpolate.splev(u, tck)
    return (px, py), [np.asarray(sx, np.float64), np.asarray(sy, np.float64)]


def spline_centroids(rng, params=DEF_PARAMS) -> Contour:
    """deepx/utils_scars.py:228-243: ``Npoints`` in [3, 7) consecutive
    vertices (cyclically, over the 12 stored points, the closing duplicate
    included) of a perturbed 12-gon of radius r0Centroid, from a start vertex
    in [0, 12)."""
    g1, g2, g3 = _split(rng, 3)
    (px, py), _ = poly_and_spline_curve(g1, r0=params["r0Centroid"],
                                        max_protrude=params["maxProtrudeFactorCentroid"],
                                        n_gon=12)
    n_points = _randint(g2, 3, 7)
    m = len(px)
    sel = (np.arange(n_points) + _randint(g3, 0, m)) % m
    return px[sel], py[sel]


def blob_contours(rng, shape, params=DEF_PARAMS, centroids=None) -> List[Contour]:
    """deepx/utils_scars.py:246-356 and :79-97 up to the rasterisation: one
    closed spline per centroid around it, the union placed at a uniform
    global centre of the field (conservatively inside it) and scaled onto
    the image; returned as polygons in pixel coordinates (rows, cols) with
    rows = shape[1] - y and cols = x, clipped to [0, shape[1]] and
    [0, shape[0]] as drawpolyintoemptycanvas clips them."""
    g1, g2, _, g4 = _split(rng, 4)
    if centroids is None:
        centroids = spline_centroids(g1, params)
    n_blobs = len(centroids[0])
    r0, mpf = params["r0"], params["maxProtrudeFactor"]
    max_ext = params["r0Centroid"] + params["maxProtrudeFactorCentroid"] + mpf + r0
    pts = []
    for i in range(n_blobs):
        g2, _ = _split(g2, 2)
        _, sp = poly_and_spline_curve(g2, r0=r0, max_protrude=mpf)
        pts.append([sp[0] + centroids[0][i], sp[1] + centroids[1][i]])
    max_field = math.ceil(max_ext / params["ScarScaleFactor"])
    ratio = shape[0] / shape[1]
    if ratio == 1:
        fx, fy = max_field, max_field
    elif ratio > 1:
        fx, fy = math.ceil(ratio) * max_field, max_field
    else:
        fx, fy = max_field, math.ceil(ratio) * max_field
    g41, g42 = _split(g4, 2)
    x, y = g41.uniform(), g42.uniform()
    if params["CONSERVATIVE_CENTROID"]:
        centre = ((fx - 2 * max_ext) * x + max_ext, (fy - 2 * max_ext) * y + max_ext)
    else:
        centre = ((fx - max_ext) * x + max_ext / 2, (fy - max_ext) * y + max_ext / 2)
    field_ratio = (fx / shape[0], fy / shape[1])
    out = []
    for p in pts:
        q = []
        for j in range(2):
            v = (p[j] + centre[j]) / field_ratio[j]
            q.append(v * (v > 0))
        rows = np.clip(shape[1] - q[1], 0, shape[1])
        cols = np.clip(q[0], 0, shape[0])
        out.append((rows, cols))
    return out


def gauss_kernel(size: int, sigma: float) -> np.ndarray:
    """deepx/utils_scars.py:64-76 (makegauss2D, MATLAB's fspecial): the
    (size, size) Gaussian of width ``sigma`` on the grid -(size-1)/2 ..
    (size-1)/2, entries below eps * max zeroed (eps of float32, the
    reference's dtype), normalised to sum 1. float64."""
    m = (size - 1) / 2.0
    a = np.arange(-m, m + 1, dtype=np.float64)
    yy, xx = np.meshgrid(a, a)
    h = np.exp(-(xx * xx + yy * yy) / (2.0 * sigma * sigma))
    h[h < np.finfo(np.float32).eps * h.max()] = 0.0
    s = h.sum()
    return h / s if s != 0 else h


def rasterize(contours: Sequence[Contour], shape, out=None) -> torch.Tensor:
    """The blobs' union as a float32 device mask (1 inside any contour):
    drawpolyintoemptycanvas per blob, summed and clipped to 1
    (deepx/utils_scars.py:79-97,377-381), by ``fk_rasterize_polygons``."""
    dev = require_gpu("scars.rasterize")
    H, W = int(shape[0]), int(shape[1])
    rows = np.ascontiguousarray(np.concatenate([np.asarray(c[0], np.float64) for c in contours]))
    cols = np.ascontiguousarray(np.concatenate([np.asarray(c[1], np.float64) for c in contours]))
    starts = np.zeros(len(contours) + 1, dtype=np.int32)
    starts[1:] = np.cumsum([len(c[0]) for c in contours])
    if out is None:
        out = torch.empty((H, W), dtype=torch.float32, device=dev)
    L = _lib.lib()
    nbytes = int(L.fk_rasterize_workspace_bytes(len(contours), int(starts[-1])))
    work = torch.empty(max(1, nbytes), dtype=torch.uint8, device=dev)
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.check(L.fk_rasterize_polygons(
        _lib.ptr(out), H, W, rows.ctypes.data_as(dp), cols.ctypes.data_as(dp),
        starts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(contours), _lib.ptr(work),
        nbytes, stream_handle()))
    return out


def convolve_same(img: torch.Tensor, kernel: np.ndarray) -> torch.Tensor:
    """convolve2d(img, kernel, "same") (jax.scipy.signal / scipy.signal) of a
    float32 device image, by ``fk_convolve_same`` (double accumulation)."""
    if img.dim() != 2 or img.dtype != torch.float32 or not img.is_cuda:
        raise ValueError("convolve_same: img must be a 2-D float32 device tensor")
    k = np.ascontiguousarray(np.asarray(kernel, dtype=np.float64))
    if k.ndim != 2 or k.size == 0:
        raise ValueError("convolve_same: the kernel must be a non-empty 2-D array")
    src = img.contiguous()
    kd = torch.from_numpy(k).to(src.device)
    dst = torch.empty_like(src)
    L = _lib.lib()
    nbytes = int(L.fk_convolve_workspace_bytes(src.shape[0], src.shape[1], k.shape[0],
                                               k.shape[1]))
    work = torch.empty(max(1, nbytes), dtype=torch.uint8, device=src.device)
    _lib.check(L.fk_convolve_same(_lib.ptr(src), src.shape[0], src.shape[1], _lib.ptr(kd),
                                  k.shape[0], k.shape[1], _lib.ptr(dst), _lib.ptr(work), nbytes,
                                  stream_handle()))
    return dst


def random_spline(rng, shape, params=DEF_PARAMS, centroids=None) -> torch.Tensor:
    """deepx/utils_scars.py:246-399 (MakeAndSumCompositeBlob's
    CompositeSplineMask, random_spline): the blobs' union as a 0/1 float32
    device mask; a mask of one value is replaced by one pixel at the centre."""
    shape = (int(shape[0]), int(shape[1]))
    mask = rasterize(blob_contours(rng, shape, params, centroids), shape)
    if not bool((mask != mask[0, 0]).any()):
        mask.zero_()
        mask[shape[0] // 2, shape[1] // 2] = 1.0
    return mask


def _blur_draws(rng, shape):
    """The key split of random_diffusivity_scar (:409-424): centroids from
    the first key, the blobs from the second, the blur width from the
    fourth; the kernel spans int(0.1 shape[0]) cells."""
    g1, g2, _, g4 = _split(rng, 4)
    size = int(shape[0] * 0.1)
    if size < 1:
        raise ValueError("the blur kernel of a %d-row image is empty (int(0.1 * %d) = 0)"
                         % (shape[0], shape[0]))
    sigma = float(g4.standard_normal()) * shape[0] / 10
    return g1, g2, size, sigma


def blurred_scar(rng, shape) -> torch.Tensor:
    """The blurred scar of random_diffusivity_scar before its inversion
    (:409-425), a float32 device tensor."""
    shape = (int(shape[0]), int(shape[1]))
    g1, g2, size, sigma = _blur_draws(rng, shape)
    params = DEF_PARAMS
    mask = random_spline(g2, shape, params, spline_centroids(g1, params))
    return convolve_same(mask, gauss_kernel(size, sigma))


def random_diffusivity_scar(rng, shape) -> torch.Tensor:
    """deepx/utils_scars.py:409-426: 1 - the blurred scar (float32, device)."""
    return 1.0 - blurred_scar(rng, shape)


def scar_to_diffusivity(blurred: torch.Tensor, domain=(0.0001, 0.001)) -> torch.Tensor:
    """diffusivity_rescale(1 - blurred, domain) (cardiax/convert.py:23-26) by
    ``fk_scar_to_diffusivity``: the map's minimum becomes domain[0] and its
    maximum domain[1]."""
    if blurred.dtype != torch.float32 or not blurred.is_cuda:
        raise ValueError("scar_to_diffusivity: a float32 device tensor is required")
    src = blurred.contiguous()
    D = torch.empty_like(src)
    work = torch.empty(4096, dtype=torch.uint8, device=src.device)
    _lib.check(_lib.lib().fk_scar_to_diffusivity(_lib.ptr(src), src.numel(), float(domain[0]),
                                                 float(domain[1]), _lib.ptr(D), _lib.ptr(work),
                                                 4096, stream_handle()))
    return D


def random_diffusivity(rng, shape, domain=(0.0001, 0.001)) -> torch.Tensor:
    """deepx/generate.py:79-83: random_diffusivity_scar rescaled onto
    ``domain`` (float32, on the current GPU)."""
    return scar_to_diffusivity(blurred_scar(rng, shape), domain)
