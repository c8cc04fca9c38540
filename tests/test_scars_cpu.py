"""CPU tests of the random scar generator (deepx/utils_scars.py:64-426): the
oracle's rasterisation and kernel pinned against closed forms and
independent libraries, and the host geometry (seeded draws in place of
jax.random keys) against the reference's construction."""
import math

import numpy as np
import pytest
from matplotlib.path import Path
from scipy import signal
from scipy.spatial import ConvexHull, Delaunay

from cardiax_amd import scars
from oracle import scar_oracle as so


def _pixels(shape):
    r, c = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    return r.ravel().astype(np.float64), c.ravel().astype(np.float64)


def test_polygon_rule_half_open_on_integer_rectangles():
    """skimage's crossing-number rule: an axis-aligned rectangle with integer
    corners covers rows [r0, r1) and columns [c0, c1), either orientation."""
    shape = (20, 24)
    for (r0, r1, c0, c1) in [(2, 9, 3, 11), (0, 20, 0, 24), (5, 6, 7, 8), (3, 19, 0, 5)]:
        want = np.zeros(shape, bool)
        want[r0:r1, c0:c1] = True
        rows = np.array([r0, r0, r1, r1], float)
        cols = np.array([c0, c1, c1, c0], float)
        assert np.array_equal(so.polygon_mask(rows, cols, shape), want)
        assert np.array_equal(so.polygon_mask(rows[::-1], cols[::-1], shape), want)
    # non-integer corners: the open interior
    m = so.polygon_mask([2.5, 2.5, 7.2, 7.2], [1.3, 6.6, 6.6, 1.3], shape)
    want = np.zeros(shape, bool)
    want[3:8, 2:7] = True
    assert np.array_equal(m, want)


def test_polygon_matches_delaunay_on_convex_polygons():
    rng = np.random.default_rng(1)
    shape = (41, 37)
    pr, pc = _pixels(shape)
    for _ in range(25):
        pts = rng.uniform([-5, -5], [shape[0] + 5, shape[1] + 5], size=(12, 2))
        hull = pts[ConvexHull(pts).vertices]
        inside = Delaunay(hull).find_simplex(np.stack([pr, pc], 1)) >= 0
        got = so.polygon_mask(hull[:, 0], hull[:, 1], shape).ravel()
        assert np.array_equal(got, inside)


def test_polygon_matches_matplotlib_on_star_polygons():
    rng = np.random.default_rng(2)
    shape = (53, 61)
    pr, pc = _pixels(shape)
    for _ in range(25):
        n = int(rng.integers(5, 40))
        th = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(3, 30, n)
        cr, cc = rng.uniform(0, shape[0]), rng.uniform(0, shape[1])
        rows, cols = cr + rad * np.sin(th), cc + rad * np.cos(th)
        path = Path(np.stack([cols, rows], 1), closed=False)
        inside = path.contains_points(np.stack([pc, pr], 1), radius=0.0)
        got = so.polygon_mask(rows, cols, shape).ravel()
        assert np.array_equal(got, inside)


def test_polygon_pixels_beyond_the_image_are_dropped():
    rng = np.random.default_rng(3)
    for _ in range(10):
        th = np.sort(rng.uniform(0, 2 * np.pi, 30))
        rad = rng.uniform(10, 40, 30)
        rows, cols = 30 + rad * np.sin(th), 25 + rad * np.cos(th)
        big = so.polygon_mask(rows, cols, (120, 120))
        assert np.array_equal(so.polygon_mask(rows, cols, (40, 33)), big[:40, :33])


def test_gauss_kernel_against_scipy_window():
    for size, sigma in [(25, 9.0), (24, 30.0), (7, 5.0), (102, 80.0), (1, 3.0)]:
        g = signal.windows.gaussian(size, std=sigma)
        ref = np.outer(g, g)
        ref /= ref.sum()
        np.testing.assert_allclose(so.gauss_kernel(size, sigma), ref, rtol=1e-13, atol=0)
        np.testing.assert_allclose(so.gauss_kernel(size, -sigma), ref, rtol=1e-13, atol=0)
        np.testing.assert_allclose(scars.gauss_kernel(size, sigma), so.gauss_kernel(size, sigma),
                                   rtol=1e-15, atol=0)


def test_gauss_kernel_threshold_and_even_sizes():
    # sigma 2 on 25 x 25: exp(-d^2/8) < 2^-23 beyond d^2 > 8 * 23 ln 2 = 127.5
    h = so.gauss_kernel(25, 2.0)
    g = np.arange(25) - 12.0
    d2 = g[:, None] ** 2 + g[None, :] ** 2
    assert (h[d2 > 128] == 0).all() and (h[d2 < 127] > 0).all()
    assert h.sum() == pytest.approx(1.0, abs=1e-15)
    w = signal.windows.gaussian(25, std=2.0)
    ref = np.outer(w, w)
    ref[ref < np.finfo(np.float32).eps * ref.max()] = 0
    np.testing.assert_allclose(h, ref / ref.sum(), rtol=1e-13, atol=0)
    # even size: centred between pixels, symmetric under every flip
    e = so.gauss_kernel(12, 2.5)
    assert np.array_equal(e, e[::-1, :]) and np.array_equal(e, e.T)
    np.testing.assert_allclose(scars.gauss_kernel(12, 2.5), e, rtol=1e-15, atol=0)


def test_convolve_same_alignment():
    """"same" keeps out[i, j] = sum k[p, q] in[i + (kh-1)//2 - p, j + (kw-1)//2 - q]:
    a unit impulse reproduces the kernel shifted by (kh-1)//2, (kw-1)//2."""
    rng = np.random.default_rng(4)
    for kh, kw in [(3, 3), (4, 6), (5, 2), (1, 7)]:
        k = rng.standard_normal((kh, kw))
        img = np.zeros((15, 17))
        img[7, 8] = 1.0
        out = so.convolve_same(img, k)
        want = np.zeros_like(img)
        for p in range(kh):
            for q in range(kw):
                want[7 - (kh - 1) // 2 + p, 8 - (kw - 1) // 2 + q] = k[p, q]
        np.testing.assert_array_equal(out, want)
        big = rng.uniform(size=(40, 33))
        np.testing.assert_allclose(so.convolve_same_fft(big, k), so.convolve_same(big, k),
                                   rtol=0, atol=1e-12)


def test_scar_to_diffusivity_endpoints():
    rng = np.random.default_rng(5)
    b = rng.uniform(size=(30, 20)).astype(np.float32)
    D = so.scar_to_diffusivity(b, (1e-4, 1e-3))
    assert D.min() == np.float32(1e-4) and D.max() == np.float32(1e-3)
    # the scar's maximum becomes the D minimum (1 - scar)
    assert D.flat[np.argmax(b)] == np.float32(1e-4) and D.flat[np.argmin(b)] == np.float32(1e-3)


def test_spline_curve_construction():
    """makePolyAndSplineCurve: vertex k at angle 2 pi k/(n-1) from +y towards
    +x (Px = r sin, Py = r cos), the last vertex repeating the first, radius
    r0 + max_protrude r0 (N + 0.5) 0.36, the spline through the end points
    sampled at 101 parameters."""
    radii = []
    for seed in range(400):
        (px, py), (sx, sy) = scars.poly_and_spline_curve(seed, r0=1.0, max_protrude=0.2)
        assert len(px) == 11 and len(sx) == 101
        assert px[-1] == px[0] and py[-1] == py[0]
        r = np.hypot(px, py)
        ang = np.mod(np.arctan2(px, py), 2 * np.pi)
        np.testing.assert_allclose(ang[:-1], 2 * np.pi * np.arange(10) / 10, atol=1e-12)
        assert sx[0] == pytest.approx(px[0], abs=1e-12) and sy[0] == pytest.approx(py[0], abs=1e-12)
        assert sx[-1] == pytest.approx(px[-1], abs=1e-12)
        radii.append(r[:-1])
    rad = np.concatenate(radii)
    # E[r] = 1 + 0.2 * 0.5 * 0.36, sd = 0.2 * 0.36
    assert rad.mean() == pytest.approx(1.036, abs=3 * 0.072 / math.sqrt(rad.size))
    assert rad.std() == pytest.approx(0.072, rel=0.05)


def test_centroids_are_consecutive_vertices():
    counts, starts = set(), set()
    for seed in range(300):
        cx, cy = scars.spline_centroids(seed)
        assert 3 <= len(cx) <= 6
        counts.add(len(cx))
        g1, _, g3 = scars._split(seed, 3)
        (px, py), _ = scars.poly_and_spline_curve(g1, r0=1.5, max_protrude=0.2, n_gon=12)
        # a run of consecutive stored vertices, modulo 12 (vertex 11 repeats vertex 0)
        ok = [s for s in range(12)
              if all(px[(s + k) % 12] == a and py[(s + k) % 12] == b
                     for k, (a, b) in enumerate(zip(cx, cy)))]
        assert ok
        starts.update(ok)
    assert counts == {3, 4, 5, 6} and starts == set(range(12))


@pytest.mark.parametrize("shape", [(64, 64), (256, 256), (300, 200), (150, 400)])
def test_blob_contours_placement(shape):
    """Blobs of 101-point contours inside the image, the conservative
    centre keeping all but strongly perturbed ones off the edges, and the
    same draws for the same seed."""
    inner = 0
    for seed in range(60):
        cs = scars.blob_contours(seed, shape)
        assert 3 <= len(cs) <= 6
        for rows, cols in cs:
            assert len(rows) == 101
            assert rows.min() >= 0 and rows.max() <= shape[1]
            assert cols.min() >= 0 and cols.max() <= shape[0]
        rows = np.concatenate([c[0] for c in cs])
        cols = np.concatenate([c[1] for c in cs])
        inner += bool(cols.min() > 0 and cols.max() < shape[0] and rows.min() > 0
                      and rows.max() < shape[1])
        assert np.array_equal(cs[0][0], scars.blob_contours(seed, shape)[0][0])
    assert inner >= 50


def test_blob_contours_scale_with_the_image():
    """On square images the field (maxField = ceil(2.9 / 0.4) = 8 units) maps
    onto the image, so the same draws give contours scaled by the side: the
    ratio of sides is a power of two here, and the scaling exact."""
    for seed in range(20):
        a = scars.blob_contours(seed, (64, 64))
        b = scars.blob_contours(seed, (256, 256))
        for (ra, ca), (rb, cb) in zip(a, b):
            np.testing.assert_array_equal(cb, 4 * ca)
            np.testing.assert_array_equal(256 - rb, 4 * (64 - ra))
        # field units: the centre lies in [2.9, 8 - 2.9], blobs within 2.9 of it
        # unless strongly perturbed
        cols = np.concatenate([c[1] for c in a]) / 8.0
        assert 0 <= cols.min() and cols.max() <= 8.0


def test_blur_draws_and_small_images():
    with pytest.raises(ValueError, match="blur kernel"):
        scars._blur_draws(0, (9, 9))
    sig = [scars._blur_draws(s, (200, 200))[3] for s in range(2000)]
    assert scars._blur_draws(0, (200, 200))[2] == 20
    # sigma = N(0, 1) * shape[0] / 10
    assert np.mean(sig) == pytest.approx(0.0, abs=3 * 20 / math.sqrt(2000))
    assert np.std(sig) == pytest.approx(20.0, rel=0.06)


def test_random_diffusivity_needs_the_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="GPU"):
        scars.random_diffusivity(0, (64, 64))
