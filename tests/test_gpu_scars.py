"""GPU tests of the random scar D-map generator (deepx/utils_scars.py:64-426,
deepx/generate.py:79-83): the device rasterisation bit for bit against the
brute-force crossing-number oracle, the blur against scipy's convolve2d
"same", the rescale bit for bit against its FP64 restatement, and the whole
generator against the oracle pipeline on the same contours."""
import numpy as np
import pytest
import torch

from cardiax_amd import generate, scars
from oracle import scar_oracle as so

pytestmark = pytest.mark.gpu


def _star(rng, shape, n, rmin, rmax):
    th = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(rmin, rmax, n)
    cr, cc = rng.uniform(0, shape[0]), rng.uniform(0, shape[1])
    return cr + rad * np.sin(th), cc + rad * np.cos(th)


def test_raster_random_polygons_bitwise(gpu):
    rng = np.random.default_rng(0)
    for shape in [(37, 53), (64, 64), (130, 70), (1, 9), (200, 333)]:
        for n_polys in (1, 3, 6):
            cs = [_star(rng, shape, int(rng.integers(3, 60)), 1.0, 0.6 * max(shape))
                  for _ in range(n_polys)]
            print("raster %s, %d polygons" % (shape, n_polys), flush=True)
            got = scars.rasterize(cs, shape).cpu().numpy()
            assert np.array_equal(got, so.union_mask(cs, shape))
    # integer rectangles: the half-open edges of the rule
    got = scars.rasterize([([2.0, 2.0, 9.0, 9.0], [3.0, 11.0, 11.0, 3.0])], (20, 24)).cpu().numpy()
    want = np.zeros((20, 24), np.float32)
    want[2:9, 3:11] = 1
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(64, 64), (256, 256), (300, 200), (150, 400), (1200, 1200)])
def test_raster_generator_contours_bitwise(gpu, shape):
    for seed in range(3):
        cs = scars.blob_contours(seed, shape)
        print("generator contours %s seed %d" % (shape, seed), flush=True)
        got = scars.rasterize(cs, shape).cpu().numpy()
        ref = so.union_mask(cs, shape)
        assert np.array_equal(got, ref)
        # square images hold the whole scar; on others the reference maps y onto
        # shape[1] rows and x onto shape[0] columns, and what falls outside is dropped
        assert ref.any() or shape[0] != shape[1]


def test_raster_rejects_bad_polygons(gpu):
    with pytest.raises(ValueError, match="fewer than 3"):
        scars.rasterize([([1.0, 2.0], [1.0, 2.0])], (8, 8))
    with pytest.raises(ValueError, match="non-finite"):
        scars.rasterize([([1.0, np.nan, 3.0], [1.0, 2.0, 3.0])], (8, 8))


@pytest.mark.parametrize("H,W,kh,kw", [(40, 33, 3, 3), (40, 33, 4, 6), (17, 90, 25, 25),
                                       (64, 64, 1, 1), (9, 5, 12, 7), (256, 256, 25, 25),
                                       (300, 200, 30, 30)])
def test_blur_matches_scipy(gpu, H, W, kh, kw):
    rng = np.random.default_rng(H * 31 + kh)
    img = rng.uniform(size=(H, W)).astype(np.float32)
    k = rng.standard_normal((kh, kw))
    got = scars.convolve_same(torch.from_numpy(img).cuda(), k).cpu().numpy()
    ref = so.convolve_same(img, k)
    # double accumulation, one rounding to float32
    np.testing.assert_allclose(got, ref, rtol=2 ** -23, atol=1e-12 * np.abs(k).sum())


def test_blur_full_size_gaussian(gpu):
    """The reference's 1200^2 default image with its 120 x 120 kernel."""
    mask = so.union_mask(scars.blob_contours(3, (1200, 1200)), (1200, 1200))
    k = so.gauss_kernel(120, 95.0)
    got = scars.convolve_same(torch.from_numpy(mask).cuda(), k).cpu().numpy()
    ref = so.convolve_same_fft(mask, k)
    assert np.abs(got - ref).max() <= 1e-6


def test_rescale_bitwise(gpu):
    rng = np.random.default_rng(9)
    for shape in [(64, 64), (1200, 1200), (3, 1000001)]:
        b = rng.uniform(size=shape).astype(np.float32)
        for dom in [(1e-4, 1e-3), (2e-4, 5e-4)]:
            got = scars.scar_to_diffusivity(torch.from_numpy(b).cuda(), dom).cpu().numpy()
            assert np.array_equal(got, so.scar_to_diffusivity(b, dom))


@pytest.mark.parametrize("shape", [(64, 64), (256, 256), (300, 200), (1200, 1200)])
def test_random_diffusivity_matches_oracle_pipeline(gpu, shape):
    for seed in range(2):
        print("random_diffusivity %s seed %d" % (shape, seed), flush=True)
        D = generate.random_diffusivity(seed, shape)
        assert D.is_cuda and D.dtype == torch.float32 and tuple(D.shape) == shape
        got = D.cpu().numpy()
        # the oracle pipeline on the generator's own draws
        g1, g2, size, sigma = scars._blur_draws(seed, shape)
        cs = scars.blob_contours(g2, shape, scars.DEF_PARAMS, scars.spline_centroids(g1))
        mask = so.union_mask(cs, shape)
        k = so.gauss_kernel(size, sigma)
        conv = so.convolve_same_fft if shape[0] * shape[1] > 300_000 else so.convolve_same
        blurred = conv(mask, k).astype(np.float32)
        ref = so.scar_to_diffusivity(blurred)
        assert got.min() == np.float32(1e-4) and got.max() == np.float32(1e-3)
        # blur values within 1e-6 move D by at most 1e-6 * 9e-4 / (max - min of the scar)
        span = float(blurred.max() - blurred.min())
        assert np.abs(got.astype(np.float64) - ref).max() <= 1e-6 * 9e-4 / span + 2e-10
        assert np.array_equal(got, generate.random_diffusivity(seed, shape).cpu().numpy())
        inv = scars.random_diffusivity_scar(seed, shape).cpu().numpy()
        np.testing.assert_allclose(inv, 1.0 - blurred, rtol=0, atol=1e-6)


def test_random_diffusivity_domain_and_mask(gpu):
    D = generate.random_diffusivity(5, (128, 128), domain=(2e-4, 5e-4)).cpu().numpy()
    assert D.min() == np.float32(2e-4) and D.max() == np.float32(5e-4)
    m = scars.random_spline(11, (96, 96)).cpu().numpy()
    assert set(np.unique(m)) == {0.0, 1.0}


def test_scar_outside_the_image_becomes_the_centre_pixel(gpu):
    """On tall images the reference's row mapping (rows = shape[1] - y) can put
    the whole scar outside; the mask of one value then becomes one pixel at the
    centre, and the D-map is the blur of that pixel, rescaled."""
    shape = (150, 400)
    seed = 12
    g1, g2, size, sigma = scars._blur_draws(seed, shape)
    cs = scars.blob_contours(g2, shape, scars.DEF_PARAMS, scars.spline_centroids(g1))
    assert not so.union_mask(cs, shape).any()  # the whole scar falls outside
    h1, h2, _, _ = scars._blur_draws(seed, shape)  # the same draws again
    m = scars.random_spline(h2, shape, scars.DEF_PARAMS, scars.spline_centroids(h1))
    m = m.cpu().numpy()
    want = np.zeros(shape, np.float32)
    want[75, 200] = 1.0
    assert np.array_equal(m, want)
    D = generate.random_diffusivity(seed, shape).cpu().numpy()
    blurred = so.convolve_same(want, so.gauss_kernel(size, sigma)).astype(np.float32)
    ref = so.scar_to_diffusivity(blurred)
    assert D.min() == np.float32(1e-4) and D.max() == np.float32(1e-3)
    np.testing.assert_allclose(D, ref, rtol=0, atol=1e-9)
