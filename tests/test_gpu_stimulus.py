"""GPU tests of the stimulus builders (cardiax/stimulus.py:31-140) on the
device: box fields bit-identical to NumPy slicing of the reference's slices,
and the triangular stimulus (fk_affine_transform) bit-identical to
scipy.ndimage.rotate, the routine the reference calls."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from cardiax_amd import stimulus as S
from oracle import stimulus_oracle as so

pytestmark = pytest.mark.gpu

P = S.Protocol(0, 2, 1e9)


def _check_rotation(got, ref):
    np.testing.assert_array_equal(got != 0, ref != 0)  # the cells a stimulus replaces
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("shape", [(64, 64), (50, 70), (33, 17), (1200, 1200), (1, 6)])
@pytest.mark.parametrize("angle", [0, 90, 270, 30, 45, -20, 135, 359.5, 1e-3])
def test_triangular_matches_scipy(gpu, shape, angle):
    for d in (S.Direction.NORTH, S.Direction.WEST, S.Direction.SOUTH):
        got = S.triangular(shape, d, angle, 0.2, 20.0, P).field
        assert got.is_cuda and got.dtype == torch.float32
        stripe = S._stripe_mask(shape, d, 0.2, 20.0)
        ref = ndimage.rotate(stripe, angle=angle, mode="nearest", prefilter=False,
                             reshape=False)
        _check_rotation(got.cpu().numpy(), ref)


def _affine(a, M, off):
    import ctypes

    from cardiax_amd import _lib
    from cardiax_amd._device import stream_handle

    src = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    dst = torch.empty_like(src)
    m = (ctypes.c_double * 4)(*[float(x) for x in np.ravel(M)])
    o = (ctypes.c_double * 2)(*[float(x) for x in off])
    _lib.check(_lib.lib().fk_affine_transform(src.data_ptr(), a.shape[0], a.shape[1], m, o,
                                              dst.data_ptr(), stream_handle()))
    return dst.cpu().numpy()


def test_rotate_random_fields(gpu):
    rng = np.random.default_rng(1)
    for shape in [(40, 57), (128, 96), (1, 5)]:
        a = rng.standard_normal(shape).astype(np.float32)
        for angle in (17.0, -123.4, 180.0, 45.0):
            ref = ndimage.rotate(a, angle=angle, mode="nearest", prefilter=False,
                                 reshape=False)
            M, off = so.rotation_map(shape, angle)
            got = _affine(a, M, off)
            _check_rotation(got, ref)
            assert np.array_equal(got, so.rotate(a, angle).astype(np.float32))


def test_affine_transform_general_maps(gpu):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((30, 41)).astype(np.float32)
    for M, off in [(np.eye(2), (0.3, -2.6)), (np.array([[0.5, 0.2], [-0.1, 1.7]]), (3.0, -4.0)),
                   (np.array([[2.0, 0.0], [0.0, -1.0]]), (-40.0, 60.0))]:
        ref = ndimage.affine_transform(a, M, offset=off, order=3, mode="nearest",
                                       prefilter=False)
        assert np.array_equal(_affine(a, M, off), ref)


@pytest.mark.parametrize("shape", [(10, 12), (64, 48), (3, 3)])
def test_box_builders_match_numpy_slicing(gpu, shape):
    H, W = shape
    for d in S.Direction:
        for cov in (0.0, 0.05, 0.3, 1.0, 1.7):
            got = S.linear(shape, d, cov, 20.0, P).field
            assert got.is_cuda
            assert np.array_equal(got.cpu().numpy(), S._stripe_mask(shape, d, cov, 20.0))
    for c, s in [((5, 6), (4, 4)), ((0, 0), (6, 6)), ((-3, 20), (4, 5)), ((9, 1), (30, 2)),
                 ((H // 2, W // 2), (H, W))]:
        ref = np.zeros(shape, np.float32)
        ref[int(c[0] - s[0] / 2):int(c[0] + s[0] / 2), int(c[1] - s[1] / 2):int(c[1] + s[1] / 2)] = 3.5
        got = S.rectangular(shape, c, s, 3.5, P).field
        assert np.array_equal(got.cpu().numpy(), ref), (c, s)


def test_bad_direction_still_raises(gpu):
    with pytest.raises(ValueError):
        S.linear((8, 8), 7, 0.2, 20.0, P)
    with pytest.raises(ValueError):
        S.triangular((8, 8), 7, 10.0, 0.2, 20.0, P)
