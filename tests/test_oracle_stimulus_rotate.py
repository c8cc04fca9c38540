"""CPU pins of the ``triangular`` stimulus rotation (cardiax/stimulus.py:
137-139) and the argument checks of the device stimulus builders.

The oracle's FP64 restatement of scipy.ndimage.rotate(mode="nearest",
prefilter=False, reshape=False, order 3) is compared with scipy itself, the
library the reference calls: bit for bit, which includes the set of nonzero
cells (the solver replaces j_ion exactly where a field is nonzero, so that set
is what a stimulus means). A closed form pins it independently of scipy.
"""
import ctypes

import numpy as np
import pytest
from scipy import ndimage

from cardiax_amd import _lib
from cardiax_amd import stimulus as S
from oracle import stimulus_oracle as so

ANGLES = [0, 90, 180, 270, -90, 30, 45, 12.5, -20, 135, 200, 359, 1e-3, 77.7, 450]
SHAPES = [(64, 64), (50, 70), (33, 17), (5, 9), (1, 6)]


def _inputs(shape, rng):
    stripe = np.zeros(shape, np.float32)
    stripe[:, : max(1, int(shape[0] * 0.2))] = 20.0
    return {"stripe": stripe, "random": rng.standard_normal(shape).astype(np.float32)}


@pytest.mark.parametrize("shape", SHAPES)
def test_rotate_restatement_matches_scipy(shape):
    rng = np.random.default_rng(sum(shape))
    for angle in ANGLES:
        for name, a in _inputs(shape, rng).items():
            ref = ndimage.rotate(a, angle=angle, mode="nearest", prefilter=False,
                                 reshape=False)
            got = so.rotate(a, angle).astype(np.float32)
            assert np.array_equal(got, ref), (angle, name, np.abs(got - ref).max())


def test_rotate_closed_forms():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((9, 12))
    # constant fields stay constant (weights sum to 1 and taps clamp to edges)
    np.testing.assert_allclose(so.rotate(np.full((9, 12), 3.0), 37.0), 3.0, atol=1e-14)
    # 180 degrees samples exact integer positions: the B-spline of unfiltered
    # samples at an integer is (a[k-1] + 4 a[k] + a[k+1]) / 6 per axis
    r = so.rotate(a, 180.0)
    k = np.array([1, 4, 1]) / 6.0
    sm = np.einsum("i,j,ijxy->xy", k, k,
                   np.stack([np.stack([np.roll(np.roll(a, -di, 0), -dj, 1)
                                       for dj in (-1, 0, 1)]) for di in (-1, 0, 1)]))
    np.testing.assert_allclose(r[1:-1, 1:-1], sm[::-1, ::-1][1:-1, 1:-1], atol=1e-13)


def test_device_builders_reject_bad_arguments():
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.fk_stimulus_box(fake, 0, 8, 0, 1, 0, 1, 1.0, None) == _lib.FK_ERR_ARG
    assert L.fk_stimulus_box(fake, 8, 8, -1, 1, 0, 1, 1.0, None) == _lib.FK_ERR_ARG
    assert L.fk_stimulus_box(fake, 8, 8, 0, 9, 0, 1, 1.0, None) == _lib.FK_ERR_ARG
    assert L.fk_stimulus_box(None, 8, 8, 0, 1, 0, 1, 1.0, None) == _lib.FK_ERR_ARG
    m = (ctypes.c_double * 4)(1.0, 0.0, 0.0, 1.0)
    o = (ctypes.c_double * 2)(0.0, 0.0)
    bad = (ctypes.c_double * 2)(float("nan"), 0.0)
    dst = ctypes.c_void_p(0x2000)
    assert L.fk_affine_transform(fake, 8, 8, m, bad, dst, None) == _lib.FK_ERR_ARG
    assert L.fk_affine_transform(fake, 8, 8, m, o, fake, None) == _lib.FK_ERR_ARG
    assert L.fk_affine_transform(fake, 0, 8, m, o, dst, None) == _lib.FK_ERR_ARG
    assert L.fk_affine_transform(fake, 8, 8, None, o, dst, None) == _lib.FK_ERR_ARG


def test_slices_normalise_like_numpy():
    # the device path fills [r0, r1) x [c0, c1) from slice.indices; check the
    # normalisation against NumPy slicing for the reference's slice forms
    for H, W in [(10, 12), (3, 3)]:
        for d in S.Direction:
            for cov in (0.0, 0.05, 0.3, 1.0, 1.7):
                rows, cols = S._stripe_slices((H, W), d, cov)
                r0, r1, _ = rows.indices(H)
                c0, c1, _ = cols.indices(W)
                box = np.zeros((H, W), np.float32)
                box[r0:max(r0, r1), c0:max(c0, c1)] = 1
                np.testing.assert_array_equal(box != 0,
                                              S._stripe_mask((H, W), d, cov, 1.0) != 0)
        for (c, s) in [((5, 6), (4, 4)), ((0, 0), (6, 6)), ((-3, 20), (4, 5)), ((9, 1), (30, 2))]:
            x1, x2 = int(c[0] - s[0] / 2), int(c[0] + s[0] / 2)
            y1, y2 = int(c[1] - s[1] / 2), int(c[1] + s[1] / 2)
            ref = np.zeros((H, W), np.float32)
            ref[x1:x2, y1:y2] = 1
            r0, r1, _ = slice(x1, x2).indices(H)
            c0, c1, _ = slice(y1, y2).indices(W)
            box = np.zeros((H, W), np.float32)
            box[r0:max(r0, r1), c0:max(c0, c1)] = 1
            np.testing.assert_array_equal(box, ref)
