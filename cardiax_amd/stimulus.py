"""Stimulus types and field builders.

Mirrors cardiax/stimulus.py:13-140. ``Protocol`` and ``Stimulus`` are
NamedTuples with the reference's fields; protocol times are in step units.
The builders return float32 fields on the default device. With a GPU they
are built there: the reference's slices (``jax.ops.index_update``) are
normalised by Python slice semantics, so negative and empty slices behave as
the reference's do, and filled by ``fk_stimulus_box``; ``triangular``'s
rotation runs in ``fk_affine_transform`` with the affine map scipy's
``rotate`` builds, bit-identical to scipy's result. Without a GPU the same
fields are built on the host with NumPy slicing and scipy's ``rotate``, as the
reference builds them.
"""
from enum import IntEnum
from typing import NamedTuple, Tuple

import numpy as np

from ._device import as_field, gpu_available

Shape = Tuple[int, ...]
Point2D = Tuple[int, int]


class Protocol(NamedTuple):
    """cardiax/stimulus.py:13-16 -- start, duration and period in steps."""
    start: int
    duration: int
    period: int


class Stimulus(NamedTuple):
    """cardiax/stimulus.py:19-21."""
    protocol: Protocol
    field: object


class Direction(IntEnum):
    """cardiax/stimulus.py:24-28."""
    NORTH = 0
    EAST = 1
    SOUTH = 2
    WEST = 3


def rectangular(shape: Shape, centre: Point2D, size: Point2D, modulus: float,
                protocol: Protocol) -> Stimulus:
    """cardiax/stimulus.py:31-60: ``modulus`` on
    [int(c0 - s0/2):int(c0 + s0/2), int(c1 - s1/2):int(c1 + s1/2)], 0 elsewhere."""
    x1 = int(centre[0] - size[0] / 2)
    x2 = int(centre[0] + size[0] / 2)
    y1 = int(centre[1] - size[1] / 2)
    y2 = int(centre[1] + size[1] / 2)
    rows, cols = slice(x1, x2), slice(y1, y2)
    if gpu_available():
        return Stimulus(protocol, _box_on_device(shape, rows, cols, modulus))
    mask = np.zeros(shape, dtype=np.float32)
    mask[rows, cols] = modulus
    return Stimulus(protocol, as_field(mask))


def _stripe_slices(shape: Shape, direction, coverage: float):
    # The stripe width is taken from shape[0] for every direction, as in
    # cardiax/stimulus.py:88. A width of 0 makes the EAST/SOUTH slices
    # ``-0:`` cover the whole field, again exactly as the reference slices do.
    s = int(shape[0] * coverage)
    every = slice(None)
    if direction == Direction.WEST:
        return every, slice(None, s)
    if direction == Direction.EAST:
        return every, slice(-s, None)
    if direction == Direction.NORTH:
        return slice(None, s), every
    if direction == Direction.SOUTH:
        return slice(-s, None), every
    raise ValueError(
        "direction mus be either 'left', 'right', 'up', or 'down' not %s" % direction)


def _box_on_device(shape: Shape, rows: slice, cols: slice, modulus: float):
    """A float32 field on the GPU with ``modulus`` on field[rows, cols]."""
    import torch

    from . import _lib
    from ._device import require_gpu, stream_handle

    H, W = (int(x) for x in shape)
    r0, r1, rs = rows.indices(H)
    c0, c1, cs = cols.indices(W)
    assert rs == 1 and cs == 1
    r1, c1 = max(r0, r1), max(c0, c1)  # an empty slice
    field = torch.empty((H, W), dtype=torch.float32, device=require_gpu("stimulus"))
    _lib.check(_lib.lib().fk_stimulus_box(field.data_ptr(), H, W, r0, r1, c0, c1,
                                          float(modulus), stream_handle()))
    return field


def _stripe_mask(shape: Shape, direction, coverage: float, modulus: float) -> np.ndarray:
    rows, cols = _stripe_slices(shape, direction, coverage)
    field = np.zeros(shape, dtype=np.float32)
    field[rows, cols] = modulus
    return field


def linear(shape: Shape, direction: Direction, coverage: float, modulus: float,
           protocol: Protocol) -> Stimulus:
    """cardiax/stimulus.py:63-106: a stripe of width int(shape[0]*coverage)
    along one edge; ValueError for an unknown direction."""
    if gpu_available():
        rows, cols = _stripe_slices(shape, direction, coverage)
        return Stimulus(protocol, _box_on_device(shape, rows, cols, modulus))
    return Stimulus(protocol, as_field(_stripe_mask(shape, direction, coverage, modulus)))


def triangular(shape: Shape, direction: Direction, angle: float, coverage: float,
               modulus: float, protocol: Protocol) -> Stimulus:
    """cardiax/stimulus.py:109-140: the ``linear`` stripe rotated by ``angle``
    degrees with a cubic spline, no prefilter, nearest-edge extension and the
    shape kept."""
    if gpu_available():
        import ctypes

        import torch
        from scipy import special

        from . import _lib
        from ._device import stream_handle

        stripe = linear(shape, direction, coverage, modulus, protocol).field
        field = torch.empty_like(stripe)
        H, W = (int(x) for x in shape)
        # the map scipy.ndimage.rotate hands to affine_transform (axes (1, 0),
        # reshape=False), built with the same functions
        c, s = special.cosdg(angle), special.sindg(angle)
        M = np.array([[c, s], [-s, c]])
        centre = (np.asarray([H, W]) - 1) / 2
        offset = centre - M @ centre
        m = (ctypes.c_double * 4)(*[float(x) for x in M.ravel()])
        o = (ctypes.c_double * 2)(*[float(x) for x in offset])
        _lib.check(_lib.lib().fk_affine_transform(stripe.data_ptr(), H, W, m, o,
                                                  field.data_ptr(), stream_handle()))
        return Stimulus(protocol, field)
    from scipy import ndimage

    stripe = _stripe_mask(shape, direction, coverage, modulus)
    field = ndimage.rotate(stripe, angle=angle, mode="nearest", prefilter=False,
                           reshape=False)
    return Stimulus(protocol, as_field(np.ascontiguousarray(field, dtype=np.float32)))
