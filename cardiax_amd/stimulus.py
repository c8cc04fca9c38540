"""Stimulus types and host-side field builders.

Mirrors cardiax/stimulus.py:13-140. ``Protocol`` and ``Stimulus`` are
NamedTuples with the reference's fields; protocol times are in step units.
Fields are built once on the host with NumPy slicing (so negative or empty
slices behave exactly as the reference's ``jax.ops.index_update`` slices do)
and handed to the solver as float32 tensors on the default device.
"""
from enum import IntEnum
from typing import NamedTuple, Tuple

import numpy as np

from ._device import as_field

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
    mask = np.zeros(shape, dtype=np.float32)
    x1 = int(centre[0] - size[0] / 2)
    x2 = int(centre[0] + size[0] / 2)
    y1 = int(centre[1] - size[1] / 2)
    y2 = int(centre[1] + size[1] / 2)
    mask[x1:x2, y1:y2] = modulus
    return Stimulus(protocol, as_field(mask))


def _stripe_mask(shape: Shape, direction, coverage: float, modulus: float) -> np.ndarray:
    # The stripe width is taken from shape[0] for every direction, as in
    # cardiax/stimulus.py:88. A width of 0 makes the EAST/SOUTH slices
    # ``-0:`` cover the whole field, again exactly as the reference slices do.
    s = int(shape[0] * coverage)
    field = np.zeros(shape, dtype=np.float32)
    if direction == Direction.WEST:
        field[:, :s] = modulus
    elif direction == Direction.EAST:
        field[:, -s:] = modulus
    elif direction == Direction.NORTH:
        field[:s, :] = modulus
    elif direction == Direction.SOUTH:
        field[-s:, :] = modulus
    else:
        raise ValueError(
            "direction mus be either 'left', 'right', 'up', or 'down' not %s" % direction)
    return field


def linear(shape: Shape, direction: Direction, coverage: float, modulus: float,
           protocol: Protocol) -> Stimulus:
    """cardiax/stimulus.py:63-106: a stripe of width int(shape[0]*coverage)
    along one edge; ValueError for an unknown direction."""
    return Stimulus(protocol, as_field(_stripe_mask(shape, direction, coverage, modulus)))


def triangular(shape: Shape, direction: Direction, angle: float, coverage: float,
               modulus: float, protocol: Protocol) -> Stimulus:
    """cardiax/stimulus.py:109-140: the ``linear`` stripe rotated by ``angle``
    degrees with a cubic spline, no prefilter, nearest-edge extension and the
    shape kept."""
    from scipy import ndimage

    stripe = _stripe_mask(shape, direction, coverage, modulus)
    field = ndimage.rotate(stripe, angle=angle, mode="nearest", prefilter=False,
                           reshape=False)
    return Stimulus(protocol, as_field(np.ascontiguousarray(field, dtype=np.float32)))
