"""Seeded synthetic inputs of the benchmark and parity workloads.

All randomness comes from ``numpy.random.default_rng(seed)``; the reference's
own experiments draw from ``jax.random`` (deepx/generate.py:16-83), which is
not available, so these inputs are project-defined with the same shapes and
magnitudes: dt = dx = 0.01 and D = 1e-3 (experiments/generate_fd_data_128.py:
13-14,116), stimulus modulus 20 with a 2-step pulse
(experiments/generate_fd_examples.py:18-19), PARAMSET_3 by default
(cardiax/params.py:60-62).

Fields are returned as NumPy float32 arrays; callers move them to the device.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import params as P
from .stimulus import Direction, Protocol, _stripe_mask

DT = 0.01
DX = 0.01
D0 = 1e-3
MODULUS = 20.0
PERIOD = 1e9  # a single pulse, as every reference call site uses


@dataclass
class HostStimulus:
    protocol: Protocol
    field: np.ndarray


@dataclass
class Sim:
    """One simulation: inputs of cardiax.solve._forward_euler."""
    params: P.Params
    D: np.ndarray
    stimuli: List[HostStimulus]
    t0: float
    t1: float
    dt: float = DT
    dx: float = DX
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.D.shape


def rect_mask(shape, centre, size, modulus):
    """Host mask of stimulus.rectangular (cardiax/stimulus.py:31-60)."""
    m = np.zeros(shape, np.float32)
    x1, x2 = int(centre[0] - size[0] / 2), int(centre[0] + size[0] / 2)
    y1, y2 = int(centre[1] - size[1] / 2), int(centre[1] + size[1] / 2)
    m[x1:x2, y1:y2] = modulus
    return m


def scar_map(shape, rng, n_min=1, n_max=5, axes=(20, 200), scale=(0.1, 0.2), base=D0,
             taper=(3, 5), taper_frac=0.1, min_taper=12.0):
    """Baseline diffusivity with elliptical scars.

    Each scar: uniform centre, full axes U[axes] cells (scaled by side/1024 for
    grids other than 1024), rotation U[0, pi), D multiplied inside by
    U[scale], with a smoothstep edge max(U[taper], taper_frac * side,
    min_taper) cells wide. Overlapping scars take the minimum factor, not the
    product. The edge scales with the grid like the reference's blur, whose
    kernel spans 10% of the image (deepx/utils_scars.py:402-426), never
    narrower than 12 cells, and the scale stays >= 0.1, the floor of the
    reference's own random D range [1e-4, 1e-3] (deepx/generate.py:80). These
    keep the dynamics bounded: the reference's non-conservative
    D (u_xx + u_yy) + D_x u_x + D_y u_y with its stencils has growing modes
    where D changes steeply (edges of ~6 cells on 64^2 grids diverge for
    PARAMSET_1E, 2, 5, 6 even in FP64), which excited tissue does not damp.
    """
    H, W = shape
    s = min(H, W) / 1024.0
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64),
                         indexing="ij")
    factor = np.ones(shape)
    for _ in range(int(rng.integers(n_min, n_max + 1))):
        cy, cx = rng.uniform(0, H), rng.uniform(0, W)
        a = max(2.0, rng.uniform(*axes) * s / 2)  # semi-axes in cells
        b = max(2.0, rng.uniform(*axes) * s / 2)
        th = rng.uniform(0, np.pi)
        f = rng.uniform(*scale)
        tw = max(rng.uniform(*taper), taper_frac * min(H, W), min_taper)  # edge width, cells
        dy, dx = yy - cy, xx - cx
        xr = dx * np.cos(th) + dy * np.sin(th)
        yr = -dx * np.sin(th) + dy * np.cos(th)
        r = np.sqrt((xr / a) ** 2 + (yr / b) ** 2)
        dist = (1.0 - r) * min(a, b)  # ~ signed distance inside the ellipse, in cells
        x = np.clip(dist / tw + 0.5, 0.0, 1.0)
        inside = x * x * (3 - 2 * x)
        # overlapping scars take the deeper one, so D never leaves
        # [scale[0] * base, base]
        factor = np.minimum(factor, 1.0 - (1.0 - f) * inside)
    return (base * factor).astype(np.float32)


def c1_inputs(seed: int = 0, shape=(128, 128), steps: int = 1000) -> Sim:
    """128^2, PARAMSET_3, uniform D = 1e-3, S1 = WEST stripe (coverage 0.1) at
    step 0, S2 = rectangle over the far quadrant at a seeded step in [300, 700)."""
    rng = np.random.default_rng(seed)
    H, W = shape
    s1 = HostStimulus(Protocol(0, 2, PERIOD), _stripe_mask(shape, Direction.WEST, 0.1, MODULUS))
    s2_start = int(rng.integers(300, 700))
    s2 = HostStimulus(Protocol(s2_start, 2, PERIOD),
                      rect_mask(shape, (3 * H // 4, 3 * W // 4), (H // 2, W // 2), MODULUS))
    D = np.full(shape, D0, np.float32)
    return Sim(P.PARAMSET_3, D, [s1, s2], 0, steps, name="c1_%dx%d" % shape,
               meta={"s2_start": s2_start})


def c2_inputs(seed: int = 0, shape=(1024, 1024), steps: int = 100_000) -> Sim:
    """1024^2, PARAMSET_3, scarred D, S1 = WEST stripe at step 0, S2 = NORTH
    half field (coverage 0.5) at a seeded step in [28k, 36k): spiral
    initiation. S2 timing is scaled with ``steps`` below 1e5."""
    rng = np.random.default_rng(seed)
    D = scar_map(shape, rng)
    s1 = HostStimulus(Protocol(0, 2, PERIOD), _stripe_mask(shape, Direction.WEST, 0.1, MODULUS))
    lo, hi = int(0.28 * 100_000), int(0.36 * 100_000)
    s2_start = int(rng.integers(lo, hi))
    if steps < 100_000:
        s2_start = int(s2_start * steps / 100_000)
    s2 = HostStimulus(Protocol(s2_start, 2, PERIOD),
                      _stripe_mask(shape, Direction.NORTH, 0.5, MODULUS))
    return Sim(P.PARAMSET_3, D, [s1, s2], 0, steps, name="c2_%dx%d" % shape,
               meta={"s2_start": s2_start})


C3_SIDES = (64, 128, 256, 384, 512)


def c3_inputs(seed: int = 0, n: int = 256, steps: int = 5000, sides=C3_SIDES) -> List[Sim]:
    """Batch of independent square simulations with sides drawn from
    ``sides``, a seeded parameter set from all named sets, a scar map and
    seeded S1/S2 timings (S2 in [0.3, 0.7) of the run, like the C2 spiral)."""
    rng = np.random.default_rng(seed)
    names = list(P.ALL_PARAMSETS)
    sims = []
    for k in range(n):
        side = int(rng.choice(sides))
        shape = (side, side)
        pname = names[int(rng.integers(len(names)))]
        D = scar_map(shape, rng)
        s1_start = int(rng.integers(0, 50))
        s2_start = int(rng.integers(int(0.3 * steps), int(0.7 * steps)))
        s1 = HostStimulus(Protocol(s1_start, 2, PERIOD),
                          _stripe_mask(shape, Direction.WEST, 0.1, MODULUS))
        s2 = HostStimulus(Protocol(s2_start, 2, PERIOD),
                          _stripe_mask(shape, Direction.NORTH, 0.5, MODULUS))
        sims.append(Sim(P.ALL_PARAMSETS[pname], D, [s1, s2], 0, steps,
                        name="c3_%d_%s_%d" % (k, pname, side), meta={"paramset": pname}))
    return sims
