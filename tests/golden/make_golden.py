"""Writes the golden fixtures tests/golden/*.npz from the FP64 oracle.

Each fixture holds the inputs of one ``_forward_euler`` call (state, D,
params, stimuli, t, t_end, dt, dx) and the oracle's FP64 outputs at a few
checkpoints. Cases: 16^2 to 64^2 grids, <= 200 steps, PARAMSET_1A, 3, 5 and
the MAXFLOAT-sentinel set 7; uniform and smoothed-scar D; one and two
overlapping stimuli; stimuli touching the boundary; a non-square grid.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cardiax_amd import params as P  # noqa: E402
from cardiax_amd import workloads as wl  # noqa: E402
from cardiax_amd.stimulus import Direction, _stripe_mask  # noqa: E402
from oracle import fk_oracle as o  # noqa: E402


def smooth_scar(shape, rng, depth=0.6):
    """D = 1e-3 with one Gaussian dip (sigma = side / 6) at a seeded centre:
    smooth enough that the reference scheme stays bounded on small grids."""
    H, W = shape
    cy, cx = rng.uniform(0.3, 0.7) * H, rng.uniform(0.3, 0.7) * W
    sig = min(H, W) / 6.0
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    dip = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig ** 2))
    return 1e-3 * (1.0 - depth * dip)


def case(name, shape, pname, D, stimuli, steps, checkpoints, seed, init="rest"):
    rng = np.random.default_rng(seed)
    if init == "rest":
        st = o.init(shape)
    else:  # a perturbed, partly excited state
        st = o.State(rng.uniform(0.3, 1.0, shape), rng.uniform(0.3, 1.0, shape),
                     rng.uniform(0.0, 0.9, shape))
    # the fixture stores float32 inputs, so the FP64 run starts from exactly
    # those values
    st = o.State(*(np.asarray(x, np.float32).astype(np.float64) for x in st))
    D = np.asarray(D, np.float32).astype(np.float64)
    stimuli = [o.Stimulus(s.protocol, np.asarray(s.field, np.float32).astype(np.float64))
               for s in stimuli]
    prm = P.ALL_PARAMSETS[pname]
    outs = []
    cur, t = st, 0
    for c in checkpoints:
        cur = o.forward_euler(cur, t, c, prm, D, stimuli, wl.DT, wl.DX)
        outs.append(np.stack(cur))
        t = c
    assert checkpoints[-1] == steps
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        state0=np.stack(st).astype(np.float32), D=D.astype(np.float32),
        params=np.array([float(x) for x in prm]), paramset=pname,
        protocols=np.array([[float(x) for x in s.protocol] for s in stimuli]).reshape(-1, 3),
        fields=np.array([s.field for s in stimuli], np.float32).reshape(-1, *shape),
        checkpoints=np.array(checkpoints), dt=wl.DT, dx=wl.DX,
        expected=np.array(outs))  # (n_checkpoints, 3, H, W) float64
    print(name, shape, pname, "max |u| =", float(np.abs(outs[-1][2]).max()))


def main():
    rng = np.random.default_rng(1234)
    prot = lambda s: o.Protocol(s, 2, 1e9)
    # 1: 16x16, PARAMSET_3, uniform D, WEST stripe at t=0
    sh = (16, 16)
    case("fk_16_p3_west", sh, "3", np.full(sh, 1e-3),
         [o.Stimulus(prot(0), _stripe_mask(sh, Direction.WEST, 0.25, 20.0))], 120,
         [40, 120], 1)
    # 2: 32x32, PARAMSET_1A, scar D, two overlapping stimuli (later overrides)
    sh = (32, 32)
    f2 = wl.rect_mask(sh, (16, 16), (12, 20), 20.0)
    f3 = np.zeros(sh, np.float32); f3[10:22, 8:12] = -4.0
    case("fk_32_p1a_overlap", sh, "1A", smooth_scar(sh, rng),
         [o.Stimulus(prot(0), f2), o.Stimulus(prot(1), f3)], 150, [50, 150], 2)
    # 3: 48x40 non-square, PARAMSET_5, perturbed initial state, boundary stimuli
    sh = (48, 40)
    fs = _stripe_mask(sh, Direction.SOUTH, 0.1, 20.0)
    fe = _stripe_mask(sh, Direction.EAST, 0.1, 20.0)
    case("fk_48x40_p5_edges", sh, "5", smooth_scar(sh, rng),
         [o.Stimulus(o.Protocol(5, 2, 60), fs), o.Stimulus(prot(30), fe)], 200,
         [100, 200], 3, init="perturbed")
    # 4: 64x64, PARAMSET_7 (MAXFLOAT sentinels), S1 + S2
    sh = (64, 64)
    case("fk_64_p7_s1s2", sh, "7", np.full(sh, 1e-3),
         [o.Stimulus(prot(0), _stripe_mask(sh, Direction.WEST, 0.1, 20.0)),
          o.Stimulus(prot(100), wl.rect_mask(sh, (48, 48), (32, 32), 20.0))], 200,
         [100, 200], 4)
    # 5: tiny 5x7 grid (boundary rules overlap), PARAMSET_3, perturbed state
    sh = (5, 7)
    case("fk_5x7_p3_tiny", sh, "3", np.full(sh, 8e-4), [], 60, [20, 60], 5,
         init="perturbed")


if __name__ == "__main__":
    main()
