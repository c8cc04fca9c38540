"""The two-steps-per-launch streaming kernels (variants streamt8/12/16/24)
against the one-step streaming kernel, bit for bit.

Both compute every cell with the same operations in the same order, so two
steps in one launch must equal two launches exactly. The shapes cover the
smallest grids the kernels accept (2R + 16 rows; 56 to 96 columns in two
strips, or 104 and more), odd sides,
grids where the middle bands and strips overlap, and 1024^2; the stimuli are
active on the first and on the second step of a launch, touch the edges and
repeat; the parameter sets include the MAXFLOAT sentinel sets; the step
counts are odd (the last step runs alone) and even.
"""
import numpy as np
import pytest
import torch

from cardiax_amd import _lib, solve
from cardiax_amd import params as P
from cardiax_amd import workloads as wl
from cardiax_amd.stimulus import Protocol

pytestmark = pytest.mark.gpu

ROWS = {"streamt8": 8, "streamt12": 12, "streamt16": 16, "streamt24": 24}


def _run(variant, st, t0, t1, prm, D, stims):
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS[variant])
    try:
        out = solve._forward_euler(st, t0, t1, prm, D, stims, 0.01, 0.01)
        torch.cuda.synchronize()
        return out
    finally:
        L.fk_set_forward_variant(prev)


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    H, W = shape
    st = tuple(torch.from_numpy(x.astype(np.float32)).cuda() for x in
               (rng.uniform(0.2, 1.0, shape), rng.uniform(0.2, 1.0, shape),
                rng.uniform(-0.05, 1.05, shape)))
    D = wl.scar_map(shape, rng)
    f1 = np.zeros(shape, np.float32); f1[:, : max(1, W // 10)] = 20.0   # left edge
    f2 = np.zeros(shape, np.float32); f2[H - H // 3:, W // 2:] = 15.0  # bottom-right corner
    f3 = np.zeros(shape, np.float32); f3[: H // 4, W // 3: W // 3 + 9] = -3.0
    stims = [(Protocol(4, 2, 1e9), f1),   # active at 4, 5: both levels of one launch
             (Protocol(7, 1, 1e9), f2),   # active at 8 only: first level of a launch
             (Protocol(2, 2, 6), f3)]     # periodic: 2, 3, then 8, 9, 14, 15, ...
    return st, D, stims


@pytest.mark.parametrize("variant", sorted(ROWS))
@pytest.mark.parametrize("shape_kind", ["min", "min_odd", "overlap", "wide", "two_strips",
                                        "two_strips_wide"])
@pytest.mark.parametrize("pname", ["3", "2", "7"])
def test_two_steps_equal_one_step_kernel(gpu, variant, shape_kind, pname):
    R = ROWS[variant]
    shape = {"min": (2 * R + 16, 104), "min_odd": (2 * R + 17, 113),
             "overlap": (5 * R + 3, 3 * 48 + 21), "wide": (97, 517),
             "two_strips": (2 * R + 19, 56), "two_strips_wide": (3 * R + 5, 96)}[shape_kind]
    st, D, stims = _inputs(shape, seed=R + len(shape_kind) + int(pname))
    prm = P.ALL_PARAMSETS[pname]
    for t0, t1 in ((1, 22), (2, 22)):  # 21 steps (odd) and 20
        ref = _run("stream2", st, t0, t1, prm, D, stims)
        got = _run(variant, st, t0, t1, prm, D, stims)
        for name, a, b in zip("vwu", got, ref):
            assert torch.isfinite(b).all(), (name, t0)
            assert torch.equal(a, b), (variant, shape, pname, t0, name,
                                       float((a - b).abs().max()))


@pytest.mark.parametrize("variant", sorted(ROWS))
def test_two_steps_equal_one_step_kernel_1024(gpu, variant):
    sim = wl.c2_inputs(0, steps=400)
    stims = [(s.protocol._replace(start=0 if k == 0 else 151), s.field)
             for k, s in enumerate(sim.stimuli)]
    st = solve.init(sim.shape)
    ref = _run("stream2", st, 0, 401, sim.params, sim.D, stims)
    got = _run(variant, st, 0, 401, sim.params, sim.D, stims)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    assert float(ref.u.max()) > 0.5


def test_small_grids_fall_back(gpu):
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS["streamt24"])
    try:
        assert L.fk_forward_variant_for(63, 512) == _lib.VARIANTS["stream2"]  # < 2R + 16 rows
        assert L.fk_forward_variant_for(64, 100) == _lib.VARIANTS["stream2"]  # 97..103 columns
        assert L.fk_forward_variant_for(64, 55) == _lib.VARIANTS["stream2"]
        assert L.fk_forward_variant_for(64, 104) == _lib.VARIANTS["streamt24"]
    finally:
        L.fk_set_forward_variant(prev)
    st, D, stims = _inputs((40, 60), 1)
    a = _run("streamt24", st, 0, 9, P.PARAMSET_3, D, stims)
    b = _run("stream2", st, 0, 9, P.PARAMSET_3, D, stims)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
