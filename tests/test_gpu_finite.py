"""GPU test of the device-side non-finite check (fk_check_finite): a clean
state passes, a single NaN or inf anywhere in any field is reported as
FK_ERR_NONFINITE (FloatingPointError in Python), and the flag is cleared
between calls."""
import numpy as np
import pytest
import torch

from cardiax_amd import solve
from cardiax_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_check_finite_flags_nan_and_inf(gpu):
    sim = wl.c1_inputs(1, shape=(300, 257), steps=100)
    st = solve._forward_euler(solve.init((300, 257)), 0, 100, sim.params, sim.D,
                              [(s.protocol, s.field) for s in sim.stimuli], 0.01, 0.01)
    solve.check_finite(st)  # a healthy solve
    rng = np.random.default_rng(0)
    for bad in (float("nan"), float("inf"), float("-inf")):
        for k in range(3):
            broken = [x.clone() for x in st]
            i, j = rng.integers(0, 300), rng.integers(0, 257)
            broken[k][i, j] = bad
            with pytest.raises(FloatingPointError):
                solve.check_finite(solve.State(*broken))
            solve.check_finite(st)  # the flag starts clear on every call
    big = torch.ones((1024, 1024), device="cuda")
    big[-1, -1] = float("nan")  # the last element of a large field
    with pytest.raises(FloatingPointError):
        solve.check_finite((big, big.clone().fill_(1.0), big.clone().fill_(0.0)))


def test_unstable_dt_trips_the_flag(gpu):
    """D dt / dx^2 above the explicit-Euler bound of the composite stencil
    (0.531) blows up, and the device-side check reports it; the same run
    well inside the bound stays finite."""
    shape = (96, 96)
    sim = wl.c1_inputs(2, shape=shape, steps=400)
    D = np.full(shape, 1e-3, np.float32)  # D / dx^2 = 10 per unit time
    stims = [(s.protocol, s.field) for s in sim.stimuli]
    bad = solve._forward_euler(solve.init(shape), 0, 400, sim.params, D, stims, 0.1, 0.01)
    with pytest.raises(FloatingPointError):
        solve.check_finite(bad)  # D dt / dx^2 = 1.0
    ok = solve._forward_euler(solve.init(shape), 0, 400, sim.params, D, stims, 0.03, 0.01)
    solve.check_finite(ok)  # 0.3, inside the bound


def test_resume_from_a_checkpoint_is_bitwise(gpu):
    """A long run cut at an arbitrary step and resumed from the saved State
    equals the uncut run bit for bit (the kernels are deterministic)."""
    shape = (128, 160)
    sim = wl.c1_inputs(4, shape=shape, steps=1500)
    stims = [(s.protocol, s.field) for s in sim.stimuli]
    whole = solve._forward_euler(solve.init(shape), 0, 1500, sim.params, sim.D, stims, 0.01,
                                 0.01)
    part = solve._forward_euler(solve.init(shape), 0, 637, sim.params, sim.D, stims, 0.01, 0.01)
    saved = [x.cpu().numpy().copy() for x in part]  # snapshot to the host and back
    resumed = solve._forward_euler(solve.State(*saved), 637, 1500, sim.params, sim.D, stims,
                                   0.01, 0.01)
    for a, b in zip(whole, resumed):
        assert torch.equal(a, b)
