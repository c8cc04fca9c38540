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
