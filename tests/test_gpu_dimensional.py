"""GPU test of forward_dimensional (cardiax/solve.py:133-165): sizes in cm and
times in ms convert with int() truncation, and the result is forward() over
the converted checkpoints, bit for bit."""
import numpy as np
import pytest
import torch

from cardiax_amd import convert, solve
from cardiax_amd import workloads as wl
from cardiax_amd.stimulus import Stimulus

pytestmark = pytest.mark.gpu


def test_forward_dimensional_equals_forward(gpu):
    sim = wl.c1_inputs(6, shape=(64, 48), steps=400)
    stims = [Stimulus(s.protocol, s.field) for s in sim.stimuli]
    # 0.64 x 0.48 cm, 3.999 ms (truncates to 399 steps), a checkpoint every 1 ms
    got = solve.forward_dimensional((0.64, 0.48), 3.999, 1, sim.params, sim.D, stims, 0.01,
                                    0.01)
    cps = list(range(0, convert.ms_to_units(3.999, 0.01), convert.ms_to_units(1, 0.01)))
    assert cps == [0, 100, 200, 300]
    ref = solve.forward(solve.init((64, 48)), cps, sim.params, sim.D, stims, 0.01, 0.01)
    assert len(got) == len(ref) == 3
    for a, b in zip(got, ref):
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    assert float(got[-1].u.max()) > 0.1
