"""Pins of the oracle's Heun integrator (oracle/fk_oracle.py step_heun,
cardiax/solve.py:73-85).

A uniform field with no stimulus has no diffusion, so every cell follows the
same ODE. On it the global error of Heun's method falls 4x when dt halves
(second order) and Euler's 2x (first order), against a reference solution at
a much finer dt; dropping a stage, mis-weighting the average or evaluating
k2 at the wrong point breaks the second-order ratio.
"""
import numpy as np
import pytest

from cardiax_amd import params as P
from oracle import fk_oracle as o


def _uniform(shape=(4, 5), u=0.3, v=0.8, w=0.9):
    return o.State(np.full(shape, v), np.full(shape, w), np.full(shape, u))


def _solve(method, dt, T):
    st = _uniform()
    n = int(round(T / dt))
    f = o.forward_heun if method == "heun" else o.forward_euler
    return f(st, 0, n, P.PARAMSET_3, np.full((4, 5), 1e-3), [], dt, 0.01)


@pytest.mark.parametrize("method,order", [("heun", 2), ("euler", 1)])
def test_convergence_order_on_the_uniform_ode(method, order):
    T = 2.0
    ref = _solve("heun", 0.0005, T)
    errs = []
    for dt in (0.02, 0.01, 0.005):
        x = _solve(method, dt, T)
        errs.append(max(float(np.abs(a - b).max()) for a, b in zip(x, ref)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert abs(np.log2(r1) - order) < 0.25 and abs(np.log2(r2) - order) < 0.25, (errs, r1, r2)


def test_one_heun_step_is_the_average_of_the_two_slopes():
    rng = np.random.default_rng(0)
    shape = (9, 11)
    st = o.State(rng.uniform(0.3, 1, shape), rng.uniform(0.3, 1, shape), rng.uniform(0, 1, shape))
    D = rng.uniform(5e-4, 1e-3, shape)
    f = np.zeros(shape); f[:, :2] = 1.0
    stims = [o.Stimulus(o.Protocol(0, 2, 1e9), f)]
    dt = 0.01
    k1 = o.step(st, 0.0, P.PARAMSET_3, D, stims, 0.01)
    x1 = o.step_euler(st, 0.0, P.PARAMSET_3, D, stims, dt, 0.01)
    k2 = o.step(x1, 0.0, P.PARAMSET_3, D, stims, 0.01)
    h = o.step_heun(st, 0.0, P.PARAMSET_3, D, stims, dt, 0.01)
    for x, a, b, y in zip(st, k1, k2, h):
        np.testing.assert_allclose(y, x + 0.5 * dt * (a + b), rtol=0, atol=1e-15)
    # Heun and Euler differ by dt^2/2 times the change of slope
    e = o.step_euler(st, 0.0, P.PARAMSET_3, D, stims, dt, 0.01)
    assert max(float(np.abs(a - b).max()) for a, b in zip(h, e)) > 0
