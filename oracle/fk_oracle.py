"""CPU restatement of the cardiax forward-Euler path: the parity oracle.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check
results. The product path (``cardiax_amd``) never imports it and has no CPU
fallback.

Every function below restates one reference function in NumPy. The default
dtype is float64 (the oracle proper); the same code run with float32 is the
op-graph-faithful fp32 restatement used to estimate fp32 rounding growth.

Reference behaviour followed (file:line under the reference tree):
  * ``State``/``init``        cardiax/solve.py:12-23
  * ``step``                  cardiax/solve.py:26-65
  * ``step_euler``            cardiax/solve.py:68-70
  * ``step_heun``             cardiax/solve.py:73-85
  * ``forward_heun``          cardiax/solve.py:103-111
  * ``forward_euler``         cardiax/solve.py:92-100 (fori_loop over [t, t_end))
  * ``forward``               cardiax/solve.py:168-222 (host segment loop)
  * ``gradient``              cardiax/solve.py:225-254
  * ``stimulate``             cardiax/solve.py:257-271
  * ``stimulus_active``       the activity predicate of cardiax/solve.py:262-267

How each part is pinned independently of this file is listed in the test
modules ``tests/test_oracle_*.py`` (exactness on polynomials, the rational
boundary tables, the stimulus known-answer sets of the reference's own tests,
reaction limits, invariants).
"""
from __future__ import annotations

import math
from typing import NamedTuple, Sequence

import numpy as np


class State(NamedTuple):
    """cardiax/solve.py:12-15 -- field order v, w, u."""
    v: np.ndarray
    w: np.ndarray
    u: np.ndarray


class Protocol(NamedTuple):
    """cardiax/stimulus.py:13-16 -- times in step units."""
    start: float
    duration: float
    period: float


class Stimulus(NamedTuple):
    """cardiax/stimulus.py:19-21."""
    protocol: Protocol
    field: np.ndarray


def init(shape, dtype=np.float64) -> State:
    """cardiax/solve.py:18-23: v = w = 1, u = 0."""
    return State(np.ones(shape, dtype), np.ones(shape, dtype), np.zeros(shape, dtype))


def _pad(a: np.ndarray) -> np.ndarray:
    """jnp.pad(a, 1, mode="edge") (cardiax/solve.py:29-32,45)."""
    return np.pad(a, 1, mode="edge")


def gradient(a: np.ndarray, axis: int) -> np.ndarray:
    """cardiax/solve.py:225-254.

    First derivative along ``axis`` without grid spacing: 4th-order central
    differences on entries 2..n-3, 3rd-order one-sided differences on the two
    outermost entries of each side. Requires n >= 5 along ``axis``.
    """
    x = np.moveaxis(np.asarray(a), axis, 0)
    n = x.shape[0]
    if n < 5:
        raise ValueError("gradient needs at least 5 entries along the axis, got %d" % n)
    dt = x.dtype
    c = lambda r: dt.type(r)  # coefficients carry the array's precision
    lo = c(-11 / 6) * x[0:2] + c(3) * x[1:3] - c(3 / 2) * x[2:4] + c(1 / 3) * x[3:5]
    mid = (c(1 / 12) * x[0:n - 4] - c(2 / 3) * x[1:n - 3]
           + c(2 / 3) * x[3:n - 1] - c(1 / 12) * x[4:n])
    hi = (c(-1 / 3) * x[n - 5:n - 3] + c(3 / 2) * x[n - 4:n - 2]
          - c(3) * x[n - 3:n - 1] + c(11 / 6) * x[n - 2:n])
    return np.moveaxis(np.concatenate((lo, mid, hi), axis=0), 0, axis)


def pymod(a: float, b: float) -> float:
    """Floor modulo with the sign of the divisor (jnp.mod / Python ``%``)."""
    return a % b


def stimulus_active(t: float, protocol: Protocol) -> bool:
    """Activity predicate of cardiax/solve.py:262-267:
    t >= start and mod(start - t + 1, period) < duration."""
    start, duration, period = (float(protocol.start), float(protocol.duration),
                               float(protocol.period))
    return (t >= start) and (pymod(start - t + 1.0, period) < duration)


def stimulate(t: float, X: np.ndarray, stimuli: Sequence[Stimulus]) -> np.ndarray:
    """cardiax/solve.py:257-271.

    Later stimuli override earlier ones where their field is nonzero; the
    result *replaces* X wherever the combined field is nonzero.
    """
    stimulated = np.zeros_like(X)
    for s in stimuli:
        active = stimulus_active(t, s.protocol)
        f = np.asarray(s.field, dtype=X.dtype)
        stimulated = np.where(f * active != 0, f, stimulated)
    return np.where(stimulated != 0, stimulated, X)


def reaction(v, w, u, params):
    """Pointwise part of cardiax/solve.py:35-42,57-58.

    Returns (d_v, d_w, j_ion) before stimulus replacement.
    """
    dt = u.dtype.type
    P = [dt(x) for x in params]
    (tau_v_plus, tau_v1_minus, tau_v2_minus, tau_w_plus, tau_w_minus, tau_d,
     tau_0, tau_r, tau_si, k, V_csi, V_c, V_v, Cm) = P
    one = dt(1)
    p = (u >= V_c).astype(u.dtype)
    q = (u >= V_v).astype(u.dtype)
    tau_v_minus = (one - q) * tau_v1_minus + q * tau_v2_minus
    j_fi = -v * p * (u - V_c) * (one - u) / tau_d
    j_so = (u * (one - p) / tau_0) + (p / tau_r)
    j_si = -(w * (one + np.tanh(k * (u - V_csi)))) / (dt(2) * tau_si)
    j_ion = -(j_fi + j_so + j_si) / Cm
    d_v = ((one - p) * (one - v) / tau_v_minus) - ((p * v) / tau_v_plus)
    d_w = ((one - p) * (one - w) / tau_w_minus) - ((p * w) / tau_w_plus)
    return d_v, d_w, j_ion


def diffusion(u: np.ndarray, D: np.ndarray, dx: float) -> np.ndarray:
    """The del_u term of cardiax/solve.py:49-55 on the padded domain."""
    dxt = u.dtype.type(dx)
    u_x = gradient(u, 0) / dxt
    u_y = gradient(u, 1) / dxt
    u_xx = gradient(u_x, 0) / dxt
    u_yy = gradient(u_y, 1) / dxt
    D_x = gradient(D, 0) / dxt
    D_y = gradient(D, 1) / dxt
    return D * (u_xx + u_yy) + (D_x * u_x) + (D_y * u_y)


def step(state: State, t: float, params, diffusivity, stimuli, dx) -> State:
    """cardiax/solve.py:26-65: time derivatives (d_v, d_w, d_u)."""
    v = _pad(state.v)
    w = _pad(state.w)
    u = _pad(state.u)
    D = _pad(np.asarray(diffusivity, dtype=u.dtype))
    d_v, d_w, j_ion = reaction(v, w, u, params)
    padded = [Stimulus(s.protocol, _pad(np.asarray(s.field, dtype=u.dtype))) for s in stimuli]
    j_ion = stimulate(t, j_ion, padded)
    d_u = diffusion(u, D, dx) + j_ion
    return State(d_v[1:-1, 1:-1], d_w[1:-1, 1:-1], d_u[1:-1, 1:-1])


def step_euler(state: State, t, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:68-70: x + dt * dx/dt, all from the same old state."""
    g = step(state, t, params, diffusivity, stimuli, dx)
    dtt = state.u.dtype.type(dt)
    return State(*(x + dxdt * dtt for x, dxdt in zip(state, g)))


def step_heun(state: State, t, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:73-85: k1 = step(x, t), x1 = x + dt k1, k2 = step(x1, t)
    (the same stimulus time), x + (k1 + k2) * (dt / 2)."""
    dtt = state.u.dtype.type(dt)
    k1 = step(state, t, params, diffusivity, stimuli, dx)
    x1 = State(*(x + d * dtt for x, d in zip(state, k1)))
    k2 = step(x1, t, params, diffusivity, stimuli, dx)
    half = dtt * state.u.dtype.type(0.5)
    return State(*(x + (a + b) * half for x, a, b in zip(state, k1, k2)))


def forward_heun(state: State, t, t_end, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:103-111: Heun steps for i in [t, t_end), step i at
    stimulus time i."""
    for i in loop_times(t, t_end):
        state = step_heun(state, i, params, diffusivity, stimuli, dt, dx)
    return state


def loop_times(t: float, t_end: float):
    """Loop indices of lax.fori_loop(t, t_end) (cardiax/solve.py:94-99)."""
    n = max(0, int(math.ceil(float(t_end) - float(t))))
    return [float(t) + i for i in range(n)]


def forward_euler(state: State, t, t_end, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:92-100: step i uses stimulus time i, i in [t, t_end)."""
    for i in loop_times(t, t_end):
        state = step_euler(state, i, params, diffusivity, stimuli, dt, dx)
    return state


def forward(state: State, checkpoints, params, diffusivity, stimuli, dt, dx):
    """cardiax/solve.py:168-222 (Euler branch): one State per checkpoint
    interval, the initial state excluded."""
    states = []
    cps = [float(c) for c in checkpoints]
    for a, b in zip(cps[:-1], cps[1:]):
        state = forward_euler(state, a, b, params, diffusivity, stimuli, dt, dx)
        states.append(state)
    return states


# ---------------------------------------------------------------------------
# Effective 1-D operators on original cells, for the adjoint and the tests.
# ---------------------------------------------------------------------------

def effective_operators(n: int):
    """Return (A1, A2), the n x n matrices of crop(G(pad(.))) and
    crop(G(G(pad(.)))) along one axis, built by applying ``gradient`` to unit
    vectors (spacing not applied)."""
    A1 = np.zeros((n, n))
    A2 = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        e2 = np.tile(e[:, None], (1, 5))  # 5 columns: gradient needs a 2-D array only for moveaxis
        g1 = gradient(_pad(e2), 0)
        g2 = gradient(g1, 0)
        A1[:, j] = g1[1:-1, 2]
        A2[:, j] = g2[1:-1, 2]
    return A1, A2
