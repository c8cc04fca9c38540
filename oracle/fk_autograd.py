"""FP64 torch-autograd restatement of the cardiax Euler step (CPU).

TEST INFRASTRUCTURE ONLY (see fk_oracle.py). jax is not available, so this
differentiable restatement of cardiax/solve.py:26-70,225-271 stands in for
jax.grad through the reference: the same operation graph (edge pad, gates as
0/1 masks from comparisons, tanh, replacement by jnp.where, six `gradient`s,
crop), differentiated by torch in float64. Comparisons carry no gradient and
torch.where routes the cotangent to the selected branch only, as in JAX.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def _pad(a):
    return F.pad(a[None, None], (1, 1, 1, 1), mode="replicate")[0, 0]


def gradient(a, axis):
    """cardiax/solve.py:225-254 along dim `axis` of a 2-D tensor."""
    x = a if axis == 0 else a.t()
    n = x.shape[0]
    lo = (-11 / 6) * x[0:2] + 3 * x[1:3] - 1.5 * x[2:4] + (1 / 3) * x[3:5]
    mid = (1 / 12) * x[0:n - 4] - (2 / 3) * x[1:n - 3] + (2 / 3) * x[3:n - 1] - (1 / 12) * x[4:n]
    hi = (-1 / 3) * x[n - 5:n - 3] + 1.5 * x[n - 4:n - 2] - 3 * x[n - 3:n - 1] + (11 / 6) * x[n - 2:n]
    g = torch.cat((lo, mid, hi), 0)
    return g if axis == 0 else g.t()


def active(t, protocol):
    start, duration, period = (float(x) for x in protocol)
    return (t >= start) and (((start - t + 1.0) % period) < duration)


def step_euler(v, w, u, t, prm, D, stimuli, dt, dx):
    """One Euler step; prm is a dict with tensors or floats for every
    Params field (cardiax/params.py:4-18)."""
    vp, wp, up, Dp = _pad(v), _pad(w), _pad(u), _pad(D)
    p = (up >= prm["V_c"]).to(up.dtype)
    q = (up >= prm["V_v"]).to(up.dtype)
    tvm = (1 - q) * prm["tau_v1_minus"] + q * prm["tau_v2_minus"]
    j_fi = -vp * p * (up - prm["V_c"]) * (1 - up) / prm["tau_d"]
    j_so = (up * (1 - p) / prm["tau_0"]) + (p / prm["tau_r"])
    j_si = -(wp * (1 + torch.tanh(prm["k"] * (up - prm["V_csi"])))) / (2 * prm["tau_si"])
    j_ion = -(j_fi + j_so + j_si) / prm["Cm"]
    stimulated = torch.zeros_like(j_ion)
    for s in stimuli:
        f = _pad(torch.as_tensor(s[1], dtype=up.dtype))
        if active(t, s[0]):
            stimulated = torch.where(f != 0, f, stimulated)
    j_ion = torch.where(stimulated != 0, stimulated, j_ion)
    u_x = gradient(up, 0) / dx
    u_y = gradient(up, 1) / dx
    u_xx = gradient(u_x, 0) / dx
    u_yy = gradient(u_y, 1) / dx
    D_x = gradient(Dp, 0) / dx
    D_y = gradient(Dp, 1) / dx
    del_u = Dp * (u_xx + u_yy) + D_x * u_x + D_y * u_y
    d_v = ((1 - p) * (1 - vp) / tvm) - ((p * vp) / prm["tau_v_plus"])
    d_w = ((1 - p) * (1 - wp) / prm["tau_w_minus"]) - ((p * wp) / prm["tau_w_plus"])
    d_u = del_u + j_ion
    c = (slice(1, -1), slice(1, -1))
    return v + dt * d_v[c], w + dt * d_w[c], u + dt * d_u[c]


def rollout_frames(state0, params, D, stimuli, t0, n_steps, frame_every, dt, dx,
                   taus=None):
    """u after every frame_every steps (stacked), float64, differentiable in
    D and in the tensors of `taus` = (tau_d, tau_r, tau_si)."""
    names = ("tau_v_plus", "tau_v1_minus", "tau_v2_minus", "tau_w_plus", "tau_w_minus",
             "tau_d", "tau_0", "tau_r", "tau_si", "k", "V_csi", "V_c", "V_v", "Cm")
    prm = {k: float(x) for k, x in zip(names, params)}
    if taus is not None:
        prm["tau_d"], prm["tau_r"], prm["tau_si"] = taus
    v, w, u = (torch.as_tensor(x, dtype=torch.float64) for x in state0)
    frames = []
    for k in range(n_steps):
        v, w, u = step_euler(v, w, u, float(t0 + k), prm, D, stimuli, dt, dx)
        if (k + 1) % frame_every == 0:
            frames.append(u)
    return torch.stack(frames)
