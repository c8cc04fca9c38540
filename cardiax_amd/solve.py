"""Drop-in replacement for the forward-Euler path of ``cardiax.solve``.

Same public names, positional signatures and field orders as
cardiax/solve.py:12-271; every numerical step runs in the HIP kernels of
``libcardiax_amd.so`` through the C-ABI of include/cardiax_amd.h. Arrays are
returned as float32 torch tensors on the current ROCm device. Inputs may be
torch tensors (any device) or NumPy arrays; they are copied to the device once
per call when needed. Like the reference's functional JAX code, no function
modifies its inputs.
"""
from __future__ import annotations

import ctypes
from typing import NamedTuple, Sequence

import torch

from . import _lib
from ._device import on_device, require_gpu, stream_handle
from . import convert


class State(NamedTuple):
    """cardiax/solve.py:12-15 -- field order v, w, u is part of the interface."""
    v: torch.Tensor
    w: torch.Tensor
    u: torch.Tensor


# ---------------------------------------------------------------------------
# argument marshalling
# ---------------------------------------------------------------------------

def _shape_of(state) -> tuple:
    shapes = {tuple(x.shape) for x in state}
    if len(shapes) != 1:
        raise ValueError("state fields have different shapes: %s" % sorted(shapes))
    shape = shapes.pop()
    if len(shape) != 2:
        raise ValueError("state fields must be 2-D, got shape %s" % (shape,))
    return shape


def _dev_state(state, dev) -> State:
    if len(state) != 3:
        raise TypeError("state must be a (v, w, u) triple")
    return State(*(on_device(x, dev, n) for x, n in zip(state, ("v", "w", "u"))))


def _dev_field(a, dev, shape, name) -> torch.Tensor:
    t = on_device(a, dev, name)
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
    return t


class _Stimuli:
    """ctypes array of fk_stimulus; keeps the device fields alive."""

    def __init__(self, stimuli, dev, shape):
        stimuli = list(stimuli) if stimuli is not None else []
        if len(stimuli) > _lib.FK_MAX_STIMULI:
            raise NotImplementedError("at most %d stimuli per call, got %d"
                                      % (_lib.FK_MAX_STIMULI, len(stimuli)))
        self.fields = []
        arr = (_lib.FkStimulus * max(1, len(stimuli)))()
        for i, s in enumerate(stimuli):
            protocol, field = s[0], s[1]
            f = _dev_field(field, dev, shape, "stimulus %d field" % i)
            self.fields.append(f)
            start, duration, period = (float(_scalar(x)) for x in protocol)
            arr[i] = _lib.FkStimulus(start, duration, period, f.data_ptr())
        self.arr = arr
        self.n = len(stimuli)

    def pointer(self):
        return ctypes.cast(self.arr, ctypes.POINTER(_lib.FkStimulus))


def _scalar(x) -> float:
    if isinstance(x, torch.Tensor):
        return float(x.reshape(-1)[0].item()) if x.numel() == 1 else _bad_scalar(x)
    try:
        import numpy as np
        if isinstance(x, np.ndarray):
            return float(x.reshape(-1)[0]) if x.size == 1 else _bad_scalar(x)
    except ImportError:  # pragma: no cover
        pass
    return float(x)


def _bad_scalar(x):
    raise TypeError("expected a scalar, got an array of shape %s" % (tuple(x.shape),))


# ---------------------------------------------------------------------------
# public API (cardiax/solve.py)
# ---------------------------------------------------------------------------

def init(shape) -> State:
    """cardiax/solve.py:18-23: v = w = 1, u = 0 (float32)."""
    dev = require_gpu("init")
    shape = tuple(int(s) for s in shape)
    v, w, u = (torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(3))
    _lib.check(_lib.lib().fk_init(v.data_ptr(), w.data_ptr(), u.data_ptr(), v.numel(),
                                  stream_handle()))
    return State(v, w, u)


def step(state, t, params, diffusivity, stimuli, dx) -> State:
    """cardiax/solve.py:26-65: the time derivatives (d_v, d_w, d_u) at time t."""
    dev = require_gpu("step")
    st = _dev_state(state, dev)
    H, W = _shape_of(st)
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    stim = _Stimuli(stimuli, dev, (H, W))
    out = State(*(torch.empty_like(st.u) for _ in range(3)))
    p = _lib.make_params(params)
    _lib.check(_lib.lib().fk_step(ctypes.byref(p), D.data_ptr(), st.v.data_ptr(),
                                  st.w.data_ptr(), st.u.data_ptr(), H, W, stim.pointer(),
                                  stim.n, _scalar(t), float(dx), out.v.data_ptr(),
                                  out.w.data_ptr(), out.u.data_ptr(), stream_handle()))
    return out


def step_euler(state, t, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:68-70: x + dt * step(x)."""
    dev = require_gpu("step_euler")
    st = _dev_state(state, dev)
    H, W = _shape_of(st)
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    stim = _Stimuli(stimuli, dev, (H, W))
    out = State(*(torch.empty_like(st.u) for _ in range(3)))
    p = _lib.make_params(params)
    _lib.check(_lib.lib().fk_step_euler(
        ctypes.byref(p), D.data_ptr(), st.v.data_ptr(), st.w.data_ptr(), st.u.data_ptr(), H,
        W, stim.pointer(), stim.n, _scalar(t), float(dt), float(dx), out.v.data_ptr(),
        out.w.data_ptr(), out.u.data_ptr(), stream_handle()))
    return out


def step_heun(state, t, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:73-85: k1 = step(x, t), k2 = step(x + dt k1, t),
    x + (k1 + k2) dt / 2, as two fused launches (fk_step_heun)."""
    dev = require_gpu("step_heun")
    st = _dev_state(state, dev)
    H, W = _shape_of(st)
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    stim = _Stimuli(stimuli, dev, (H, W))
    out = State(*(torch.empty_like(st.u) for _ in range(3)))
    L = _lib.lib()
    nbytes = L.fk_heun_workspace_bytes(H, W)
    work = torch.empty(max(1, nbytes // 4), dtype=torch.float32, device=dev)
    p = _lib.make_params(params)
    _lib.check(L.fk_step_heun(
        ctypes.byref(p), D.data_ptr(), st.v.data_ptr(), st.w.data_ptr(), st.u.data_ptr(), H,
        W, stim.pointer(), stim.n, _scalar(t), float(dt), float(dx), out.v.data_ptr(),
        out.w.data_ptr(), out.u.data_ptr(), work.data_ptr(), nbytes, stream_handle()))
    return out


def _forward_impl(entry, state, t, t_end, params, diffusivity, stimuli, dt, dx,
                  workspace="fk_forward_workspace_bytes") -> State:
    dev = require_gpu(entry)
    st = _dev_state(state, dev)
    H, W = _shape_of(st)
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    stim = _Stimuli(stimuli, dev, (H, W))
    out = State(*(torch.empty_like(st.u) for _ in range(3)))
    L = _lib.lib()
    nbytes = getattr(L, workspace)(H, W)
    work = torch.empty(max(1, nbytes // 4), dtype=torch.float32, device=dev)
    p = _lib.make_params(params)
    _lib.check(getattr(L, entry)(
        ctypes.byref(p), D.data_ptr(), st.v.data_ptr(), st.w.data_ptr(), st.u.data_ptr(), H,
        W, stim.pointer(), stim.n, _scalar(t), _scalar(t_end), float(dt), float(dx),
        out.v.data_ptr(), out.w.data_ptr(), out.u.data_ptr(), work.data_ptr(), nbytes,
        stream_handle()))
    return out


def _forward_euler(state, t, t_end, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:92-100: Euler steps for i = t, t+1, ... while i < t_end;
    step i uses stimulus time i. Returns the final state."""
    return _forward_impl("fk_forward_euler", state, t, t_end, params, diffusivity, stimuli,
                         dt, dx)


def _forward_euler_plain(state, t, t_end, params, diffusivity, stimuli, dt, dx) -> State:
    """Same contract as ``_forward_euler`` on the unoptimised one-thread-per-
    cell kernel; the fp32 baseline the tuned kernel is measured against."""
    return _forward_impl("fk_forward_euler_plain", state, t, t_end, params, diffusivity,
                         stimuli, dt, dx)


def _unsupported_integrator(name):
    def f(*args, **kwargs):
        raise NotImplementedError(
            "%s integration is not implemented: this package provides the forward-Euler "
            "and Heun integrators" % name)
    f.__name__ = "_forward_" + name.lower()
    return f


def _forward_heun(state, t, t_end, params, diffusivity, stimuli, dt, dx) -> State:
    """cardiax/solve.py:103-111: Heun steps for i = t, t+1, ... while i < t_end;
    both stages of step i use stimulus time i."""
    return _forward_impl("fk_forward_heun", state, t, t_end, params, diffusivity, stimuli,
                         dt, dx, workspace="fk_heun_workspace_bytes")


_forward_dormandprince = _unsupported_integrator("DORMANDPRINCE")


def step_rk(state, t, params, diffusivity, stimuli, dt, dx):
    """cardiax/solve.py:86-87 integrates with jax.experimental.ode.odeint
    (adaptive Dormand-Prince), which this package does not implement; raises
    NotImplementedError, as TimeIntegrator.DORMANDPRINCE does."""
    raise NotImplementedError("step_rk (jax.experimental.ode.odeint) is not implemented: this "
                              "package provides the forward-Euler and Heun integrators")


class TimeIntegrator:
    """cardiax/solve.py:127-130. As in the reference, each attribute *is* the
    integrator callable, so ``forward(..., integrator=TimeIntegrator.EULER)``
    calls it directly and any callable with the same signature plugs in."""
    EULER = staticmethod(_forward_euler)
    HEUN = staticmethod(_forward_heun)
    DORMANDPRINCE = staticmethod(_forward_dormandprince)


def forward(state, checkpoints, params, diffusivity, stimuli, dt, dx,
            integrator=TimeIntegrator.EULER, plot_while=False):
    """cardiax/solve.py:168-222: run each checkpoint interval with
    ``integrator`` and return the list of states at checkpoints[1:] (the
    initial state is not included)."""
    if plot_while:
        raise NotImplementedError("plot_while: plotting is not part of this package")
    f = integrator
    if f is TimeIntegrator.DORMANDPRINCE:
        return f(state, checkpoints, params, diffusivity, stimuli, dt, dx)
    cps = [_scalar(c) for c in (checkpoints.tolist() if hasattr(checkpoints, "tolist")
                                else checkpoints)]
    dev = require_gpu("forward")
    st = _dev_state(state, dev)
    H, W = _shape_of(st)
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    # Move stimulus fields once for all segments.
    stimuli = [type(s)(s[0], _dev_field(s[1], dev, (H, W), "stimulus field"))
               if hasattr(s, "_fields") else (s[0], _dev_field(s[1], dev, (H, W), "field"))
               for s in (stimuli or [])]
    states = []
    for i in range(len(cps) - 1):
        print("Solving at: %dms/%dms\t\t with %d passages"
              % (cps[i + 1] * dt, cps[-1] * dt, cps[i + 1] - cps[i]), end="\r")
        st = f(st, cps[i], cps[i + 1], params, D, stimuli, dt, dx)
        states.append(st)
    return states


def forward_dimensional(tissue_size, final_time, ms_step, params, diffusivity, stimuli, dt,
                        dx, integrator=TimeIntegrator.EULER, plot_while=False):
    """cardiax/solve.py:133-165: sizes in cm and times in ms, truncated to
    grid units with ``int()``; asserts the D-map and stimulus shapes."""
    shape = convert.realsize_to_shape(tissue_size, dx)
    start = 0
    stop = convert.ms_to_units(final_time, dt)
    stp = convert.ms_to_units(ms_step, dt)
    assert tuple(shape) == tuple(diffusivity.shape)
    assert all([tuple(s.field.shape) == tuple(shape) for s in stimuli])
    state = init(shape)
    checkpoints = list(range(start, stop, stp))
    return forward(state, checkpoints, params, diffusivity, stimuli, dt, dx, integrator,
                   plot_while)


def check_finite(state) -> None:
    """Raise FloatingPointError when any field of ``state`` (on the GPU)
    holds NaN or inf (fk_check_finite; synchronises the stream). No
    reference counterpart: JAX carries NaN through silently."""
    dev = require_gpu("check_finite")
    fields = [on_device(x, dev, "field") for x in state]
    n = fields[0].numel()
    if any(f.numel() != n for f in fields):
        raise ValueError("check_finite: fields differ in size")
    ptrs = (ctypes.c_void_p * len(fields))(*[f.data_ptr() for f in fields])
    flag = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().fk_check_finite(ptrs, len(fields), n, flag.data_ptr(),
                                          stream_handle()))


def gradient(a, axis):
    """cardiax/solve.py:225-254: first derivative of a 2-D array along
    ``axis`` without spacing (4th-order central inside, 3rd-order one-sided on
    the two outermost entries of each side)."""
    dev = require_gpu("gradient")
    t = on_device(a, dev, "a")
    if t.dim() != 2:
        raise ValueError("gradient takes a 2-D array, got %d-D" % t.dim())
    axis = int(axis)
    if axis < 0:
        axis += 2
    out = torch.empty_like(t)
    _lib.check(_lib.lib().fk_gradient(t.data_ptr(), out.data_ptr(), t.shape[0], t.shape[1],
                                      axis, stream_handle()))
    return out


def stimulate(t, X, stimuli):
    """cardiax/solve.py:257-271: X with cells replaced by the fields of the
    stimuli active at time t (later stimuli win where their field is nonzero)."""
    dev = require_gpu("stimulate")
    x = on_device(X, dev, "X")
    stim = _Stimuli(stimuli, dev, tuple(x.shape))
    out = torch.empty_like(x)
    _lib.check(_lib.lib().fk_stimulate(_scalar(t), x.data_ptr(), out.data_ptr(), x.numel(),
                                       stim.pointer(), stim.n, stream_handle()))
    return out
