"""Autotuner of the forward step kernel (fk_autotune_forward,
fk_autotune_forward_state).

The reference leaves kernel choice to XLA. Here the step kernel has several
implementations (LDS tiles, streaming wave bands of 8/16/32 rows, temporal
blocking of 2 or 4 steps on 32^2 / 64^2 tiles); ``tune(shape)`` times each on
the GPU for one grid shape and records the fastest, which every automatic
launch on that shape uses afterwards (process-wide, guarded in the library).
"""
from __future__ import annotations

import ctypes
from typing import Dict

from . import _lib
from ._device import require_gpu, stream_handle

NAMES = {v: k for k, v in _lib.VARIANTS.items()}


def tune(shape, n_steps: int = 200) -> Dict[str, object]:
    """Time every variant over ``n_steps`` Euler steps of a ``shape`` grid and
    make the fastest the automatic choice for that shape. Returns
    ``{"best": name, "times_ms": {name: ms}}``."""
    require_gpu("autotune")
    H, W = (int(x) for x in shape)
    L = _lib.lib()
    times = (ctypes.c_float * len(_lib.VARIANTS))()
    best = ctypes.c_int32(0)
    _lib.check(L.fk_autotune_forward(H, W, int(n_steps), stream_handle(), times,
                                     ctypes.byref(best)))
    return {"best": NAMES[best.value],
            "times_ms": {NAMES[i]: round(float(t), 4) for i, t in enumerate(times) if t >= 0}}


def tune_on(state, t, params, diffusivity, stimuli, dt, dx, n_steps: int = 200
            ) -> Dict[str, object]:
    """Time every variant over ``n_steps`` Euler steps of the caller's own
    problem (the arguments of ``_forward_euler`` at time ``t``, which are not
    modified; fk_autotune_forward_state) and make the fastest the automatic
    choice for its shape. Kernels whose gates branch cost more where waves
    hold excited or mixed cells, so a state from the run to be made picks
    the kernel for that run. Returns ``{"best": name, "times_ms": {...}}``."""
    from .solve import State, _Stimuli, _dev_field, _scalar

    dev = require_gpu("autotune")
    st = State(*(_dev_field(x, dev, tuple(state[2].shape), "state") for x in state))
    H, W = st.u.shape
    D = _dev_field(diffusivity, dev, (H, W), "diffusivity")
    stim = _Stimuli(stimuli, dev, (H, W))
    p = _lib.make_params(params)
    L = _lib.lib()
    times = (ctypes.c_float * len(_lib.VARIANTS))()
    best = ctypes.c_int32(0)
    _lib.check(L.fk_autotune_forward_state(
        ctypes.byref(p), D.data_ptr(), st.v.data_ptr(), st.w.data_ptr(), st.u.data_ptr(), H, W,
        stim.pointer(), stim.n, float(_scalar(t)), int(n_steps), float(dt), float(dx),
        stream_handle(), times, ctypes.byref(best)))
    return {"best": NAMES[best.value],
            "times_ms": {NAMES[i]: round(float(x), 4) for i, x in enumerate(times) if x >= 0}}


def lookup(shape) -> str:
    """Name of the variant tuned for ``shape`` ("auto" when not tuned)."""
    H, W = (int(x) for x in shape)
    return NAMES[int(_lib.lib().fk_autotune_lookup(H, W))]


def resolved(shape) -> str:
    """Name of the variant that forward solves of ``shape`` run now: the
    process-wide setting, else the tuned one, else the size-based default."""
    H, W = (int(x) for x in shape)
    v = int(_lib.lib().fk_forward_variant_for(H, W))
    if v < 0:
        _lib.check(v)
    return NAMES[v]


def reset() -> None:
    """Forget every tuned shape."""
    _lib.lib().fk_autotune_reset()
