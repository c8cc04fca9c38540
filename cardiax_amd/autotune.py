"""Autotuner of the forward step kernel (fk_autotune_forward).

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
