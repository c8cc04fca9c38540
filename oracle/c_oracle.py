"""ctypes front end of oracle/fk_oracle.c (libfk_oracle.so).

TEST INFRASTRUCTURE ONLY (see fk_oracle.py): the FP64 build checks long
horizons; the float32 build is the CPU baseline bench.py times.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .fk_oracle import State

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfk_oracle.so")
_lib = None


def build() -> None:
    """Compile libfk_oracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-B", "libfk_oracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        for suf in ("f64", "f32"):
            f = getattr(L, "fk_oracle_forward_" + suf)
            f.restype = ctypes.c_int64
            f.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_int]
            g = getattr(L, "fk_oracle_scratch_bytes_" + suf)
            g.restype = ctypes.c_size_t
            g.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def forward_euler(state, t, t_end, params, diffusivity, stimuli, dt, dx, dtype=np.float64,
                  nthreads=0, gates=None, gate_mode=0) -> State:
    """cardiax/solve.py:92-100 in C. Returns a new State (inputs untouched).

    gates: a uint8 array of (steps, H, W) for the Heaviside gates of each
    cell-step (bit 0 = [u >= V_c], bit 1 = [u >= V_v]); gate_mode 1 records
    them there, 2 replays them in place of the comparisons (the fixed-mask
    function the adjoint differentiates), 0 ignores them."""
    dtype = np.dtype(dtype)
    suf = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}[dtype]
    v, w, u = (np.array(x, dtype=dtype, order="C", copy=True) for x in state)
    H, W = u.shape
    D = np.ascontiguousarray(diffusivity, dtype=dtype)
    n = len(stimuli)
    proto = np.zeros((max(n, 1), 3), np.float64)
    fields = np.zeros((max(n, 1), H, W), dtype)
    for i, s in enumerate(stimuli):
        proto[i] = [float(x) for x in s.protocol]
        fields[i] = np.asarray(s.field, dtype=dtype)
    L = lib()
    scratch = np.empty(getattr(L, "fk_oracle_scratch_bytes_" + suf)(H, W), np.uint8)
    params_arr = np.array([float(x) for x in params], np.float64)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    gp = None
    if gate_mode:
        steps = max(0, int(np.ceil(float(t_end) - float(t))))
        if (gates is None or gates.dtype != np.uint8 or not gates.flags.c_contiguous
                or gates.size < steps * H * W):
            raise ValueError("gates must be a C-contiguous uint8 array of >= %d x %d x %d"
                             % (steps, H, W))
        gp = ptr(gates)
    getattr(L, "fk_oracle_forward_" + suf)(
        ptr(params_arr), ptr(D), ptr(v), ptr(w), ptr(u), H, W, n, ptr(proto), ptr(fields),
        float(t), float(t_end), float(dt), float(dx), ptr(scratch), int(nthreads), gp,
        int(gate_mode))
    return State(v, w, u)
