"""ctypes binding of ``libcardiax_amd.so`` (declared in include/cardiax_amd.h).

torch is imported first so that the library's ``libamdhip64.so.7`` dependency
resolves to the HIP runtime torch already loaded: one runtime per process.
Loading fails loudly when the library has not been built; there is no
fallback path.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL below)

from .params import Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcardiax_amd.so")

FK_OK = 0
FK_ERR_ARG = -1
FK_ERR_HIP = -2
FK_ERR_UNSUPPORTED = -3
FK_ERR_NONFINITE = -4
FK_MAX_STIMULI = 32
FK_NOALLOC_MAX_CELLS = 1 << 20

# Forward kernel variants (fk_set_forward_variant).
VARIANTS = {"auto": 0, "plain": 1, "tile16": 2, "tile32": 3, "stream8": 4, "stream16": 5,
            "stream32": 6, "tb64x2": 7, "tb64x4": 8, "tb32x2": 9, "tb32x4": 10,
            "streamc8": 11, "streamc16": 12, "stream4": 13, "streamc4": 14,
            "stream2": 15, "streamc2": 16, "pair2": 17, "pair4": 18, "streamt8": 19,
            "streamt12": 20, "streamt16": 21, "streamt24": 22}

c_f = ctypes.c_float
c_d = ctypes.c_double
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p


class FkParams(ctypes.Structure):
    """struct fk_params -- the 14 Params fields in order (cardiax/params.py:4-18)."""
    _fields_ = [(name, c_f) for name in Params._fields]


class FkStimulus(ctypes.Structure):
    """struct fk_stimulus (cardiax/stimulus.py:13-21)."""
    _fields_ = [("start", c_d), ("duration", c_d), ("period", c_d), ("field", c_vp)]


FK_BATCH_MAX_STIMULI = 4
FK_BATCH_N_OPTIONS = 6
FK_BATCH_TWO_STEP = 256


class FkSim(ctypes.Structure):
    """struct fk_sim: one member of a batch."""
    _fields_ = [("params", FkParams), ("D", c_vp), ("v", c_vp), ("w", c_vp), ("u", c_vp),
                ("H", c_i32), ("W", c_i32), ("n_stimuli", c_i32),
                ("stimuli", FkStimulus * FK_BATCH_MAX_STIMULI), ("t", c_d), ("dt", c_f),
                ("dx", c_f)]


class FkRolloutSpec(ctypes.Structure):
    """struct fk_rollout_spec: one problem of a batched rollout."""
    _fields_ = [("params", FkParams), ("D", c_vp), ("v0", c_vp), ("w0", c_vp), ("u0", c_vp),
                ("H", c_i32), ("W", c_i32), ("n_stimuli", c_i32),
                ("stimuli", FkStimulus * FK_BATCH_MAX_STIMULI), ("t0", c_d), ("dt", c_f),
                ("dx", c_f), ("frames", c_vp), ("g_frames", c_vp), ("grad_D", c_vp)]


# name -> (restype, argtypes)
_SIGNATURES = {
    "fk_last_error": (ctypes.c_char_p, []),
    "fk_abi_version": (c_i32, []),
    "fk_set_forward_variant": (c_i32, [c_i32]),
    "fk_autotune_forward": (c_i32, [c_i32, c_i32, c_i32, c_vp, ctypes.POINTER(c_f),
                                    ctypes.POINTER(c_i32)]),
    "fk_autotune_forward_state": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp,
                                          c_i32, c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d,
                                          c_i32, c_f, c_f, c_vp, ctypes.POINTER(c_f),
                                          ctypes.POINTER(c_i32)]),
    "fk_autotune_lookup": (c_i32, [c_i32, c_i32]),
    "fk_forward_variant_for": (c_i32, [c_i32, c_i32]),
    "fk_autotune_reset": (None, []),
    "fk_init": (c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "fk_check_finite": (c_i32, [ctypes.POINTER(c_vp), c_i32, c_i64, c_vp, c_vp]),
    "fk_step": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                        ctypes.POINTER(FkStimulus), c_i32, c_d, c_f, c_vp, c_vp, c_vp, c_vp]),
    "fk_step_euler": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                              ctypes.POINTER(FkStimulus), c_i32, c_d, c_f, c_f, c_vp, c_vp,
                              c_vp, c_vp]),
    "fk_forward_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "fk_forward_euler": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32,
                                 c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d, c_d, c_f, c_f,
                                 c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "fk_forward_euler_noalloc": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp,
                                         c_i32, c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d,
                                         c_d, c_f, c_f, c_vp, c_vp, c_vp, c_vp]),
    "fk_forward_euler_plain": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp,
                                       c_i32, c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d,
                                       c_d, c_f, c_f, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "fk_forward_trajectory": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32,
                                      c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d, c_i32,
                                      c_f, c_f, c_vp, c_vp]),
    "fk_heun_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "fk_resize_workspace_bytes": (c_sz, [c_i32, c_i32, c_i32, c_i32, c_i32]),
    "fk_resize": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_vp, c_i32, c_i32, c_vp, c_sz, c_vp]),
    "fk_stimulus_box": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_f, c_vp]),
    "fk_affine_transform": (c_i32, [c_vp, c_i32, c_i32, ctypes.POINTER(c_d),
                                    ctypes.POINTER(c_d), c_vp, c_vp]),
    "fk_step_heun": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                             ctypes.POINTER(FkStimulus), c_i32, c_d, c_f, c_f, c_vp, c_vp,
                             c_vp, c_vp, c_sz, c_vp]),
    "fk_forward_heun": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32,
                                c_i32, ctypes.POINTER(FkStimulus), c_i32, c_d, c_d, c_f, c_f,
                                c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "fk_gradient": (c_i32, [c_vp, c_vp, c_i32, c_i32, c_i32, c_vp]),
    "fk_stimulate": (c_i32, [c_d, c_vp, c_vp, c_i64, ctypes.POINTER(FkStimulus), c_i32, c_vp]),
    "fk_adjoint_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "fk_adjoint_prepare": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_sz, c_vp]),
    "fk_adjoint_step": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                                ctypes.POINTER(FkStimulus), c_i32, c_d, c_f, c_f, c_vp, c_vp,
                                c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "fk_adjoint_segment": (c_i32, [ctypes.POINTER(FkParams), c_vp, c_i32, c_i32,
                                   ctypes.POINTER(FkStimulus), c_i32, c_d, c_i32, c_f, c_f,
                                   c_vp, ctypes.POINTER(c_vp), c_vp, c_vp, c_vp, c_vp, c_vp,
                                   c_sz, c_vp]),
    "fk_adjoint_tau_grads": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_vp]),
    "fk_rasterize_workspace_bytes": (c_sz, [c_i32, c_i32]),
    "fk_rasterize_polygons": (c_i32, [c_vp, c_i32, c_i32, ctypes.POINTER(c_d),
                                      ctypes.POINTER(c_d), ctypes.POINTER(c_i32), c_i32, c_vp,
                                      c_sz, c_vp]),
    "fk_convolve_same": (c_i32, [c_vp, c_i32, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_sz,
                                 c_vp]),
    "fk_convolve_workspace_bytes": (c_sz, [c_i32, c_i32, c_i32, c_i32]),
    "fk_scar_to_diffusivity": (c_i32, [c_vp, c_i64, c_d, c_d, c_vp, c_vp, c_sz, c_vp]),
    "fk_batch_create": (c_i32, [ctypes.POINTER(FkSim), c_i32, ctypes.POINTER(c_vp)]),
    "fk_rollout_batch_create": (c_i32, [ctypes.POINTER(FkRolloutSpec), c_i32, c_i32, c_i32,
                                        c_i32, ctypes.POINTER(c_vp)]),
    "fk_rollout_batch_device_bytes": (c_sz, [c_vp]),
    "fk_rollout_batch_forward": (c_i32, [c_vp, c_vp]),
    "fk_rollout_batch_backward": (c_i32, [c_vp, c_vp, c_vp]),
    "fk_rollout_batch_destroy": (c_i32, [c_vp]),
    "fk_batch_run": (c_i32, [c_vp, c_i64, c_i32, c_vp]),
    "fk_batch_autotune": (c_i32, [c_vp, c_i32, c_vp, ctypes.POINTER(c_f),
                                  ctypes.POINTER(c_i32)]),
    "fk_batch_set_rows": (c_i32, [c_vp, c_i32]),
    "fk_batch_n_items": (c_i32, [c_vp]),
    "fk_batch_destroy": (c_i32, [c_vp]),
}

_lib = None


def exported_symbols():
    """Names of every entry point this binding declares."""
    return sorted(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load (once) and return the native library; RuntimeError if missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "native library %s is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise the Python error matching a nonzero C-ABI return code."""
    if rc == FK_OK:
        return
    msg = lib().fk_last_error().decode("utf-8", "replace")
    if rc == FK_ERR_ARG:
        raise ValueError(msg)
    if rc == FK_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == FK_ERR_NONFINITE:
        raise FloatingPointError(msg)
    raise RuntimeError("HIP error: " + msg)


def make_params(params) -> FkParams:
    vals = tuple(params)
    if len(vals) != 14:
        raise TypeError("params must have the 14 Params fields, got %d values" % len(vals))
    return FkParams(*[float(x) for x in vals])


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()
