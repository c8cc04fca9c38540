"""CPU tests of the C-ABI boundary and the host logic (no GPU compute).

The native library must load and export every function include/cardiax_amd.h
declares; argument errors must come back as FK_ERR_ARG with a message before
any device work; and the Python layer must refuse to run without a GPU
instead of falling back to a CPU path.
"""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from cardiax_amd import _lib, solve, stimulus
from cardiax_amd.params import PARAMSET_3, Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "cardiax_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _header_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), name
    # the ctypes binding covers exactly the declared entry points
    assert set(_lib.exported_symbols()) == set(declared)
    assert L.fk_abi_version() >= 1


def test_workspace_query():
    L = _lib.lib()
    # ping-pong state (3 fields, 256-byte aligned) + one float4 of coefficients per cell
    assert L.fk_forward_workspace_bytes(128, 64) == 3 * 128 * 64 * 4 + 16 * 128 * 64
    assert L.fk_forward_workspace_bytes(5, 7) == 512 + 16 * 35
    assert L.fk_forward_workspace_bytes(0, 5) == 0


def test_default_variant_by_grid_size():
    """Host logic only: which step kernel an automatic solve of a shape runs."""
    from cardiax_amd import autotune

    L = _lib.lib()
    prev = L.fk_set_forward_variant(0)
    try:
        assert autotune.resolved((1024, 1024)) == "stream2"  # up to 1.1 M cells
        assert autotune.resolved((1200, 1200)) == "streamt8"  # two steps per launch
        assert autotune.resolved((2048, 2048)) == "streamt16"
        assert autotune.resolved((4096, 4096)) == "streamt16"
        assert autotune.resolved((40000, 100)) == "stream4"  # too narrow for two steps
        assert autotune.resolved((100000, 100)) == "tile16"
        assert autotune.resolved((20000, 64)) == "streamt8"
        L.fk_set_forward_variant(_lib.VARIANTS["streamt12"])
        assert autotune.resolved((1024, 1024)) == "streamt12"
        assert autotune.resolved((39, 1024)) == "stream2"  # below 2R + 16 rows
        assert autotune.resolved((1024, 100)) == "stream2"  # 97..103 columns do not fit
        assert autotune.resolved((1024, 55)) == "stream2"   # below 56 columns
        assert autotune.resolved((1024, 64)) == "streamt12"  # two strips
        assert autotune.resolved((1024, 104)) == "streamt12"
        L.fk_set_forward_variant(_lib.VARIANTS["pair2"])
        assert autotune.resolved((64, 64)) == "pair2"
        assert autotune.resolved((64, 63)) == "stream2"      # odd W: the one-column kernel
        L.fk_set_forward_variant(_lib.VARIANTS["pair4"])
        assert autotune.resolved((64, 63)) == "stream4"
        with pytest.raises(ValueError):
            autotune.resolved((2, 64))
    finally:
        L.fk_set_forward_variant(prev)


def _call_step(H=8, W=8, params=None, n_stim=0, stim=None, dx=0.01, null_D=False):
    L = _lib.lib()
    p = _lib.make_params(params or PARAMSET_3)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    D = None if null_D else fake
    return L.fk_step(ctypes.byref(p), D, fake, fake, fake, H, W, stim, n_stim, 0.0, dx,
                     fake, fake, fake, None)


def test_argument_errors_are_reported_before_any_device_work():
    L = _lib.lib()
    assert _call_step(H=2) == _lib.FK_ERR_ARG
    assert b"too small" in L.fk_last_error()
    assert _call_step(null_D=True) == _lib.FK_ERR_ARG
    assert b"null" in L.fk_last_error()
    assert _call_step(dx=0.0) == _lib.FK_ERR_ARG
    bad = Params(*([float("nan")] + list(PARAMSET_3[1:])))
    assert _call_step(params=bad) == _lib.FK_ERR_ARG
    assert _call_step(n_stim=33, stim=(_lib.FkStimulus * 33)()) == _lib.FK_ERR_UNSUPPORTED
    s = (_lib.FkStimulus * 1)()
    s[0] = _lib.FkStimulus(0.0, 2.0, 1e9, None)
    assert _call_step(n_stim=1, stim=s) == _lib.FK_ERR_ARG
    assert b"null field" in L.fk_last_error()
    assert L.fk_gradient(ctypes.c_void_p(16), ctypes.c_void_p(16), 4, 9, 0, None) == \
        _lib.FK_ERR_ARG
    assert L.fk_gradient(ctypes.c_void_p(16), ctypes.c_void_p(16), 9, 9, 2, None) == \
        _lib.FK_ERR_ARG
    fake = ctypes.c_void_p(16)
    p = _lib.make_params(PARAMSET_3)
    rc = L.fk_forward_euler(ctypes.byref(p), fake, fake, fake, fake, 8, 8, None, 0, 0.0,
                            float("inf"), 0.01, 0.01, fake, fake, fake, fake, 0, None)
    assert rc == _lib.FK_ERR_ARG
    rc = L.fk_forward_euler(ctypes.byref(p), fake, fake, fake, fake, 8, 8, None, 0, 0.0,
                            10.0, 0.01, 0.01, fake, fake, fake, None, 0, None)
    assert rc == _lib.FK_ERR_ARG and b"workspace" in L.fk_last_error()


def test_time_indices_bounded_by_float32_counter():
    L = _lib.lib()
    p = _lib.make_params(PARAMSET_3)
    fake = ctypes.c_void_p(0x1000)

    def fwd(t, t_end):
        return L.fk_forward_euler(ctypes.byref(p), fake, fake, fake, fake, 8, 8, None, 0, t,
                                  t_end, 0.01, 0.01, fake, fake, fake, fake, 1 << 20, None)

    assert fwd(0.0, 2.0 ** 24 + 1) == _lib.FK_ERR_ARG
    assert b"2^24" in L.fk_last_error()
    assert fwd(2.0 ** 24, 2.0 ** 24 + 5) == _lib.FK_ERR_ARG
    assert fwd(-2.0 ** 24, 0.0) == _lib.FK_ERR_ARG
    assert L.fk_step(ctypes.byref(p), fake, fake, fake, fake, 8, 8, None, 0, 3e7, 0.01, fake,
                     fake, fake, None) == _lib.FK_ERR_ARG


def test_check_finite_arguments():
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    ptrs = (ctypes.c_void_p * 2)(0x1000, None)
    assert L.fk_check_finite(ptrs, 2, 10, fake, None) == _lib.FK_ERR_ARG
    assert b"null" in L.fk_last_error()
    assert L.fk_check_finite(ptrs, 0, 10, fake, None) == _lib.FK_ERR_ARG
    assert L.fk_check_finite(None, 1, 10, fake, None) == _lib.FK_ERR_ARG
    assert L.fk_check_finite(ptrs, 1, -1, fake, None) == _lib.FK_ERR_ARG
    assert L.fk_check_finite(ptrs, 1, 10, None, None) == _lib.FK_ERR_ARG
    with pytest.raises(FloatingPointError):
        _lib.check(_lib.FK_ERR_NONFINITE)


def test_io_and_stimulus_argument_errors():
    """Overlapping buffers, small workspaces and bad boxes are refused before
    any launch (the pointers are never dereferenced)."""
    L = _lib.lib()
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x1000 + 4 * 8 * 8 - 4)
    far = ctypes.c_void_p(0x100000)
    m = (ctypes.c_double * 4)(1.0, 0.0, 0.0, 1.0)
    off = (ctypes.c_double * 2)(0.0, 0.0)
    # 8 x 8 floats from 0x1000 overlap a dst starting one float before their end
    assert L.fk_affine_transform(a, 8, 8, m, off, b, None) == _lib.FK_ERR_ARG
    assert b"overlap" in L.fk_last_error()
    assert L.fk_affine_transform(b, 8, 8, m, off, a, None) == _lib.FK_ERR_ARG
    bad = (ctypes.c_double * 4)(float("nan"), 0.0, 0.0, 1.0)
    assert L.fk_affine_transform(a, 8, 8, bad, off, far, None) == _lib.FK_ERR_ARG
    need = L.fk_resize_workspace_bytes(1, 8, 8, 4, 4)
    assert need == 4 * 8 * 8
    assert L.fk_resize(a, 1, 8, 8, far, 4, 4, far, need - 1, None) == _lib.FK_ERR_ARG
    assert b"workspace" in L.fk_last_error()
    assert L.fk_resize(a, 1, 8, 8, b, 4, 4, far, need, None) == _lib.FK_ERR_ARG
    assert b"overlap" in L.fk_last_error()
    assert L.fk_resize(a, 0, 8, 8, b, 4, 4, None, 0, None) == _lib.FK_OK  # nothing to do
    assert L.fk_stimulus_box(a, 8, 8, 0, 9, 0, 4, 1.0, None) == _lib.FK_ERR_ARG
    assert L.fk_stimulus_box(a, 8, 8, -1, 4, 0, 4, 1.0, None) == _lib.FK_ERR_ARG


def test_check_maps_codes_to_python_errors():
    _call_step(H=1)
    with pytest.raises(ValueError):
        _lib.check(_lib.FK_ERR_ARG)
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.FK_ERR_UNSUPPORTED)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.FK_ERR_HIP)
    _lib.check(_lib.FK_OK)
    with pytest.raises(TypeError):
        _lib.make_params((1.0, 2.0))


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU behaviour")
def test_no_cpu_fallback():
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        solve.init((8, 8))
    st = (np.ones((8, 8)), np.ones((8, 8)), np.zeros((8, 8)))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        solve._forward_euler(st, 0, 10, PARAMSET_3, np.full((8, 8), 1e-3), [], 0.01, 0.01)
    with pytest.raises(RuntimeError):
        solve.gradient(np.ones((8, 8)), 0)


def test_time_integrator_members_are_the_callables():
    assert solve.TimeIntegrator.EULER is solve._forward_euler
    assert solve.TimeIntegrator.HEUN is solve._forward_heun
    st = (np.ones((8, 8)), np.ones((8, 8)), np.zeros((8, 8)))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        solve.TimeIntegrator.HEUN(st, 0, 10, PARAMSET_3, np.full((8, 8), 1e-3), [], 0.01, 0.01)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        solve.step_heun(st, 0, PARAMSET_3, np.full((8, 8), 1e-3), [], 0.01, 0.01)
    with pytest.raises(NotImplementedError):
        solve.TimeIntegrator.DORMANDPRINCE(None, [0, 1], None, None, None, 0.01, 0.01)
    with pytest.raises(NotImplementedError):
        solve.step_rk(st, 0, PARAMSET_3, np.full((8, 8), 1e-3), [], 0.01, 0.01)


def test_public_surface_matches_reference_names():
    for name in ("State", "init", "step", "step_euler", "step_heun", "_forward_euler",
                 "_forward_heun", "TimeIntegrator",
                 "forward", "forward_dimensional", "gradient", "stimulate", "step_rk"):
        assert hasattr(solve, name), name
    assert solve.State._fields == ("v", "w", "u")
    assert stimulus.Protocol._fields == ("start", "duration", "period")
    assert stimulus.Stimulus._fields == ("protocol", "field")
    assert Params._fields[0] == "tau_v_plus" and Params._fields[-1] == "Cm"
    assert len(Params._fields) == 14
