"""fk_forward_euler_noalloc against fk_forward_euler, through the C-ABI.

The no-allocation entry point keeps its ping-pong state in a fixed device
array; its results must equal fk_forward_euler's bit for bit on the same
inputs and kernel variant. Two calls run back to back, the second on a
smaller grid (the array keeps the first call's contents), then the largest
accepted grid and the refusal of a larger one.
"""
import ctypes

import numpy as np
import pytest
import torch

from cardiax_amd import _lib, workloads as wl
from cardiax_amd.params import PARAMSET_3

pytestmark = pytest.mark.gpu


def _inputs(shape, seed, dev):
    rng = np.random.default_rng(seed)
    H, W = shape
    D = torch.from_numpy(wl.scar_map(shape, rng).astype(np.float32)).to(dev)
    v = torch.from_numpy(rng.uniform(0.2, 1.0, shape).astype(np.float32)).to(dev)
    w = torch.from_numpy(rng.uniform(0.2, 1.0, shape).astype(np.float32)).to(dev)
    u = torch.from_numpy(rng.uniform(-0.05, 1.05, shape).astype(np.float32)).to(dev)
    f1 = torch.zeros(shape, dtype=torch.float32, device=dev)
    f1[:, : max(1, W // 10)] = 20.0
    f2 = torch.zeros(shape, dtype=torch.float32, device=dev)
    f2[H // 2:, W // 3: W // 3 + max(1, W // 4)] = 20.0
    stims = (_lib.FkStimulus * 2)(_lib.FkStimulus(0.0, 2.0, 1e9, f1.data_ptr()),
                                  _lib.FkStimulus(5.0, 2.0, 9.0, f2.data_ptr()))
    return D, (v, w, u), (f1, f2), stims


def _run(fn_name, shape, seed, t0, t1, dev, with_work):
    L = _lib.lib()
    D, x, keep, stims = _inputs(shape, seed, dev)
    H, W = shape
    out = [torch.full(shape, float("nan"), dtype=torch.float32, device=dev) for _ in range(3)]
    p = _lib.make_params(PARAMSET_3)
    args = [ctypes.byref(p), D.data_ptr(), x[0].data_ptr(), x[1].data_ptr(), x[2].data_ptr(), H,
            W, ctypes.cast(stims, ctypes.POINTER(_lib.FkStimulus)), 2, float(t0), float(t1),
            0.01, 0.01, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr()]
    s = torch.cuda.current_stream().cuda_stream
    if with_work:
        nb = L.fk_forward_workspace_bytes(H, W)
        work = torch.empty(nb // 4 + 1, dtype=torch.float32, device=dev)
        _lib.check(getattr(L, fn_name)(*args, work.data_ptr(), nb, s))
    else:
        _lib.check(getattr(L, fn_name)(*args, s))
    torch.cuda.synchronize()
    del keep
    return [o.cpu().numpy() for o in out]


def _compare(shape, seed, t0, t1, dev):
    print("no-alloc vs fk_forward_euler: %dx%d, steps %d..%d, seed %d"
          % (shape[0], shape[1], t0, t1, seed), flush=True)
    ref = _run("fk_forward_euler", shape, seed, t0, t1, dev, True)
    got = _run("fk_forward_euler_noalloc", shape, seed, t0, t1, dev, False)
    for name, a, b in zip("vwu", got, ref):
        assert np.isfinite(b).all(), name
        np.testing.assert_array_equal(a, b, err_msg="field %s differs" % name)


@pytest.fixture
def pinned(gpu):
    """Pin the stream2 kernel (the default for these sizes) for both paths."""
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS["stream2"])
    yield
    L.fk_set_forward_variant(prev)


def test_two_calls_second_smaller(gpu, pinned):
    _compare((1024, 1024), 1, 0, 40, gpu)  # the largest grid accepted
    _compare((96, 80), 2, 3, 34, gpu)      # a smaller one: odd step count


@pytest.mark.parametrize("variant", ["stream4", "tb64x2", "pair2", "streamt8"])
def test_variants(gpu, variant):
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS[variant])
    try:
        _compare((200, 136), 3, 0, 23, gpu)
        _compare((64, 48), 4, 0, 10, gpu)
    finally:
        L.fk_set_forward_variant(prev)


def test_coefficient_variant_runs_its_streaming_counterpart(gpu):
    # streamc2 keeps precomputed coefficients in the workspace; the no-alloc
    # path runs stream2 instead, so its result is stream2's
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS["streamc2"])
    try:
        got = _run("fk_forward_euler_noalloc", (130, 100), 5, 0, 17, gpu, False)
    finally:
        L.fk_set_forward_variant(prev)
    prev = L.fk_set_forward_variant(_lib.VARIANTS["stream2"])
    try:
        ref = _run("fk_forward_euler", (130, 100), 5, 0, 17, gpu, True)
    finally:
        L.fk_set_forward_variant(prev)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_single_and_zero_steps(gpu, pinned):
    _compare((50, 70), 6, 4, 5, gpu)
    _compare((50, 70), 7, 4, 4, gpu)


def test_too_large_is_refused(gpu):
    L = _lib.lib()
    H, W = 1025, 1024
    assert H * W > _lib.FK_NOALLOC_MAX_CELLS
    p = _lib.make_params(PARAMSET_3)
    # refused before any pointer is used: the arguments are never read
    dummy = torch.zeros(1, dtype=torch.float32, device=gpu)
    ptr = dummy.data_ptr()
    rc = L.fk_forward_euler_noalloc(ctypes.byref(p), ptr, ptr, ptr, ptr, H, W, None, 0, 0.0,
                                    10.0, 0.01, 0.01, ptr, ptr, ptr,
                                    torch.cuda.current_stream().cuda_stream)
    assert rc == _lib.FK_ERR_ARG
    assert b"above" in L.fk_last_error()
