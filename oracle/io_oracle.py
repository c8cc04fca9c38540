"""CPU restatement of the dataset generator's output: the resize applied to
every checkpoint and the file layout it is written in.

TEST INFRASTRUCTURE ONLY. Only ``tests/`` may import this module, to check
results; the product path (``cardiax_amd``) never does.

Reference behaviour followed (file:line under the reference tree):
  * ``imresize``   cardiax/io.py:118-124: ``jax.image.resize(a, a.shape[:-2] +
    size, "bilinear")``. jax is a third-party dependency absent from the
    reference tree (unpinned; ``install_jax.sh:2`` pins jaxlib 0.1.64). Its
    ``resize`` defaults to ``antialias=True`` and then runs JAX's published
    ``scale_and_translate`` with translation 0 on every axis whose size
    changes: output o of an axis n_in -> n_out samples the input at
    s = (o + 1/2) n_in / n_out - 1/2 with the triangle kernel
    max(0, 1 - |s - j| / k), k = max(n_in / n_out, 1); the weights are
    divided by their sum (set to 0 when the sum is <= 1000 float32 eps) and
    zeroed where s lies outside [-1/2, n_in - 1/2]. Restated below in FP64.
    Pinned by torch's antialiased bilinear ``interpolate`` (the same
    PIL-style triangle filter) and by closed forms in
    ``tests/test_oracle_io.py``.
  * ``frames``     deepx/generate.py:173-207: the state after each checkpoint
    segment [c_i, c_{i+1}) of ``checkpoints = arange(start, stop, step)``,
    resized to the output shape; ``len(checkpoints) - 1`` frames.
  * ``layout``     cardiax/io.py:13-64 with deepx/generate.py:167-172,207: the
    datasets the generator writes and their shapes, including the frame slot
    that stays zero (the file holds ``len(checkpoints)`` frames, the loop
    writes one fewer) and the ``stimuli`` dataset that is created and never
    written.
"""
from __future__ import annotations

import numpy as np

from . import fk_oracle as o

_F32_EPS = float(np.finfo(np.float32).eps)


def weight_matrix(n_in: int, n_out: int) -> np.ndarray:
    """(n_in, n_out) resampling weights of one axis (JAX scale_and_translate,
    antialias=True, translation 0)."""
    inv_scale = n_in / n_out
    k = max(inv_scale, 1.0)
    s = (np.arange(n_out, dtype=np.float64) + 0.5) * inv_scale - 0.5
    x = np.abs(s[None, :] - np.arange(n_in, dtype=np.float64)[:, None]) / k
    w = np.maximum(0.0, 1.0 - x)
    tot = w.sum(axis=0, keepdims=True)
    w = np.where(np.abs(tot) > 1000.0 * _F32_EPS, w / np.where(tot != 0, tot, 1.0), 0.0)
    inside = (s >= -0.5) & (s <= n_in - 0.5)
    return np.where(inside[None, :], w, 0.0)


def imresize(a, size) -> np.ndarray:
    """cardiax/io.py:118-124 in FP64: the last two axes of ``a`` resampled to
    ``size``; leading axes untouched, an axis of unchanged size copied."""
    a = np.asarray(a, dtype=np.float64)
    h, w = (int(x) for x in size)
    H, W = a.shape[-2:]
    out = a
    if h != H:
        out = np.einsum("...iq,ih->...hq", out, weight_matrix(H, h))
    if w != W:
        out = np.einsum("...hq,qw->...hw", out, weight_matrix(W, w))
    return np.array(out)


def frames(start, stop, step, dt, dx, params, diffusivity, stimuli, reshape=None,
           forward_euler=o.forward_euler):
    """deepx/generate.py:173-207: resized states after each checkpoint
    segment, as a (len(checkpoints) - 1, 3, h, w) FP64 array."""
    D = np.asarray(diffusivity, dtype=np.float64)
    out_shape = tuple(reshape) if reshape is not None else D.shape
    cps = list(range(int(start), int(stop), int(step)))
    st = o.init(D.shape)
    res = []
    for i in range(len(cps) - 1):
        st = forward_euler(st, cps[i], cps[i + 1], params, D, stimuli, dt, dx)
        res.append(imresize(np.stack([st.v, st.w, st.u]), out_shape))
    return np.stack(res) if res else np.zeros((0, 3) + out_shape)


def layout(n_checkpoints: int, n_stimuli: int, out_shape, params_fields):
    """cardiax/io.py:13-64 as used by deepx/generate.py:167-172: dataset name
    -> shape of what the generator's file holds."""
    h, w = out_shape
    lay = {"states": (n_checkpoints, 3, h, w), "stimuli": (n_stimuli, h, w),
           "params/D": (h, w), "params/dt": (), "params/dx": (), "diffusivity": (h, w),
           "field": (n_stimuli, h, w), "start": (n_stimuli,), "duration": (n_stimuli,),
           "period": (n_stimuli,)}
    for f in params_fields:
        lay["params/" + f] = ()
    return lay
