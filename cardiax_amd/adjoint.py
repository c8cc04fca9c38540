"""Differentiable forward-Euler rollout: gradients of a loss on sampled u
frames with respect to the diffusivity map D and tau_d, tau_r, tau_si.

``rollout(...)`` runs cardiax/solve.py:92-100 on the GPU and returns the u
fields after every ``frame_every`` steps. Its backward pass is the
hand-written adjoint of the Euler step: the states are recomputed segment by
segment from checkpoints kept every ~sqrt(N) steps (fk_forward_trajectory),
so memory is O(sqrt(N)) states instead of N, and each segment is reversed by
one fk_adjoint_segment call. Gates are fixed masks and the
stimulus replacement routes no cotangent to j_ion -- what jax.grad computes
through the reference's greater_equal and where.
"""
from __future__ import annotations

import ctypes
import math
from typing import Sequence

import torch

from . import _lib
from ._device import on_device, require_gpu, stream_handle
from .params import Params
from .solve import State, _Stimuli, _scalar

# Params fields that carry gradients, in the order of the returned gradients.
GRAD_FIELDS = ("tau_d", "tau_r", "tau_si")


def _with_taus(params, tau_d, tau_r, tau_si) -> Params:
    p = Params(*[float(x) for x in params])
    return p._replace(tau_d=float(tau_d), tau_r=float(tau_r), tau_si=float(tau_si))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, tau_d, tau_r, tau_si, cfg):
        dev = D.device
        # host values of the taus: given floats need no device round trip
        hv = cfg["tau_host"]
        prm = _with_taus(cfg["params"], *(h if h is not None else t.item()
                                          for h, t in zip(hv, (tau_d, tau_r, tau_si))))
        st = State(*(on_device(x, dev).contiguous() for x in cfg["state0"]))
        H, W = st.u.shape
        n, fe, t0 = cfg["n_steps"], cfg["frame_every"], cfg["t0"]
        S = cfg["checkpoint_every"] or max(1, int(math.ceil(math.sqrt(n))))
        L = _lib.lib()
        stim = _Stimuli(cfg["stimuli"], dev, (H, W))
        p = _lib.make_params(prm)
        s = stream_handle()
        frames = torch.empty((n // fe, H, W), dtype=torch.float32, device=dev)
        # Each checkpoint interval runs through fk_forward_trajectory, the call
        # the backward pass uses to recompute it, so the backward linearises
        # around exactly these states (same kernels, same rounding), whatever
        # step kernel is set or tuned.
        seg = torch.empty((min(S, max(n, 1)), 3, H, W), dtype=torch.float32, device=dev)
        ckpts = {0: st}
        cur = st
        k = 0
        while k < n:
            e = min(n, k + S)
            _lib.check(L.fk_forward_trajectory(
                ctypes.byref(p), D.data_ptr(), cur.v.data_ptr(), cur.w.data_ptr(),
                cur.u.data_ptr(), H, W, stim.pointer(), stim.n, float(t0 + k), e - k,
                float(cfg["dt"]), float(cfg["dx"]), seg.data_ptr(), s))
            for j in range((k // fe + 1) * fe, e + 1, fe):  # x_j is seg[j - k - 1]
                frames[j // fe - 1].copy_(seg[j - k - 1, 2])
            last = seg[e - k - 1]
            cur, k = State(last[0].clone(), last[1].clone(), last[2].clone()), e
            if k < n:
                ckpts[k] = cur
        ctx.cfg = cfg
        ctx.prm = prm
        ctx.S = S
        ctx.ckpts = ckpts
        ctx.save_for_backward(D)
        return frames

    @staticmethod
    def backward(ctx, g_frames):
        (D,) = ctx.saved_tensors
        cfg, prm, S = ctx.cfg, ctx.prm, ctx.S
        dev = D.device
        H, W = D.shape
        n, fe, t0, dt, dx = cfg["n_steps"], cfg["frame_every"], cfg["t0"], cfg["dt"], cfg["dx"]
        L = _lib.lib()
        stim = _Stimuli(cfg["stimuli"], dev, (H, W))
        p = _lib.make_params(prm)
        nb = L.fk_adjoint_workspace_bytes(H, W)
        work = torch.empty(max(1, nb // 8), dtype=torch.float64, device=dev)
        s = stream_handle()
        _lib.check(L.fk_adjoint_prepare(D.data_ptr(), H, W, work.data_ptr(), nb, s))
        g_frames = g_frames.contiguous().to(torch.float32)
        lam = [torch.zeros((H, W), dtype=torch.float32, device=dev) for _ in range(3)]
        gD = torch.zeros((H, W), dtype=torch.float32, device=dev)
        # states x_c .. x_{e-1} of one segment, (v, w, u) planes per state
        states = torch.empty((S, 3, H, W), dtype=torch.float32, device=dev)
        for c in sorted(ctx.ckpts, reverse=True):
            e = min(n, c + S)
            x = ctx.ckpts[c]
            states[0, 0].copy_(x.v)
            states[0, 1].copy_(x.w)
            states[0, 2].copy_(x.u)
            if e - c > 1:  # recompute x_{c+1} .. x_{e-1}
                _lib.check(L.fk_forward_trajectory(
                    ctypes.byref(p), D.data_ptr(), x.v.data_ptr(), x.w.data_ptr(),
                    x.u.data_ptr(), H, W, stim.pointer(), stim.n, float(t0 + c), e - c - 1,
                    float(dt), float(dx), states[1].data_ptr(), s))
            # frame cotangents enter lam_u of x_{k+1} before step k is reversed
            inject = (ctypes.c_void_p * (e - c))()
            for k in range(c, e):
                if (k + 1) % fe == 0:
                    inject[k - c] = g_frames[(k + 1) // fe - 1].data_ptr()
            _lib.check(L.fk_adjoint_segment(
                ctypes.byref(p), D.data_ptr(), H, W, stim.pointer(), stim.n, float(t0 + c),
                e - c, float(dt), float(dx), states.data_ptr(), inject, lam[0].data_ptr(),
                lam[1].data_ptr(), lam[2].data_ptr(), gD.data_ptr(), work.data_ptr(), nb, s))
        gt = torch.empty(3, dtype=torch.float32, device=dev)
        _lib.check(L.fk_adjoint_tau_grads(work.data_ptr(), H, W, gt.data_ptr(), s))
        return gD, gt[0], gt[1], gt[2], None


def rollout(D: torch.Tensor, tau_d, tau_r, tau_si, state0, params, stimuli: Sequence, t0,
            n_steps: int, frame_every: int, dt: float, dx: float,
            checkpoint_every: int | None = None) -> torch.Tensor:
    """u after every ``frame_every`` Euler steps of cardiax/solve.py:92-100,
    shape (n_steps // frame_every, H, W), differentiable with respect to D
    (H x W) and the 0-d tensors tau_d, tau_r, tau_si (which replace those
    fields of ``params``; pass floats for constants, which also spares the
    host a device synchronisation). Other fields of ``params``, the initial
    state and the stimuli are constants."""
    if int(frame_every) <= 0 or int(n_steps) < 0:
        raise ValueError("frame_every must be positive and n_steps non-negative")
    shape = tuple(state0[2].shape)  # the u field of State(v, w, u)
    if tuple(D.shape) != shape:
        raise ValueError("D has shape %s, the state has %s" % (tuple(D.shape), shape))
    dev = require_gpu("rollout")
    D = on_device(D, dev) if not isinstance(D, torch.Tensor) else D
    if D.dtype != torch.float32 or D.device != dev:
        raise ValueError("D must be a float32 tensor on the GPU")
    # the kernels read D as a row-major H x W array; a strided view is copied
    # (differentiably, so its gradient maps back onto the caller's view)
    D = D.contiguous()
    taus, host = [], []
    for x in (tau_d, tau_r, tau_si):
        if isinstance(x, torch.Tensor):
            taus.append(x.to(device=dev, dtype=torch.float32).reshape(()))
            host.append(None)
        else:  # a constant: no gradient, value known on the host
            taus.append(torch.full((), float(x), dtype=torch.float32, device=dev))  # no sync
            host.append(float(x))
    cfg = {"params": params, "state0": state0, "stimuli": list(stimuli or []),
           "t0": _scalar(t0), "n_steps": int(n_steps), "frame_every": int(frame_every),
           "dt": float(dt), "dx": float(dx), "checkpoint_every": checkpoint_every,
           "tau_host": host}
    return _Rollout.apply(D, taus[0], taus[1], taus[2], cfg)
