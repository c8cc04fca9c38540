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


def checkpoint_interval(n_steps: int, checkpoint_every: int | None = None) -> int:
    """Steps between checkpoints: ``checkpoint_every`` or ceil(sqrt(n_steps))."""
    return int(checkpoint_every) if checkpoint_every else max(1, int(math.ceil(math.sqrt(n_steps))))


class RolloutBatch:
    """Many ``rollout`` problems planned together (fk_rollout_batch_*): the
    forward solves and the reverse passes of all of them run as one launch
    per Euler step and one per reverse step, over work lists that hold every
    tile of every grid. Each problem is ``rollout(D, tau_d, tau_r, tau_si,
    state0, params, stimuli, t0, n_steps, frame_every, dt, dx)`` with the taus
    of ``params`` held constant; the results (frames and dL/dD) are bitwise
    those of ``rollout`` per problem with the same checkpoint interval.

    ``problems``: (state0, params, stimuli, t0) per problem; the grids may
    differ in size. The plan owns the device buffers: ``D`` (one tensor per
    problem, overwritten by ``forward``), ``frames``, ``g_frames`` and
    ``grad_D``. Use it as a context manager or call ``close``. A plan runs on
    the current stream of each call; calls must not overlap."""

    def __init__(self, problems, n_steps: int, frame_every: int, dt: float, dx: float,
                 checkpoint_every: int | None = None):
        dev = require_gpu("RolloutBatch")
        n_steps, fe = int(n_steps), int(frame_every)
        if n_steps < fe or fe < 1:
            raise ValueError("need n_steps >= frame_every >= 1 (got %d, %d)" % (n_steps, fe))
        self.n_steps, self.frame_every = n_steps, fe
        self.S = checkpoint_interval(n_steps, checkpoint_every)
        n_frames = n_steps // fe
        self.D, self.frames, self.g_frames, self.grad_D, self.g_tau = [], [], [], [], None
        self._keep = []
        specs = (_lib.FkRolloutSpec * len(problems))()
        for i, (state0, params, stimuli, t0) in enumerate(problems):
            st = State(*(on_device(x, dev).contiguous() for x in state0))
            H, W = st.u.shape
            stim = _Stimuli(stimuli, dev, (H, W))
            if stim.n > _lib.FK_BATCH_MAX_STIMULI:
                raise NotImplementedError("at most %d stimuli per problem, got %d"
                                          % (_lib.FK_BATCH_MAX_STIMULI, stim.n))
            D = torch.empty((H, W), dtype=torch.float32, device=dev)
            fr = torch.empty((n_frames, H, W), dtype=torch.float32, device=dev)
            gf = torch.zeros((n_frames, H, W), dtype=torch.float32, device=dev)
            gD = torch.empty((H, W), dtype=torch.float32, device=dev)
            self.D.append(D)
            self.frames.append(fr)
            self.g_frames.append(gf)
            self.grad_D.append(gD)
            self._keep.append((st, stim))
            sp = specs[i]
            sp.params = _lib.make_params(params)
            sp.D, sp.v0, sp.w0, sp.u0 = D.data_ptr(), st.v.data_ptr(), st.w.data_ptr(), \
                st.u.data_ptr()
            sp.H, sp.W, sp.n_stimuli = H, W, stim.n
            for k in range(stim.n):
                sp.stimuli[k] = stim.arr[k]
            sp.t0, sp.dt, sp.dx = float(_scalar(t0)), float(dt), float(dx)
            sp.frames, sp.g_frames, sp.grad_D = fr.data_ptr(), gf.data_ptr(), gD.data_ptr()
        torch.cuda.synchronize()  # initial states made on other streams are complete
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().fk_rollout_batch_create(specs, len(problems), n_steps, fe, self.S,
                                                      ctypes.byref(self._h)))
        self.n = len(problems)

    @property
    def device_bytes(self) -> int:
        """Device memory of the plan itself (checkpoints, segment states,
        cotangents, workspaces), bytes."""
        return int(_lib.lib().fk_rollout_batch_device_bytes(self._h))

    def forward(self, Ds):
        """Frames of every problem at the D-maps ``Ds`` (copied into the plan's
        D buffers): the plan's ``frames`` tensors, overwritten."""
        self._require_open()
        Ds = list(Ds)
        if len(Ds) != self.n:
            raise ValueError("%d D-maps for %d problems" % (len(Ds), self.n))
        for src, dst in zip(Ds, self.D):
            if tuple(src.shape) != tuple(dst.shape):
                raise ValueError("D has shape %s, the problem's grid is %s"
                                 % (tuple(src.shape), tuple(dst.shape)))
        with torch.no_grad():
            for dst, src in zip(self.D, Ds):
                dst.copy_(src)
        _lib.check(_lib.lib().fk_rollout_batch_forward(self._h, stream_handle()))
        return self.frames

    def backward(self, g_frames=None, tau_grads: bool = False):
        """dLoss/dD of every problem from the frame cotangents (the plan's
        ``g_frames`` when None) of the last ``forward``. Returns the plan's
        ``grad_D`` tensors and, with ``tau_grads``, an (n, 3) tensor of
        gradients with respect to (tau_d, tau_r, tau_si)."""
        self._require_open()
        if g_frames is not None:
            g_frames = list(g_frames)
            if len(g_frames) != self.n or any(tuple(g.shape) != tuple(f.shape)
                                              for g, f in zip(g_frames, self.g_frames)):
                raise ValueError("g_frames must match the plan's frames in number and shape")
            with torch.no_grad():
                for dst, src in zip(self.g_frames, g_frames):
                    dst.copy_(src)
        gt = None
        if tau_grads:
            gt = torch.empty((self.n, 3), dtype=torch.float32, device=self.grad_D[0].device)
        _lib.check(_lib.lib().fk_rollout_batch_backward(
            self._h, gt.data_ptr() if gt is not None else None, stream_handle()))
        return (self.grad_D, gt) if tau_grads else self.grad_D

    def _require_open(self):
        if self._h is None or not self._h.value:
            raise RuntimeError("the rollout batch is closed")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.check(_lib.lib().fk_rollout_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass


class _RolloutBatchFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, *Ds):
        ctx.plan = plan
        frames = plan.forward(Ds)
        return tuple(f.clone() for f in frames)

    @staticmethod
    def backward(ctx, *g_frames):
        plan = ctx.plan
        zeros = [torch.zeros_like(f) if g is None else g for g, f in zip(g_frames, plan.frames)]
        grads = plan.backward([g.contiguous().to(torch.float32) for g in zeros])
        return (None,) + tuple(g.clone() for g in grads)


def rollout_batch(plan: RolloutBatch, Ds):
    """The frames of every problem of ``plan`` at ``Ds``, differentiable with
    respect to each D (the batched ``rollout``)."""
    if len(Ds) != plan.n:
        raise ValueError("%d D-maps for %d problems" % (len(Ds), plan.n))
    for D, ref in zip(Ds, plan.D):
        if tuple(D.shape) != tuple(ref.shape):
            raise ValueError("D has shape %s, the problem's grid is %s"
                             % (tuple(D.shape), tuple(ref.shape)))
    return _RolloutBatchFn.apply(plan, *Ds)
