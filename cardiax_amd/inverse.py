"""Batched inverse problem: recover scar D-maps from sampled u frames.

Each instance holds a hidden diffusivity map D* and the u frames that the
forward-Euler solve (cardiax/solve.py:92-100) produces from it. Starting
from the healthy map D = 1e-3, Adam minimises the frame MSE (accumulated in
float64) by gradients from the hand-written adjoint (``adjoint.rollout``),
keeping D inside the reference's range [1e-4, 1e-3] (deepx/generate.py:80).

D is parameterised by a coarse grid (one node per ``coarse`` cells) and its
bilinear interpolation: Adam's first steps move every free value by about
the learning rate whatever the gradient's size, and a cell-by-cell D of such
steps has the checkerboard structure that the reference's explicit scheme
with its non-conservative D_x u_x term does not damp. The interpolation is
two small matrix products (deterministic, unlike scattered atomics).

The reference has no such solver loop (its deepx package learns the inverse map
with a network instead). By default all instances advance together through
one ``adjoint.RolloutBatch`` plan: each Euler step of the forward solves, and
each reverse step of the adjoint, is one launch over a work list holding every
tile of every instance, with results bitwise equal to the per-instance
rollouts. ``batched=False`` runs the per-instance rollouts instead,
interleaved on several HIP streams. No step of the loop synchronises the host
with the device.
"""
from __future__ import annotations

import time
from typing import List, NamedTuple, Sequence

import numpy as np
import torch

from . import adjoint, solve
from . import workloads as wl
from ._device import require_gpu
from .params import PARAMSET_3
from .stimulus import Direction, Protocol, _stripe_mask

D_MIN, D_MAX = 1e-4, 1e-3


class Instance(NamedTuple):
    """One recovery problem: the hidden map, its target frames and inputs."""
    D_true: torch.Tensor
    target: torch.Tensor      # (n_frames, H, W)
    state0: solve.State
    params: object
    stimuli: list
    n_steps: int
    frame_every: int


def make_instances(seed: int = 0, n: int = 32, sides=(128, 256, 384, 512),
                   n_steps: int = 1000, frame_every: int = 50, dt: float = wl.DT,
                   dx: float = wl.DX) -> List[Instance]:
    """``n`` square instances with sides drawn from ``sides``: a seeded scar
    map, S1 (WEST stripe at step 0) and S2 (NORTH half at a seeded step in
    the first half of the run), and the target frames of the hidden map."""
    dev = require_gpu("inverse")
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        side = int(rng.choice(sides))
        shape = (side, side)
        Dt = torch.from_numpy(wl.scar_map(shape, rng)).to(dev)
        f1 = torch.from_numpy(_stripe_mask(shape, Direction.WEST, 0.1, wl.MODULUS)).to(dev)
        f2 = torch.from_numpy(_stripe_mask(shape, Direction.NORTH, 0.5, wl.MODULUS)).to(dev)
        s2 = int(rng.integers(n_steps // 5, n_steps // 2))
        stims = [(Protocol(0, 2, wl.PERIOD), f1), (Protocol(s2, 2, wl.PERIOD), f2)]
        st = x = solve.init(shape)
        frames = []
        for t in range(0, n_steps - frame_every + 1, frame_every):
            x = solve._forward_euler(x, t, t + frame_every, PARAMSET_3, Dt, stims, dt, dx)
            frames.append(x.u)
        out.append(Instance(Dt, torch.stack(frames), st, PARAMSET_3, stims, n_steps,
                            frame_every))
    return out


def interp_matrix(n: int, m: int, device) -> torch.Tensor:
    """(n, m) matrix of 1-D linear interpolation from m nodes spanning [0, n-1]
    (end points aligned) to the n cell centres."""
    x = torch.arange(n, dtype=torch.float64) * ((m - 1) / max(n - 1, 1))
    lo = torch.clamp(x.floor().long(), 0, m - 2)
    f = x - lo
    M = torch.zeros((n, m), dtype=torch.float64)
    M[torch.arange(n), lo] = 1.0 - f
    M[torch.arange(n), lo + 1] += f
    return M.to(device=device, dtype=torch.float32)


class _Param(NamedTuple):
    theta: torch.Tensor  # coarse nodes (requires grad)
    Ry: torch.Tensor
    Rx: torch.Tensor

    def D(self) -> torch.Tensor:
        return self.Ry @ self.theta @ self.Rx.t()


# Device memory the batched plan may spend on keeping every state of a rollout
# (no checkpoint recompute); above it, states are recomputed from sqrt(N)
# checkpoints. The gradients are the same bits either way.
TRAJECTORY_BUDGET_BYTES = 64 << 30


def _plan(instances: Sequence[Instance], dt: float, dx: float) -> adjoint.RolloutBatch:
    n_steps, fe = instances[0].n_steps, instances[0].frame_every
    if any(x.n_steps != n_steps or x.frame_every != fe for x in instances):
        raise ValueError("a batched plan needs one n_steps and frame_every for all instances")
    cells = sum(int(x.D_true.numel()) for x in instances)
    problems = [(x.state0, x.params, x.stimuli, 0) for x in instances]
    if 12 * cells * n_steps <= TRAJECTORY_BUDGET_BYTES:
        try:
            return adjoint.RolloutBatch(problems, n_steps, fe, dt, dx, checkpoint_every=n_steps)
        except RuntimeError:  # the device cannot hold every state: checkpoint instead
            pass
    return adjoint.RolloutBatch(problems, n_steps, fe, dt, dx)


def _frame_losses(frames, instances) -> List[torch.Tensor]:
    """The frame MSE of each instance (accumulated in float64)."""
    out = []
    for fr, inst in zip(frames, instances):
        diff = (fr - inst.target).double()
        out.append((diff * diff).mean())
    return out


def recover(instances: Sequence[Instance], iters: int = 50, lr: float = 2e-5,
            n_streams: int = 8, coarse: int = 8, dt: float = wl.DT,
            dx: float = wl.DX, batched: bool = True, d_err_at=()) -> dict:
    """Runs ``iters`` Adam iterations of forward + adjoint on every instance.
    Returns the estimates and per-instance losses and D errors (relative L2
    distance to the hidden map, before and after), plus the wall time.
    ``batched``: all instances through one RolloutBatch plan (they must share
    n_steps and frame_every); otherwise per-instance rollouts on
    ``n_streams`` streams. ``d_err_at``: iteration counts after which the
    batched path also records every instance's D error (``res["d_err_at"]``)."""
    dev = require_gpu("inverse")
    if batched:
        return _recover_batched(instances, iters, lr, coarse, dt, dx, dev, tuple(d_err_at))
    streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
    params = _init_params(instances, coarse, dev)
    opts = [torch.optim.Adam([prm.theta], lr=lr) for prm in params]
    losses = torch.zeros((len(instances), iters), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(iters):
        for k, (inst, prm, opt) in enumerate(zip(instances, params, opts)):
            s = streams[k % len(streams)]
            with torch.cuda.stream(s):
                opt.zero_grad(set_to_none=True)
                loss = _loss_backward(inst, prm.D().contiguous(), dt, dx)
                opt.step()
                with torch.no_grad():  # interpolation keeps D within the nodes' range
                    prm.theta.clamp_(D_MIN, D_MAX)
                losses[k, it] = loss.detach()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return _result(instances, params, losses, wall)


def _init_params(instances, coarse, dev):
    params, opts = [], []
    for inst in instances:
        H, W = inst.D_true.shape
        h, w = max(2, -(-H // coarse) + 1), max(2, -(-W // coarse) + 1)
        th = torch.full((h, w), D_MAX, dtype=torch.float32, device=dev, requires_grad=True)
        params.append(_Param(th, interp_matrix(H, h, dev), interp_matrix(W, w, dev)))
    return params


def _recover_batched(instances, iters, lr, coarse, dt, dx, dev, d_err_at=()) -> dict:
    params = _init_params(instances, coarse, dev)
    marks = {}
    opts = [torch.optim.Adam([prm.theta], lr=lr) for prm in params]
    losses = torch.zeros((len(instances), iters), dtype=torch.float64, device=dev)
    with _plan(instances, dt, dx) as plan:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(iters):
            for opt in opts:
                opt.zero_grad(set_to_none=True)
            Ds = [prm.D().contiguous() for prm in params]
            frames = adjoint.rollout_batch(plan, Ds)
            ls = _frame_losses(frames, instances)
            torch.stack(ls).sum().backward()
            for k, (opt, prm) in enumerate(zip(opts, params)):
                opt.step()
                with torch.no_grad():  # interpolation keeps D within the nodes' range
                    prm.theta.clamp_(D_MIN, D_MAX)
                losses[k, it] = ls[k].detach()
            if it + 1 in d_err_at:  # D errors after this iteration (on the device)
                marks[it + 1] = [_d_err(inst, prm) for inst, prm in zip(instances, params)]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    res = _result(instances, params, losses, wall)
    res["d_err_at"] = {k: [float(x) for x in v] for k, v in marks.items()}
    return res


def _d_err(inst, prm) -> torch.Tensor:
    with torch.no_grad():
        ref = inst.D_true.double()
        return torch.linalg.norm(prm.D().double() - ref) / torch.linalg.norm(ref)


def reached(inst: Instance, threshold: float = 0.5) -> torch.Tensor:
    """Cells some target frame shows excited (u >= threshold): where a wave
    passed while the frames were sampled, the only cells whose D the frames
    constrain."""
    return inst.target.max(dim=0).values >= threshold


def _result(instances, params, losses, wall) -> dict:
    err0, err1, act0, act1, frac, Ds = [], [], [], [], [], []
    for inst, prm in zip(instances, params):
        with torch.no_grad():
            D = prm.D()
            Ds.append(D)
            ref = inst.D_true.double()
            D0 = torch.full_like(ref, D_MAX)
            nrm = float(torch.linalg.norm(ref))
            err0.append(float(torch.linalg.norm(D0 - ref)) / nrm)
            err1.append(float(torch.linalg.norm(D.double() - ref)) / nrm)
            m = reached(inst)
            frac.append(float(m.double().mean()))
            nm = float(torch.linalg.norm(ref[m])) if bool(m.any()) else 0.0
            act0.append(float(torch.linalg.norm((D0 - ref)[m])) / nm if nm > 0 else 0.0)
            act1.append(float(torch.linalg.norm((D.double() - ref)[m])) / nm if nm > 0 else 0.0)
    return {"D": Ds, "loss": losses.cpu().numpy(), "seconds": wall,
            "d_err_initial": err0, "d_err_final": err1,
            "d_err_reached_initial": act0, "d_err_reached_final": act1, "reached_frac": frac}


def _loss_backward(inst: Instance, D: torch.Tensor, dt: float, dx: float) -> torch.Tensor:
    """Frame MSE of one instance at D (accumulated in float64) and its
    backward pass through the adjoint: the gradient reaches whatever D was
    computed from. Runs on the current stream."""
    p = inst.params
    fr = adjoint.rollout(D, p.tau_d, p.tau_r, p.tau_si, inst.state0, p, inst.stimuli, 0,
                         inst.n_steps, inst.frame_every, dt, dx)
    diff = (fr - inst.target).double()
    loss = (diff * diff).mean()
    loss.backward()
    return loss


def d_gradients(instances: Sequence[Instance], Ds: Sequence[torch.Tensor], n_streams: int = 8,
                dt: float = wl.DT, dx: float = wl.DX, batched: bool = False,
                plan: adjoint.RolloutBatch | None = None) -> List[torch.Tensor]:
    """The gradient of each instance's frame MSE with respect to its full D
    field at Ds[k], by the pass ``recover`` makes in every iteration: through
    one RolloutBatch plan (``batched``; ``plan`` may be given to reuse one),
    or per instance, interleaved on ``n_streams`` HIP streams."""
    require_gpu("inverse")
    if batched or plan is not None:
        own = plan is None
        plan = _plan(instances, dt, dx) if own else plan
        try:
            Dk = [D.detach().clone().requires_grad_(True) for D in Ds]
            frames = adjoint.rollout_batch(plan, Dk)
            torch.stack(_frame_losses(frames, instances)).sum().backward()
            torch.cuda.synchronize()
            return [D.grad for D in Dk]
        finally:
            if own:
                plan.close()
    streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
    torch.cuda.synchronize()  # inputs made on other streams are complete
    out = []
    for k, (inst, D) in enumerate(zip(instances, Ds)):
        with torch.cuda.stream(streams[k % len(streams)]):
            Dk = D.detach().clone().requires_grad_(True)
            _loss_backward(inst, Dk, dt, dx)
            out.append(Dk.grad)
    torch.cuda.synchronize()
    return out


def forward_only_seconds(instances: Sequence[Instance], n_streams: int = 8,
                         dt: float = wl.DT, dx: float = wl.DX) -> float:
    """Wall time of one forward rollout of every instance (no gradient), the
    denominator of the forward + adjoint cost ratio."""
    streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        for k, inst in enumerate(instances):
            with torch.cuda.stream(streams[k % len(streams)]):
                p = inst.params
                adjoint.rollout(torch.full_like(inst.D_true, D_MAX), p.tau_d, p.tau_r, p.tau_si,
                                inst.state0, p, inst.stimuli, 0, inst.n_steps,
                                inst.frame_every, dt, dx)
    torch.cuda.synchronize()
    return time.perf_counter() - t0
