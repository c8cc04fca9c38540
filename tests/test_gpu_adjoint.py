"""GPU tests of the hand-written adjoint (fk_adjoint_step, cardiax_amd.adjoint).

References: the FP64 torch-autograd restatement of the reference step
(oracle/fk_autograd.py, the stand-in for jax.grad) and central finite
differences of the FP64 oracle. Tolerance: 1e-3 relative for the D-map
gradient of a rollout; single steps are checked tighter.
"""
import ctypes

import numpy as np
import pytest
import torch

from cardiax_amd import _lib, adjoint, solve
from cardiax_amd import params as P
from cardiax_amd import workloads as wl
from cardiax_amd.stimulus import Protocol
from oracle import fk_autograd as fa
from oracle import fk_oracle as o

pytestmark = pytest.mark.gpu


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    H, W = shape
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    u = 0.5 + 0.45 * np.sin(ii / 3.1) * np.cos(jj / 4.3)
    v = 0.7 + 0.2 * rng.uniform(size=shape)
    w = 0.8 + 0.1 * rng.uniform(size=shape)
    return tuple(np.asarray(x, np.float32) for x in (v, w, u))


@pytest.mark.parametrize("shape", [(3, 5), (9, 20), (40, 33), (70, 70)])
@pytest.mark.parametrize("pname", ["3", "1A", "7"])
def test_single_step_vjp_matches_autograd(gpu, shape, pname):
    prm = P.ALL_PARAMSETS[pname]
    st = _state(shape, 1)
    rng = np.random.default_rng(2)
    D = rng.uniform(5e-4, 1e-3, shape).astype(np.float32)
    lam = [rng.normal(size=shape).astype(np.float32) for _ in range(3)]
    f = np.zeros(shape, np.float32); f[: max(1, shape[0] // 3)] = 3.0
    stims = [(Protocol(4, 2, 1e9), f)]
    t = 5.0  # stimulus active
    dt, dx = 0.01, 0.01
    # reference: VJP of one step in float64
    Dt = torch.tensor(D.astype(np.float64), requires_grad=True)
    x = [torch.tensor(a.astype(np.float64), requires_grad=True) for a in st]
    taus = [torch.tensor(float(getattr(prm, k)), dtype=torch.float64, requires_grad=True)
            for k in adjoint.GRAD_FIELDS]
    names = Params_names()
    pd = {k: float(v) for k, v in zip(names, prm)}
    pd["tau_d"], pd["tau_r"], pd["tau_si"] = taus
    y = fa.step_euler(x[0], x[1], x[2], t, pd, Dt, stims, dt, dx)
    L = sum((yy * torch.tensor(ll.astype(np.float64))).sum() for yy, ll in zip(y, lam))
    L.backward()
    # GPU
    dev = gpu
    g = lambda a: torch.from_numpy(a).to(dev)
    Lb = _lib.lib()
    H, W = shape
    nb = Lb.fk_adjoint_workspace_bytes(H, W)
    work = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
    Dg, sg = g(D), [g(a) for a in st]
    lg = [g(a) for a in lam]
    out = [torch.empty_like(lg[0]) for _ in range(3)]
    gD = torch.zeros_like(Dg)
    gt = torch.empty(3, device=dev)
    p = _lib.make_params(prm)
    stim = solve._Stimuli(stims, dev, shape)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(Lb.fk_adjoint_prepare(Dg.data_ptr(), H, W, work.data_ptr(), nb, s))
    _lib.check(Lb.fk_adjoint_step(ctypes.byref(p), Dg.data_ptr(), sg[0].data_ptr(),
                                  sg[1].data_ptr(), sg[2].data_ptr(), H, W, stim.pointer(),
                                  stim.n, t, dt, dx, lg[0].data_ptr(), lg[1].data_ptr(),
                                  lg[2].data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                  out[2].data_ptr(), gD.data_ptr(), work.data_ptr(), nb, s))
    _lib.check(Lb.fk_adjoint_tau_grads(work.data_ptr(), H, W, gt.data_ptr(), s))
    torch.cuda.synchronize()
    for name, a, b in zip("vwu", out, x):
        ref = b.grad.numpy()
        err = np.abs(a.cpu().numpy() - ref).max() / max(1e-30, np.abs(ref).max())
        assert err < 2e-5, (name, err)
    ref = Dt.grad.numpy()
    err = np.abs(gD.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 2e-5, ("D", err)
    for k, tt in enumerate(taus):
        ref = float(tt.grad)
        assert abs(float(gt[k]) - ref) <= 1e-4 * abs(ref) + 1e-7, (adjoint.GRAD_FIELDS[k],
                                                                    float(gt[k]), ref)


def Params_names():
    return P.Params._fields


def _c4_small(shape=(64, 64), n=200, fe=50, seed=3):
    """C4 in miniature: hidden scar D*, target frames from it, loss = MSE
    between the frames of the guess D = 1e-3 and the targets."""
    rng = np.random.default_rng(seed)
    Dstar = wl.scar_map(shape, rng)
    f1 = np.zeros(shape, np.float32); f1[:, :6] = 20.0
    f2 = np.zeros(shape, np.float32); f2[: shape[0] // 2] = 20.0
    stims = [(Protocol(0, 2, 1e9), f1), (Protocol(n // 3, 2, 1e9), f2)]
    st = tuple(np.asarray(x, np.float32) for x in o.init(shape))
    return st, Dstar, stims


def test_rollout_gradient_matches_fp64_autograd(gpu):
    shape, n, fe = (64, 64), 200, 50
    st, Dstar, stims = _c4_small(shape, n, fe)
    prm = P.PARAMSET_3
    dev = gpu
    target = solve.forward(st, list(range(0, n + 1, fe)), prm, Dstar, stims, 0.01, 0.01)
    target = torch.stack([x.u for x in target])
    D0 = np.full(shape, 1e-3, np.float32)
    D = torch.tensor(D0, device=dev, requires_grad=True)
    taus = [torch.tensor(float(getattr(prm, k)), device=dev, requires_grad=True)
            for k in adjoint.GRAD_FIELDS]
    fr = adjoint.rollout(D, *taus, st, prm, stims, 0, n, fe, 0.01, 0.01)
    loss = ((fr - target) ** 2).mean()
    loss.backward()
    # FP64 autograd through the restatement of the reference
    Dr = torch.tensor(D0.astype(np.float64), requires_grad=True)
    tr = [torch.tensor(float(getattr(prm, k)), dtype=torch.float64, requires_grad=True)
          for k in adjoint.GRAD_FIELDS]
    frr = fa.rollout_frames(st, prm, Dr, stims, 0, n, fe, 0.01, 0.01, tr)
    lr = ((frr - target.cpu().double()) ** 2).mean()
    lr.backward()
    gref = Dr.grad.numpy()
    ggpu = D.grad.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(ggpu - gref) / np.linalg.norm(gref)
    assert rel <= 1e-3, rel
    for k, (a, b) in enumerate(zip(taus, tr)):
        ra, rb = float(a.grad), float(b.grad)
        assert abs(ra - rb) <= 1e-3 * abs(rb) + 1e-9, (adjoint.GRAD_FIELDS[k], ra, rb)
    assert abs(loss.item() - lr.item()) <= 1e-4 * lr.item()
    print("rollout D-gradient relative L2 error vs FP64 autograd: %.2e" % rel)


def test_rollout_gradient_matches_central_differences(gpu):
    shape, n, fe = (48, 40), 120, 40
    st, Dstar, stims = _c4_small(shape, n, fe, seed=5)
    prm = P.PARAMSET_3
    target = np.stack([x.u for x in o.forward(o.State(*[a.astype(np.float64) for a in st]),
                                              list(range(0, n + 1, fe)), prm,
                                              Dstar.astype(np.float64),
                                              [o.Stimulus(o.Protocol(*p), f) for p, f in stims],
                                              0.01, 0.01)])
    D0 = np.full(shape, 1e-3)
    D = torch.tensor(D0.astype(np.float32), device=gpu, requires_grad=True)
    fr = adjoint.rollout(D, prm.tau_d, prm.tau_r, prm.tau_si, st, prm, stims, 0, n, fe, 0.01,
                         0.01)
    ((fr - torch.tensor(target, dtype=torch.float32, device=gpu)) ** 2).mean().backward()
    g = D.grad.cpu().numpy().astype(np.float64)
    ost = [o.Stimulus(o.Protocol(*p), f) for p, f in stims]

    def L(Dm):
        fs = o.forward(o.State(*[a.astype(np.float64) for a in st]), list(range(0, n + 1, fe)),
                       prm, Dm, ost, 0.01, 0.01)
        return np.mean([(x.u - tt) ** 2 for x, tt in zip(fs, target)])

    rng = np.random.default_rng(9)
    for trial in range(2):
        dD = rng.normal(size=shape)
        dD = 1e-4 * dD / np.abs(dD).max()
        # D moves by <= 1e-8 (1e-5 relative): small enough that no Heaviside
        # gate flips between the two FP64 runs (a flip shifts the loss by a
        # jump that the gradient, with gates as masks, correctly ignores)
        h = 1e-4
        fd = (L(D0 + h * dD) - L(D0 - h * dD)) / (2 * h)
        ad = float((g * dD).sum())
        assert abs(fd - ad) <= 1e-3 * abs(fd), (trial, fd, ad)


def _dev_stims(stims, dev, shape):
    return solve._Stimuli(stims, dev, shape)


def test_trajectory_matches_forward_bitwise(gpu):
    shape, n = (37, 70), 9
    st, Dstar, stims = _c4_small(shape, 30, 10, seed=7)
    stims = [(s[0]._replace(start=3), s[1]) for s in stims]
    prm = P.PARAMSET_1A
    dev = gpu
    Lb = _lib.lib()
    H, W = shape
    D = torch.from_numpy(Dstar).to(dev)
    x0 = [torch.from_numpy(a).to(dev) for a in st]
    traj = torch.empty((n, 3, H, W), device=dev)
    stim = _dev_stims(stims, dev, shape)
    p = _lib.make_params(prm)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(Lb.fk_forward_trajectory(ctypes.byref(p), D.data_ptr(), x0[0].data_ptr(),
                                        x0[1].data_ptr(), x0[2].data_ptr(), H, W,
                                        stim.pointer(), stim.n, 2.0, n, 0.01, 0.01,
                                        traj.data_ptr(), s))
    for k in range(n):
        ref = solve._forward_euler(st, 2, 2 + k + 1, prm, Dstar, stims, 0.01, 0.01)
        for f in range(3):
            assert torch.equal(traj[k, f], ref[f]), (k, f)


def test_segment_equals_step_by_step_reversal(gpu):
    shape, n = (29, 41), 7
    st, Dstar, stims = _c4_small(shape, 30, 10, seed=8)
    stims = [(s[0]._replace(start=2), s[1]) for s in stims]
    prm = P.PARAMSET_3
    dev = gpu
    Lb = _lib.lib()
    H, W = shape
    D = torch.from_numpy(Dstar).to(dev)
    x0 = [torch.from_numpy(a).to(dev) for a in st]
    states = torch.empty((n, 3, H, W), device=dev)
    for f in range(3):
        states[0, f].copy_(x0[f])
    stim = _dev_stims(stims, dev, shape)
    p = _lib.make_params(prm)
    s = torch.cuda.current_stream().cuda_stream
    t0 = 1.0
    _lib.check(Lb.fk_forward_trajectory(ctypes.byref(p), D.data_ptr(), x0[0].data_ptr(),
                                        x0[1].data_ptr(), x0[2].data_ptr(), H, W,
                                        stim.pointer(), stim.n, t0, n - 1, 0.01, 0.01,
                                        states[1].data_ptr(), s))
    g = torch.Generator(device="cpu").manual_seed(3)
    lam0 = [torch.randn(shape, generator=g).to(dev) for _ in range(3)]
    inj = {2: torch.randn(shape, generator=g).to(dev), 5: torch.randn(shape, generator=g).to(dev)}
    nb = Lb.fk_adjoint_workspace_bytes(H, W)

    def fresh():
        work = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
        _lib.check(Lb.fk_adjoint_prepare(D.data_ptr(), H, W, work.data_ptr(), nb, s))
        return work

    # one segment call
    work = fresh()
    lam = [x.clone() for x in lam0]
    gD = torch.zeros_like(D)
    ptrs = (ctypes.c_void_p * n)()
    for k, x in inj.items():
        ptrs[k] = x.data_ptr()
    _lib.check(Lb.fk_adjoint_segment(ctypes.byref(p), D.data_ptr(), H, W, stim.pointer(),
                                     stim.n, t0, n, 0.01, 0.01, states.data_ptr(), ptrs,
                                     lam[0].data_ptr(), lam[1].data_ptr(), lam[2].data_ptr(),
                                     gD.data_ptr(), work.data_ptr(), nb, s))
    gt = torch.empty(3, device=dev)
    _lib.check(Lb.fk_adjoint_tau_grads(work.data_ptr(), H, W, gt.data_ptr(), s))
    # step by step
    work2 = fresh()
    cur = [x.clone() for x in lam0]
    gD2 = torch.zeros_like(D)
    for k in range(n - 1, -1, -1):
        if k in inj:
            cur[2] = cur[2] + inj[k]
        out = [torch.empty_like(x) for x in cur]
        st_k = states[k]
        _lib.check(Lb.fk_adjoint_step(ctypes.byref(p), D.data_ptr(), st_k[0].data_ptr(),
                                      st_k[1].data_ptr(), st_k[2].data_ptr(), H, W,
                                      stim.pointer(), stim.n, t0 + k, 0.01, 0.01,
                                      cur[0].data_ptr(), cur[1].data_ptr(), cur[2].data_ptr(),
                                      out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                      gD2.data_ptr(), work2.data_ptr(), nb, s))
        cur = out
    gt2 = torch.empty(3, device=dev)
    _lib.check(Lb.fk_adjoint_tau_grads(work2.data_ptr(), H, W, gt2.data_ptr(), s))
    torch.cuda.synchronize()
    for a, b in zip(lam, cur):
        assert torch.equal(a, b)
    assert torch.equal(gD, gD2)
    assert torch.equal(gt, gt2)


@pytest.mark.parametrize("variant", ["auto", "streamc2", "tb64x2", "pair2"])
def test_rollout_forward_is_the_recomputed_trajectory(gpu, variant):
    """The frames rollout returns are the states the backward pass recomputes
    (fk_forward_trajectory from the same checkpoint), bit for bit, whichever
    step kernel is set: the adjoint linearises around the forward's own
    trajectory."""
    shape, n, fe, S = (40, 52), 30, 5, 7
    st, Dstar, stims = _c4_small(shape, n, fe, seed=11)
    prm = P.PARAMSET_3
    H, W = shape
    Lb = _lib.lib()
    prev = Lb.fk_set_forward_variant(_lib.VARIANTS[variant])
    try:
        D = torch.from_numpy(Dstar).to(gpu)
        with torch.no_grad():
            fr = adjoint.rollout(D, prm.tau_d, prm.tau_r, prm.tau_si, st, prm, stims, 0, n, fe,
                                 0.01, 0.01, checkpoint_every=S)
        traj = torch.empty((n, 3, H, W), device=gpu)
        x0 = [torch.from_numpy(a).to(gpu) for a in st]
        stim = _dev_stims(stims, gpu, shape)
        p = _lib.make_params(prm)
        _lib.check(Lb.fk_forward_trajectory(ctypes.byref(p), D.data_ptr(), x0[0].data_ptr(),
                                            x0[1].data_ptr(), x0[2].data_ptr(), H, W,
                                            stim.pointer(), stim.n, 0.0, n, 0.01, 0.01,
                                            traj.data_ptr(),
                                            torch.cuda.current_stream().cuda_stream))
        for j in range(fe, n + 1, fe):
            assert torch.equal(fr[j // fe - 1], traj[j - 1, 2]), j
    finally:
        Lb.fk_set_forward_variant(prev)


def test_rollout_strided_D_gets_its_gradient(gpu):
    """A non-contiguous D (a column slice of a wider array) is read as the
    H x W field it denotes, and its gradient lands on the view."""
    shape, n, fe = (36, 30), 60, 20
    st, Dstar, stims = _c4_small(shape, n, fe, seed=12)
    prm = P.PARAMSET_3
    target = torch.stack([x.u for x in solve.forward(st, list(range(0, n + 1, fe)), prm, Dstar,
                                                     stims, 0.01, 0.01)])

    def grad_of(Dt):
        fr = adjoint.rollout(Dt, prm.tau_d, prm.tau_r, prm.tau_si, st, prm, stims, 0, n, fe,
                             0.01, 0.01)
        ((fr - target) ** 2).mean().backward()

    D = torch.full(shape, 1e-3, device=gpu, requires_grad=True)
    grad_of(D)
    wide = torch.full((shape[0], shape[1] + 9), 1e-3, device=gpu, requires_grad=True)
    view = wide[:, 4:4 + shape[1]]
    assert not view.is_contiguous()
    grad_of(view)
    assert torch.equal(wide.grad[:, 4:4 + shape[1]], D.grad)
    assert torch.count_nonzero(wide.grad[:, :4]) == 0
    assert torch.count_nonzero(wide.grad[:, 4 + shape[1]:]) == 0
