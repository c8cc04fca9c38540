"""CPU pins of the FP64 autograd restatement (oracle/fk_autograd.py), the
stand-in for jax.grad through the reference solver.

Its forward values must equal the numpy oracle's, and its gradients must
match central finite differences of the numpy oracle where no gate flips
(gates are masks with zero derivative, as jax.grad treats greater_equal).
"""
import numpy as np
import torch

from cardiax_amd import params as P
from oracle import fk_autograd as fa
from oracle import fk_oracle as o


def _case(shape=(12, 14), seed=0):
    rng = np.random.default_rng(seed)
    st = o.State(rng.uniform(0.4, 1.0, shape), rng.uniform(0.4, 1.0, shape),
                 rng.uniform(0.0, 0.9, shape))
    D = rng.uniform(6e-4, 1e-3, shape)
    f = np.zeros(shape); f[:, :2] = 5.0
    stim = [(o.Protocol(2, 2, 1e9), f)]
    return st, D, stim


def test_forward_matches_numpy_oracle():
    st, D, stim = _case()
    prm = P.PARAMSET_3
    fr = fa.rollout_frames(st, prm, torch.tensor(D), stim, 0, 12, 4, 0.01, 0.01)
    ost = [o.Stimulus(o.Protocol(*p), f) for p, f in stim]
    cur = st
    for k, t in enumerate(range(0, 12, 4)):
        cur = o.forward_euler(cur, t, t + 4, prm, D, ost, 0.01, 0.01)
        np.testing.assert_allclose(fr[k].numpy(), cur.u, rtol=1e-13, atol=1e-13)


def test_gradients_match_central_differences():
    st, D, stim = _case(seed=1)
    prm = P.PARAMSET_3
    Dt = torch.tensor(D, requires_grad=True)
    taus = [torch.tensor(float(x), dtype=torch.float64, requires_grad=True)
            for x in (prm.tau_d, prm.tau_r, prm.tau_si)]
    rng = np.random.default_rng(2)
    target = rng.uniform(0, 1, (3,) + D.shape)
    fr = fa.rollout_frames(st, prm, Dt, stim, 0, 9, 3, 0.01, 0.01, taus)
    loss = ((fr - torch.tensor(target)) ** 2).mean()
    loss.backward()
    ost = [o.Stimulus(o.Protocol(*p), f) for p, f in stim]

    def L(Dm, prm_):
        cur, acc = st, 0.0
        for k in range(3):
            cur = o.forward_euler(cur, 3 * k, 3 * k + 3, prm_, Dm, ost, 0.01, 0.01)
            acc += ((cur.u - target[k]) ** 2).sum()
        return acc / target.size

    dD = rng.normal(size=D.shape) * 1e-4
    h = 1e-3
    fd = (L(D + h * dD, prm) - L(D - h * dD, prm)) / (2 * h)
    ad = float((Dt.grad.numpy() * dD).sum())
    assert abs(fd - ad) <= 2e-5 * abs(ad), (fd, ad)  # O(h^2) truncation
    for name, tt in zip(("tau_d", "tau_r", "tau_si"), taus):
        x0 = getattr(prm, name)
        hh = 1e-6 * x0
        fd = (L(D, prm._replace(**{name: x0 + hh})) - L(D, prm._replace(**{name: x0 - hh}))) \
            / (2 * hh)
        assert abs(fd - float(tt.grad)) <= 1e-5 * abs(fd) + 1e-12, (name, fd, float(tt.grad))
