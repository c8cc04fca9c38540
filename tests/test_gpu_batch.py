"""GPU tests of the batched forward path (cardiax_amd.batch).

Every simulation of a batch must evolve exactly as one _forward_euler call
on it would (cardiax/solve.py:92-100): both run the same streaming tile code,
so the comparison is bitwise.
"""
import numpy as np
import pytest
import torch

from cardiax_amd import params as P
from cardiax_amd import solve, workloads as wl
from cardiax_amd.batch import Batch, BatchSim
from cardiax_amd.stimulus import Protocol

pytestmark = pytest.mark.gpu


def _sims(seed=0):
    rng = np.random.default_rng(seed)
    shapes = [(64, 64), (37, 129), (130, 200), (8, 300), (96, 57), (256, 256)]
    names = sorted(P.ALL_PARAMSETS)
    out = []
    for k, sh in enumerate(shapes):
        D = wl.scar_map(sh, rng)
        f1 = np.zeros(sh, np.float32); f1[:, :max(1, sh[1] // 10)] = 20.0
        f2 = np.zeros(sh, np.float32); f2[: sh[0] // 2] = 20.0
        f3 = np.zeros(sh, np.float32); f3[sh[0] // 3: sh[0] // 2, 2:sh[1] // 2] = -3.0
        stims = [(Protocol(int(rng.integers(0, 5)), 2, 1e9), f1),
                 (Protocol(int(rng.integers(20, 60)), 2, 1e9), f2)]
        if k % 2:
            stims.append((Protocol(7, 3, 25), f3))
        st = solve.init(sh)
        out.append(BatchSim(st, P.ALL_PARAMSETS[names[(3 * k) % len(names)]], D, stims,
                            float(k), 0.01, 0.01))
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_per_simulation_runs(gpu, use_graph):
    sims = _sims()
    n = 100
    with Batch(sims) as b:
        assert b.n_tiles > len(sims)
        b.run(n, use_graph=use_graph)
        torch.cuda.synchronize()
        for s, got in zip(sims, b.states):
            ref = solve._forward_euler(s.state, s.t, s.t + n, s.params, s.diffusivity,
                                       s.stimuli, s.dt, s.dx)
            for a, r in zip(got, ref):
                assert torch.equal(a, r)


def test_batch_runs_continue_in_time(gpu):
    sims = _sims(1)
    with Batch(sims) as a, Batch(sims) as b:
        a.run(70)
        a.run(30)
        b.run(100)
        torch.cuda.synchronize()
        assert a.times == b.times == [s.t + 100 for s in sims]
        for x, y in zip(a.states, b.states):
            for p, q in zip(x, y):
                assert torch.equal(p, q)


def test_batch_argument_errors(gpu):
    sims = _sims(2)
    bad = sims[0]._replace(diffusivity=np.ones((3, 3), np.float32))
    with pytest.raises(ValueError):
        Batch([bad])
    f = np.zeros((64, 64), np.float32)
    many = sims[0]._replace(stimuli=[(Protocol(0, 2, 1e9), f)] * 5)
    with pytest.raises(NotImplementedError):
        Batch([many])
    with pytest.raises(ValueError):
        Batch([])


def _sims_two_step(seed=4):
    """Simulations that every two-step option fits (>= 48 rows; 56..96 or >= 104
    columns), with stimuli on odd and even steps and a periodic one."""
    rng = np.random.default_rng(seed)
    shapes = [(64, 64), (48, 56), (130, 200), (96, 104), (256, 256), (73, 96), (49, 131)]
    names = sorted(P.ALL_PARAMSETS)
    out = []
    for k, sh in enumerate(shapes):
        D = wl.scar_map(sh, rng)
        f1 = np.zeros(sh, np.float32); f1[:, :max(1, sh[1] // 10)] = 20.0
        f2 = np.zeros(sh, np.float32); f2[: sh[0] // 2] = 20.0
        f3 = np.zeros(sh, np.float32); f3[sh[0] // 3: sh[0] // 2, 2:sh[1] // 2] = -3.0
        stims = [(Protocol(int(rng.integers(0, 5)), 2, 1e9), f1),
                 (Protocol(int(rng.integers(20, 60)), 1, 1e9), f2),
                 (Protocol(7, 3, 25), f3)]
        st = tuple(torch.from_numpy(x.astype(np.float32)).cuda() for x in
                   (rng.uniform(0.3, 1.0, sh), rng.uniform(0.3, 1.0, sh),
                    rng.uniform(0.0, 1.0, sh)))
        out.append(BatchSim(st, P.ALL_PARAMSETS[names[(5 * k) % len(names)]], D, stims,
                            float(k), 0.01, 0.01))
    return out


def _per_sim_variant(sims, n, variant):
    from cardiax_amd import _lib

    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS[variant])
    try:
        return [solve._forward_euler(s.state, s.t, s.t + n, s.params, s.diffusivity,
                                     s.stimuli, s.dt, s.dx) for s in sims]
    finally:
        L.fk_set_forward_variant(prev)


@pytest.mark.parametrize("rows", [8, 16])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("n", [77, 64])
def test_batch_two_steps_per_launch_match_per_simulation_runs(gpu, rows, use_graph, n):
    """Two steps per launch over the whole batch (odd counts end with one
    single step) equal each simulation's own solve bit for bit."""
    sims = _sims_two_step()
    with Batch(sims) as b:
        b.set_rows(rows, two_step=True)
        b.run(n, use_graph=use_graph)
        torch.cuda.synchronize()
        for got, ref in zip(b.states, _per_sim_variant(sims, n, "streamt%d" % rows)):
            for a, r in zip(got, ref):
                assert torch.isfinite(r).all()
                assert torch.equal(a, r)


def test_batch_two_step_option_needs_every_simulation_to_fit(gpu):
    sims = _sims()  # holds an 8 x 300 simulation
    with Batch(sims) as b:
        with pytest.raises(NotImplementedError):
            b.set_rows(8, two_step=True)
    with Batch(_sims_two_step()) as b:
        res = b.autotune(n_steps=6)
        assert set(res["times_ms"]) == {"2", "4", "8", "16", "t8", "t16"}
