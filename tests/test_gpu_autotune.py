"""GPU tests of the autotuners (fk_autotune_forward, fk_batch_autotune).

Tuning only changes which kernel runs: results must stay within the parity
bar of the FP64 oracle (forward) and bit-identical to per-simulation runs
(batch), and the tuned choice must be what automatic launches use.
"""
import numpy as np
import pytest
import torch

from cardiax_amd import _lib, autotune, solve
from cardiax_amd import workloads as wl
from cardiax_amd.batch import Batch
from oracle import c_oracle
from oracle import fk_oracle as o

from test_gpu_batch import _sims

pytestmark = pytest.mark.gpu


def _per_sim(sims, n, rows):
    """Per-simulation runs on the single-grid kernel with the same band height
    (bitwise comparisons need the same instantiation of the step code)."""
    L = _lib.lib()
    prev = L.fk_set_forward_variant(_lib.VARIANTS["stream%d" % rows])
    try:
        return [solve._forward_euler(s.state, s.t, s.t + n, s.params, s.diffusivity,
                                     s.stimuli, s.dt, s.dx) for s in sims]
    finally:
        L.fk_set_forward_variant(prev)


def test_forward_autotune_records_and_is_used(gpu):
    shape = (96, 130)
    autotune.reset()
    try:
        assert autotune.lookup(shape) == "auto"
        res = autotune.tune(shape, n_steps=20)
        timed = set(res["times_ms"])
        assert timed == {"stream2", "stream4", "stream8", "stream16", "stream32", "streamc2",
                         "streamc4", "streamc8", "streamc16", "pair2", "pair4", "tile16",
                         "tile32", "tb64x2", "tb64x4", "tb32x2", "tb32x4", "streamt8",
                         "streamt12", "streamt16", "streamt24"}
        assert all(t > 0 for t in res["times_ms"].values())
        # the library takes the strict minimum of the unrounded times; the
        # reported times are rounded to 1e-4 ms, so near-ties may reorder
        assert res["times_ms"][res["best"]] <= min(res["times_ms"].values()) + 1e-4
        assert autotune.lookup(shape) == res["best"]
        assert autotune.lookup((97, 130)) == "auto"
        # an automatic solve on the tuned shape stays within the parity bar
        sim = wl.c1_inputs(3, shape=shape, steps=150)
        stims = [(s.protocol, s.field) for s in sim.stimuli]
        ostims = [o.Stimulus(o.Protocol(*s.protocol), s.field) for s in sim.stimuli]
        g = solve._forward_euler(solve.init(shape), 0, 150, sim.params, sim.D, stims, 0.01, 0.01)
        r = c_oracle.forward_euler(o.init(shape), 0, 150, sim.params, sim.D, ostims, 0.01, 0.01)
        for a, b in zip(g, r):
            assert np.abs(a.cpu().numpy() - b).max() <= 1e-4
    finally:
        autotune.reset()
    assert autotune.lookup(shape) == "auto"


def test_autotune_rejects_bad_arguments(gpu):
    L = _lib.lib()
    assert L.fk_autotune_forward(2, 50, 10, None, None, None) == _lib.FK_ERR_ARG
    assert L.fk_autotune_forward(50, 50, 0, None, None, None) == _lib.FK_ERR_ARG


def test_batch_autotune_keeps_results_bitwise(gpu):
    sims = _sims(2)
    n = 80
    with Batch(sims) as b:
        before = [torch.stack(list(s)).clone() for s in b.states]
        res = b.autotune(n_steps=8)
        # an 8-row simulation rules out two steps per launch
        assert res["rows"] in (2, 4, 8, 16) and not res["two_step"]
        assert set(res["times_ms"]) == {"2", "4", "8", "16"}
        assert all(t > 0 for t in res["times_ms"].values())
        # tuning never touches the simulations' states
        for x, s in zip(before, b.states):
            assert torch.equal(x, torch.stack(list(s)))
        b.run(n)
        torch.cuda.synchronize()
        for got, ref in zip(b.states, _per_sim(sims, n, res["rows"])):
            for a, r in zip(got, ref):
                assert torch.equal(a, r)


@pytest.mark.parametrize("rows", [2, 4, 8, 16])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_band_heights_match_per_simulation_runs(gpu, rows, use_graph):
    sims = _sims(3)
    n = 90
    with Batch(sims) as b:
        b.set_rows(rows)
        b.run(n, use_graph=use_graph)
        torch.cuda.synchronize()
        for got, ref in zip(b.states, _per_sim(sims, n, rows)):
            for a, r in zip(got, ref):
                assert torch.equal(a, r)
        with pytest.raises(ValueError):
            b.set_rows(12)


def test_autotune_on_the_callers_problem(gpu):
    """fk_autotune_forward_state times every candidate on the given state,
    parameters, D-map and stimuli, leaves them unmodified, and records the
    fastest for the shape like fk_autotune_forward."""
    shape = (130, 140)
    sim = wl.c1_inputs(4, shape=shape, steps=400)
    stims = [(s.protocol._replace(start=0 if k == 0 else 150), torch.from_numpy(s.field).cuda())
             for k, s in enumerate(sim.stimuli)]
    D = torch.from_numpy(sim.D).cuda()
    st = solve._forward_euler(solve.init(shape), 0, 200, sim.params, D, stims, 0.01, 0.01)
    keep = [x.clone() for x in st]
    autotune.reset()
    try:
        res = autotune.tune_on(st, 200, sim.params, D, stims, 0.01, 0.01, n_steps=30)
        assert set(res["times_ms"]) == set(_lib.VARIANTS) - {"auto", "plain"}
        assert all(t > 0 for t in res["times_ms"].values())
        assert res["times_ms"][res["best"]] <= min(res["times_ms"].values()) + 1e-4
        assert autotune.lookup(shape) == res["best"]
        for a, b in zip(st, keep):
            assert torch.equal(a, b)
        with pytest.raises(ValueError):  # a D-map of another shape
            autotune.tune_on(st, 200, sim.params, D[:, :-1], stims, 0.01, 0.01, n_steps=5)
    finally:
        autotune.reset()
