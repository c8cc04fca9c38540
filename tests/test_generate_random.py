"""CPU tests of the random inputs of the dataset generator
(deepx/generate.py:16-132), drawn from seeded NumPy generators in place of
jax.random keys: the reference's ranges, distributions and quirks, and
determinism per seed. The builders run on the host path here (no GPU)."""
import numpy as np
import pytest
import torch

from cardiax_amd import generate as G
from cardiax_amd import stimulus as S


def _np(f):
    return f.cpu().numpy() if hasattr(f, "cpu") else np.asarray(f)


def test_protocol_ranges_and_determinism():
    for seed in range(200):
        p = G.random_protocol(seed, min_start=10, max_start=20)
        assert 10 <= p.start < 20 and p.duration == 2 and 400 <= p.period < 1e9
        assert p == G.random_protocol(seed, min_start=10, max_start=20)
    # an empty range returns its lower bound, as jax.random.randint does
    assert G.random_protocol(0, min_start=7, max_start=7).start == 7
    assert G.random_protocol(0, min_start=7, max_start=3).start == 7


def test_linear_directions_exclude_west_and_coverage_is_scaled():
    seen = set()
    for seed in range(90):
        f = _np(G.random_linear_stimulus(seed, (40, 30), S.Protocol(0, 2, 1e9), 1.0).field)
        rows, cols = np.nonzero(f)
        g1, _ = G._split(seed, 2)
        cov = abs(float(g1.normal()))
        d = S.Direction(int(g1.integers(0, 3)))
        seen.add(d)
        np.testing.assert_array_equal(f, S._stripe_mask((40, 30), d, cov * 0.2, 1.0))
    assert seen == {S.Direction.NORTH, S.Direction.EAST, S.Direction.SOUTH}


def test_rectangle_sizes_follow_shape0():
    shape = (300, 120)
    for seed in range(50):
        f = _np(G.random_rectangular_stimulus(seed, shape, S.Protocol(0, 2, 1e9), 0.6).field)
        assert set(np.unique(f)) <= {0.0, np.float32(0.6)}
        if f.any():
            r, c = np.nonzero(f)
            # sizes are drawn in [max(300 // 100, 10), 300 // 3) and clipped by the grid
            assert r.max() - r.min() + 1 < 100 and c.max() - c.min() + 1 < 100
    # small grids: the empty size range gives the lower bound (10)
    f = _np(G.random_rectangular_stimulus(1, (24, 24), S.Protocol(0, 2, 1e9), 1.0).field)
    assert f.shape == (24, 24)


def test_stimulus_builders_all_drawn_and_start_honoured():
    kinds = set()
    for seed in range(60):
        st = G.random_stimulus(seed, (48, 48), min_start=123, max_start=124)
        assert st.protocol.start == 123 and st.protocol.duration == 2
        g1, g2, g3 = G._split(seed, 3)
        kinds.add(int(g2.integers(0, 3)))
    assert kinds == {0, 1, 2}
    a = _np(G.random_stimulus(5, (48, 48)).field)
    b = _np(G.random_stimulus(5, (48, 48)).field)
    assert np.array_equal(a, b)


def test_random_sequence_refuses_without_gpu():
    from cardiax_amd.params import PARAMSET_3

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        G.random_sequence(0, PARAMSET_3, None, shape=(32, 32), n_stimuli=1, stop=2, step=1,
                          plot_while=False)
