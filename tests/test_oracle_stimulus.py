"""Pins of the oracle's stimulus law (CPU).

Known answers: the activity sets of the reference's own macro test
(tests/macro/stimulate_test.py:16-18) and the mean-0.6 check of
tests/unittests/stimulus_test.py:23, both written against an older API but
still valid for cardiax/solve.py:257-271. Further cases follow from the law
t >= start and mod(start - t + 1, period) < duration.
"""
import numpy as np
import pytest

from cardiax_amd import stimulus as S
from oracle import fk_oracle as o


def _active_set(start, duration, period, times=range(0, 300)):
    p = o.Protocol(start, duration, period)
    return {t for t in times if o.stimulus_active(t, p)}


def test_reference_macro_known_answer_sets():
    A = {0, 1, 50, 51, 100, 101, 150, 151, 200, 201, 250, 251}
    B = {10, 11, 60, 61, 110, 111, 160, 161, 210, 211, 260, 261}
    C = {30, 31}
    assert _active_set(0, 2, 50) == A
    assert _active_set(10, 2, 50) == B
    assert _active_set(30, 2, 1000000) == C


def test_float_time_and_huge_period():
    # forward() passes float times; every call site uses period 1e9
    assert _active_set(0.0, 2.0, 1e9, [float(t) for t in range(0, 300)]) == {0.0, 1.0}
    assert _active_set(40000, 2, 1e9, range(39990, 40010)) == {40000, 40001}


def test_derived_law_for_other_durations():
    # first pulse {start+1} only when d = 1; later pulses start+1+nP-j, j<d
    assert _active_set(5, 1, 20, range(0, 60)) == {6, 26, 46}
    assert _active_set(5, 3, 20, range(0, 60)) == {5, 6, 24, 25, 26, 44, 45, 46}


def test_reference_unit_mean_check():
    # stimulus_test.py:12-25: NORTH stripe, coverage 0.5, modulus 0.6, t = 1
    shape = (10, 10)
    s = S.linear(shape, S.Direction.NORTH, 0.5, 0.6, S.Protocol(0, 2, 1e9))
    field = s.field.cpu().numpy()
    out = o.stimulate(1, np.zeros(shape), [o.Stimulus(o.Protocol(0, 2, 1e9), field)])
    assert abs(np.mean(out[:5]) - 0.6) <= 1e-3
    assert np.all(out[5:] == 0)


def test_replacement_override_order_zero_and_negative_fields():
    X = np.full((4, 4), 7.0)
    f1 = np.zeros((4, 4)); f1[:, :2] = 2.0
    f2 = np.zeros((4, 4)); f2[:, 1:3] = -5.0
    p = o.Protocol(0, 2, 1e9)
    out = o.stimulate(0, X, [o.Stimulus(p, f1), o.Stimulus(p, f2)])
    # replacement (not addition), later stimulus wins where nonzero,
    # a zero entry never stimulates, a negative amplitude does
    np.testing.assert_array_equal(out[0], [2.0, -5.0, -5.0, 7.0])
    out = o.stimulate(5, X, [o.Stimulus(p, f1), o.Stimulus(p, f2)])
    np.testing.assert_array_equal(out, X)


def test_builders_match_reference_slicing():
    shape = (20, 30)
    p = S.Protocol(0, 2, 1e9)
    f = lambda d, c: S.linear(shape, d, c, 1.0, p).field.cpu().numpy()
    # stripe width uses shape[0] for every direction (cardiax/stimulus.py:88)
    assert f(S.Direction.WEST, 0.2)[:, :4].all() and f(S.Direction.WEST, 0.2)[:, 4:].sum() == 0
    assert f(S.Direction.EAST, 0.2)[:, -4:].all() and f(S.Direction.EAST, 0.2).sum() == 80
    assert f(S.Direction.NORTH, 0.2)[:4].all() and f(S.Direction.NORTH, 0.2).sum() == 120
    assert f(S.Direction.SOUTH, 0.2)[-4:].all() and f(S.Direction.SOUTH, 0.2).sum() == 120
    # width 0 makes the "-0:" slice cover the whole field, as in the reference
    assert f(S.Direction.EAST, 0.01).sum() == 600
    with pytest.raises(ValueError):
        S.linear(shape, 7, 0.2, 1.0, p)
    r = S.rectangular(shape, (10, 15), (4, 6), 3.0, p).field.cpu().numpy()
    assert r[8:12, 12:18].all() and r.sum() == 3.0 * 24
    # a rectangle reaching past the top edge yields a negative start index,
    # which numpy-style slicing turns into an empty slice
    r = S.rectangular(shape, (1, 15), (6, 6), 3.0, p).field.cpu().numpy()
    assert r.sum() == 0
