"""Seeded stand-ins for the ``jax.random`` keys of the reference's data
generator (deepx/generate.py, deepx/utils_scars.py), which are absent here: a
seed, NumPy Generator or SeedSequence takes the place of a key and
``SeedSequence.spawn`` the place of ``jax.random.split``."""
from __future__ import annotations

import numpy as np


def _seq(rng) -> np.random.SeedSequence:
    if isinstance(rng, np.random.SeedSequence):
        return rng
    if isinstance(rng, np.random.Generator):
        return np.random.SeedSequence(int(rng.integers(0, 2 ** 63)))
    return np.random.SeedSequence(int(rng))


def _split(rng, n):
    """jax.random.split: n independent child generators."""
    return [np.random.default_rng(c) for c in _seq(rng).spawn(n)]


def _randint(g, lo, hi, size=None):
    """jax.random.randint(key, size, lo, hi): [lo, hi), lo when hi <= lo."""
    lo, hi = int(lo), int(hi)
    if hi <= lo:
        return lo if size is None else np.full(size, lo, dtype=np.int64)
    return int(g.integers(lo, hi)) if size is None else g.integers(lo, hi, size=size)
