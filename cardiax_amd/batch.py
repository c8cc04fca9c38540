"""Batched forward Euler over many independent simulations.

Each simulation is what one ``cardiax.solve._forward_euler`` call advances
(cardiax/solve.py:92-100): its own state, parameter set, D-map, stimuli and
time. ``Batch`` builds one work list over all of them; every Euler step of
the whole batch is one kernel launch, and the launches are replayed from a
recorded hipGraph (one graph launch per 32 steps); with a two-step option
each launch advances the batch two steps. The reference has no
batched path: its experiments loop over simulations one call at a time
(experiments/generate_fd_data_128.py:25-56).
"""
from __future__ import annotations

import ctypes
from typing import List, NamedTuple, Sequence

import torch

from . import _lib
from ._device import on_device, require_gpu, stream_handle
from .solve import State

FK_BATCH_MAX_STIMULI = _lib.FK_BATCH_MAX_STIMULI
# fk_batch_autotune's option order
BATCH_OPTIONS = ("2", "4", "8", "16", "t8", "t16")
FkSim = _lib.FkSim


class BatchSim(NamedTuple):
    """One member of a batch: the arguments of one _forward_euler call
    without t_end (cardiax/solve.py:93)."""
    state: object
    params: object
    diffusivity: object
    stimuli: Sequence
    t: float
    dt: float
    dx: float


class Batch:
    """Advances every simulation of ``sims`` together.

    ``states`` holds one ``State`` per simulation (device tensors owned by
    the batch, updated in place by ``run``); ``times`` the time of each
    simulation's next step.
    """

    def __init__(self, sims: Sequence[BatchSim]):
        dev = require_gpu("Batch")
        if len(sims) == 0:
            raise ValueError("empty batch")
        self._keep = []  # device tensors the descriptors point to
        self.states: List[State] = []
        self.times: List[float] = []
        arr = (FkSim * len(sims))()
        for i, s in enumerate(sims):
            st = State(*(on_device(x, dev).clone() for x in s.state))
            H, W = st.u.shape
            if any(tuple(x.shape) != (H, W) for x in st):
                raise ValueError("simulation %d: state fields differ in shape" % i)
            D = on_device(s.diffusivity, dev)
            if tuple(D.shape) != (H, W):
                raise ValueError("simulation %d: diffusivity has shape %s, expected %s"
                                 % (i, tuple(D.shape), (H, W)))
            stims = list(s.stimuli or [])
            if len(stims) > FK_BATCH_MAX_STIMULI:
                raise NotImplementedError("simulation %d: %d stimuli, the batch limit is %d"
                                          % (i, len(stims), FK_BATCH_MAX_STIMULI))
            d = arr[i]
            d.params = _lib.make_params(s.params)
            d.D = D.data_ptr()
            d.v, d.w, d.u = st.v.data_ptr(), st.w.data_ptr(), st.u.data_ptr()
            d.H, d.W = H, W
            d.n_stimuli = len(stims)
            for k, stim in enumerate(stims):
                f = on_device(stim[1], dev)
                if tuple(f.shape) != (H, W):
                    raise ValueError("simulation %d: stimulus %d field has the wrong shape"
                                     % (i, k))
                p = stim[0]
                d.stimuli[k] = _lib.FkStimulus(float(p[0]), float(p[1]), float(p[2]),
                                               f.data_ptr())
                self._keep.append(f)
            d.t = float(s.t)
            d.dt, d.dx = float(s.dt), float(s.dx)
            self._keep.append(D)
            self.states.append(st)
            self.times.append(float(s.t))
        self.cells = sum(int(x.u.numel()) for x in self.states)
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        _lib.check(_lib.lib().fk_batch_create(arr, len(sims), ctypes.byref(h)))
        self._h = h

    @property
    def n_tiles(self) -> int:
        return int(_lib.lib().fk_batch_n_items(self._h))

    def autotune(self, n_steps: int = 64) -> dict:
        """Time every available work-list option (bands of 2, 4, 8 or 16 rows
        advanced one step per launch; bands of 8 or 16 rows advanced two steps
        per launch when every simulation fits) and keep the fastest for later
        runs (fk_batch_autotune); the simulations' states are not touched.
        Returns ``{"rows": R, "two_step": bool, "times_ms": {option: ms}}``
        with options "2", "4", "8", "16", "t8", "t16" (unavailable ones left
        out)."""
        if self._h is None:
            raise RuntimeError("batch is closed")
        times = (ctypes.c_float * _lib.FK_BATCH_N_OPTIONS)()
        rows = ctypes.c_int32(0)
        _lib.check(_lib.lib().fk_batch_autotune(self._h, int(n_steps), stream_handle(), times,
                                            ctypes.byref(rows)))
        r = int(rows.value)
        return {"rows": r & ~_lib.FK_BATCH_TWO_STEP,
                "two_step": bool(r & _lib.FK_BATCH_TWO_STEP),
                "times_ms": {name: round(float(t), 4)
                             for name, t in zip(BATCH_OPTIONS, times) if t >= 0}}

    def set_rows(self, rows: int, two_step: bool = False) -> None:
        """Pin the work-list option of later runs (fk_batch_set_rows): a band
        height of 2, 4, 8 or 16 rows, or 8 or 16 with ``two_step``
        (NotImplementedError when some simulation does not fit it)."""
        if self._h is None:
            raise RuntimeError("batch is closed")
        flag = _lib.FK_BATCH_TWO_STEP if two_step else 0
        _lib.check(_lib.lib().fk_batch_set_rows(self._h, int(rows) | flag))

    def run(self, n_steps: int, use_graph: bool = True) -> List[State]:
        """Advance all simulations by ``n_steps`` Euler steps (asynchronous on
        torch's current stream)."""
        if self._h is None:
            raise RuntimeError("batch is closed")
        _lib.check(_lib.lib().fk_batch_run(self._h, int(n_steps), 1 if use_graph else 0,
                                       stream_handle()))
        self.times = [t + int(n_steps) for t in self.times]
        return self.states

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.check(_lib.lib().fk_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
