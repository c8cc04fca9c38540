"""Mirror of cardiax/io.py: the HDF5 layout the dataset generator writes, and
``imresize`` on the GPU.

Storage (``init``, ``add_*``, ``load*``) keeps the reference's dataset names,
shapes and dtypes (cardiax/io.py:13-115) and uses h5py, the reference's own
dependency, imported when a function first needs it. These functions take
host arrays (NumPy or anything ``np.asarray`` accepts, including torch
tensors), and the module body imports only NumPy, so the storage functions
that need no resize run wherever h5py and NumPy do.

``imresize`` (cardiax/io.py:118-124) runs ``fk_resize`` on the GPU and returns
a float32 torch tensor on ``cuda:0``. Differences from the reference, where it
would fail instead: ``add_state`` resizes the state to the output image size
(the reference passes ``(3, h, w)`` as the size, which ``imresize`` prefixes
with the leading axis and jax rejects), and only the ``"bilinear"`` method
(the reference's default and only one used) is built.
"""
from __future__ import annotations

import os

import numpy as np


def _h5py():
    try:
        import h5py
    except ImportError as e:  # the reference fails at import time without it
        raise ImportError("cardiax_amd.io storage needs h5py, which is not installed "
                          "in this interpreter") from e
    return h5py


def _host(a):
    """Host NumPy view of an array, a torch tensor (any device) or a scalar."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _sized(a, shape):
    """``a`` on the host, resized to ``shape`` (on the GPU) when given."""
    if shape is None:
        return _host(a)
    return _host(imresize(a, tuple(shape)))


def init(path, shape, n_iter, n_stimuli, n_variables=3):
    """cardiax/io.py:13-23 -- open ``path`` for writing (truncating) with the
    ``states`` (n_iter, n_variables, *shape) and ``stimuli`` (n_stimuli,
    *shape) float32 datasets."""
    h5py = _h5py()
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    hdf5 = h5py.File(path, "w")
    if "states" not in hdf5:
        hdf5.create_dataset("states", shape=(n_iter, n_variables, *shape), dtype="float32")
    if "stimuli" not in hdf5:
        hdf5.create_dataset("stimuli", shape=(n_stimuli, *shape), dtype="float32")
    return hdf5


def add_params(hdf5, params, diffusivity, dt, dx, shape=None):
    """cardiax/io.py:26-36 -- ``params/D`` (resized to ``shape`` if given),
    ``params/dt``, ``params/dx`` and one scalar per Params field."""
    hdf5.create_dataset("params/D", data=_sized(diffusivity, shape))
    hdf5.create_dataset("params/dt", data=dt)
    hdf5.create_dataset("params/dx", data=dx)
    for i in range(len(params)):
        hdf5.create_dataset("params/" + params._fields[i], data=params[i])
    return True


def add_diffusivity(hdf5, diffusivity, shape=None):
    """cardiax/io.py:39-44."""
    hdf5.create_dataset("diffusivity", data=_sized(diffusivity, shape))
    return True


def add_stimuli(hdf5, stimuli, shape=None):
    """cardiax/io.py:47-64 -- ``field`` (resized if ``shape``) and the
    protocol arrays ``start``, ``duration``, ``period``."""
    fields = [_sized(stimuli[i].field, shape) for i in range(len(stimuli))]
    hdf5.create_dataset("field", data=fields)
    hdf5.create_dataset("start", data=[_host(s.protocol.start) for s in stimuli])
    hdf5.create_dataset("duration", data=[_host(s.protocol.duration) for s in stimuli])
    hdf5.create_dataset("period", data=[_host(s.protocol.period) for s in stimuli])
    return True


def add_state(dset, state, t, shape=None):
    """cardiax/io.py:67-72 -- frame ``t`` of ``dset``. ``shape`` is the frame
    shape (n_variables, h, w), as the generator passes it; the state is
    resized to (h, w)."""
    if shape is not None:
        state = imresize(_stack(state), tuple(shape)[-2:])
    dset[t] = _host(_stack(state))
    return True


def add_states(dset, states, start, end):
    """cardiax/io.py:75-78 -- frames start..end-1, shape (t, 3, h, w)."""
    dset[start:end] = np.stack([_host(_stack(s)) for s in states]) if len(states) else states
    return True


def load(path, start=None, end=None, step=None):
    """cardiax/io.py:81-83 -- every top-level dataset sliced [start:end:step],
    in the file's key order. The reference also slices the ``params`` group,
    which h5py refuses; groups are skipped here."""
    h5py = _h5py()
    with h5py.File(path, "r") as file:
        return [file[k][start:end:step] for k in file if not isinstance(file[k], h5py.Group)]


def load_state(dset, start, end, step):
    """cardiax/io.py:86-87."""
    return dset[start:end:step]


def load_stimuli(file):
    """cardiax/io.py:90-96."""
    from .stimulus import Protocol, Stimulus

    stimuli = []
    for i in range(len(file["field"])):
        protocol = Protocol(file["start"][i], file["duration"][i], file["period"][i])
        stimuli.append(Stimulus(protocol, file["field"][i]))
    return stimuli


def load_params(filepath):
    """cardiax/io.py:99-110 -- (Params, D). The reference passes every other
    ``params/*`` entry positionally, in the file's (alphabetical) key order
    and including ``dt`` and ``dx``, so it fails on the files ``add_params``
    writes; here the Params fields are read by name."""
    from .params import Params

    params, D = {}, None
    with _h5py().File(filepath, "r") as file:
        stored = file["params"]
        for key in stored:
            if key == "D":
                D = stored[key][...]
            else:
                params[key] = stored[key][...]
    return Params(*[params[f] for f in Params._fields]), D


def load_diffusivity(filepath):
    """cardiax/io.py:113-115."""
    with _h5py().File(filepath, "r") as f:
        return f["diffusivity"][:]


def _stack(state):
    """A State (or any sequence of equal-shape fields) as one array."""
    if hasattr(state, "_fields") or isinstance(state, (list, tuple)):
        parts = list(state)
        if parts and hasattr(parts[0], "detach"):
            import torch

            return torch.stack([p for p in parts])
        return np.stack([_host(p) for p in parts])
    return state


def imresize(a, size, method="bilinear", out=None, work=None):
    """cardiax/io.py:118-124 -- ``jax.image.resize(a, a.shape[:-2] + size,
    "bilinear")`` (antialiased) on the GPU: the last two axes of ``a``
    resampled to ``size``. Returns a float32 torch tensor on ``cuda:0``.
    ``out`` (float32, contiguous, on the device) and ``work`` (a device
    tensor of at least ``fk_resize_workspace_bytes``) may be passed to reuse
    memory."""
    import torch

    from . import _lib
    from ._device import require_gpu, stream_handle

    if method not in ("bilinear", "linear"):
        raise NotImplementedError("imresize: only the bilinear method is built")
    dev = require_gpu("imresize")
    x = torch.as_tensor(_stack(a), dtype=torch.float32)
    if x.dim() < 2:
        raise ValueError("imresize: needs at least 2 dimensions, got shape %s"
                         % (tuple(x.shape),))
    x = x.to(dev).contiguous()
    h, w = (int(v) for v in size)
    H, W = x.shape[-2:]
    lead = tuple(x.shape[:-2])
    n = int(np.prod(lead)) if lead else 1
    if out is None:
        out = torch.empty(lead + (h, w), dtype=torch.float32, device=dev)
    elif tuple(out.shape) != lead + (h, w) or out.dtype != torch.float32 \
            or not out.is_contiguous() or out.device != x.device:
        raise ValueError("imresize: out must be a contiguous float32 device tensor of shape %s"
                         % (lead + (h, w),))
    L = _lib.lib()
    nbytes = int(L.fk_resize_workspace_bytes(n, H, W, h, w))
    if work is not None and (work.device != x.device or not work.is_contiguous()):
        raise ValueError("imresize: work must be a contiguous tensor on %s" % (x.device,))
    if work is None or work.numel() * work.element_size() < nbytes:
        work = torch.empty(max(1, nbytes), dtype=torch.uint8, device=dev)
    _lib.check(L.fk_resize(x.data_ptr(), n, H, W, out.data_ptr(), h, w, work.data_ptr(),
                           nbytes, stream_handle()))
    return out


__all__ = ["init", "add_params", "add_diffusivity", "add_stimuli", "add_state", "add_states",
           "load", "load_state", "load_stimuli", "load_params", "load_diffusivity",
           "imresize"]
