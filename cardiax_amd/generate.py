"""Mirror of deepx.generate.sequence (deepx/generate.py:135-210): the dataset
generator, the production caller of the forward-Euler path.

The reference runs ``_forward_euler`` per checkpoint interval, pulls each
state to the host and resizes it there with ``jax.image.resize`` before
writing it to HDF5. Here the state stays in HBM for the whole run, and each
checkpoint is resized on the GPU (``fk_resize``) into one of two device frame
buffers. The frame is then copied to pinned host memory on a copy stream,
while the next interval's kernels are already queued. The host writes frame
i - 1 (h5py) while the GPU works on interval i.
"""
from __future__ import annotations

import numpy as np
import torch

from . import convert, io, solve
from ._device import require_gpu


def sequence(start, stop, step, dt, dx, params, diffusivity, stimuli, filename,
             reshape=None, use_memory=False, plot_while=True):
    """deepx/generate.py:135-210. Checkpoints are ``range(start, stop, step)``
    in steps; the state after each of the ``len(checkpoints) - 1`` intervals
    is resized to ``reshape`` (default: the grid shape) and written to frame
    i of the ``states`` dataset of ``filename``. As in the reference, the file
    holds ``len(checkpoints)`` frames (the last stays zero), ``params/*``,
    ``diffusivity`` and the stimuli, all at the output shape; ``use_memory``
    writes the frames in one call at the end instead of one per checkpoint.

    With ``use_memory`` or ``filename=None`` (no file; this package's
    addition) it returns the frames as a float32 NumPy array (n, 3, h, w);
    otherwise it writes frame by frame through two pinned host slots and
    returns None, as the reference does. ``plot_while`` (matplotlib plots) is
    not built and must be False."""
    if plot_while:
        raise NotImplementedError("plot_while: plotting is not part of this package")
    dev = require_gpu("sequence")
    D = solve._dev_field(diffusivity, dev, tuple(np.shape(diffusivity)), "diffusivity")
    shape = tuple(D.shape)
    out_shape = tuple(int(x) for x in reshape) if reshape is not None else shape
    stim = [type(s)(s[0], solve._dev_field(s[1], dev, shape, "stimulus field"))
            if hasattr(s, "_fields") else (s[0], solve._dev_field(s[1], dev, shape, "field"))
            for s in (stimuli or [])]
    checkpoints = list(range(int(start), int(stop), int(step)))

    print("Tissue size is: {} - Computing on grid {}".format(
        convert.shape_to_realsize(shape, dx), shape))
    print("Checkpointing every {} steps".format(step))
    print("Cell parameters", params)

    hdf5 = None
    if filename is not None:
        hdf5 = io.init(filename, out_shape, n_iter=len(checkpoints), n_stimuli=len(stim))
        io.add_params(hdf5, params, D, dt, dx, shape=out_shape)
        io.add_stimuli(hdf5, stim, shape=out_shape)
        io.add_diffusivity(hdf5, D, shape=out_shape)

    n = max(len(checkpoints) - 1, 0)
    h, w = out_shape
    keep = hdf5 is None or use_memory  # every frame stays on the host
    host = torch.empty((n if keep else 2, 3, h, w), dtype=torch.float32, pin_memory=True)
    dev_frames = [torch.empty((3, h, w), dtype=torch.float32, device=dev) for _ in range(2)]
    work = torch.empty(max(1, io_workspace(shape, out_shape)), dtype=torch.uint8, device=dev)
    compute = torch.cuda.current_stream()
    copy = torch.cuda.Stream()
    copied = [None, None]  # copy-done event of the frame each buffer last held

    state = solve.init(shape)
    for i in range(n):
        print("Solving at: %dms/%dms\t\t" % (convert.units_to_ms(checkpoints[i + 1], dt),
                                             convert.units_to_ms(checkpoints[-1], dt)),
              end="\r")
        state = solve._forward_euler(state, checkpoints[i], checkpoints[i + 1], params, D,
                                     stim, dt, dx)
        buf = dev_frames[i % 2]
        if copied[i % 2] is not None:  # the buffer's previous frame has reached the host
            compute.wait_event(copied[i % 2])
        for k in range(3):
            io.imresize(state[k], out_shape, out=buf[k], work=work)
        ready = torch.cuda.Event()
        ready.record(compute)
        copy.wait_event(ready)
        slot = i if keep else i % 2  # a slot is reused after its frame was written
        with torch.cuda.stream(copy):
            host[slot].copy_(buf, non_blocking=True)
        done = torch.cuda.Event()
        done.record(copy)
        copied[i % 2] = done
        if not keep and i > 0:  # write the previous frame while the GPU runs this one
            copied[(i - 1) % 2].synchronize()
            io.add_state(hdf5["states"], host[(i - 1) % 2].numpy(), i - 1)
    torch.cuda.synchronize()
    if hdf5 is not None:
        if use_memory:
            io.add_states(hdf5["states"], host.numpy(), 0, n)
        elif n > 0:
            io.add_state(hdf5["states"], host[(n - 1) % 2].numpy(), n - 1)
    print()
    if hdf5 is not None:
        hdf5.close()
    return host.numpy() if keep else None


def io_workspace(shape, out_shape) -> int:
    """Bytes of fk_resize workspace for one H x W -> h x w field."""
    from . import _lib

    return int(_lib.lib().fk_resize_workspace_bytes(1, shape[0], shape[1], out_shape[0],
                                                     out_shape[1]))
