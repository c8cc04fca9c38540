"""Mirror of deepx/generate.py: the dataset generator, the production caller
of the forward-Euler path (``sequence``, deepx/generate.py:135-210), and the
random inputs ``random_sequence`` draws for it (:16-132).

The reference draws from ``jax.random`` keys, which are absent here, so the
random draws are project-defined: a seed (or NumPy Generator or
SeedSequence) takes the place of the key and ``SeedSequence.spawn`` the place
of ``jax.random.split``. The distributions, ranges and quirks are the
reference's: the rectangle size and centre both follow ``shape[0]``, the
linear and triangular directions are drawn from {NORTH, EAST, SOUTH} (never
WEST), the coverage and angle come from the same key as the direction,
``random_sequence`` reuses the last stimulus key for the D-map, and an
empty integer range returns its lower bound as ``jax.random.randint`` does.
The D-map is the reference's spline scar (deepx/utils_scars.py:246-426,
``scars.random_diffusivity``): contours on the host, rasterisation, blur and
rescale onto ``domain`` on the GPU.

The reference runs ``_forward_euler`` per checkpoint interval, pulls each
state to the host and resizes it there with ``jax.image.resize`` before
writing it to HDF5. Here the state stays in HBM for the whole run, and each
checkpoint is resized on the GPU (``fk_resize``) into one of two device frame
buffers. The frame is then copied to one of two pinned host slots on a copy
stream, while the next interval's kernels are already queued. The host moves
frame i - 1 out of its slot (into the returned array, or into the file) while
the GPU works on interval i.
"""
from __future__ import annotations

import numpy as np
import torch

from . import convert, io, scars, solve, stimulus
from ._device import require_gpu
from ._keys import _randint, _seq, _split  # noqa: F401  (jax.random stand-ins)


def random_protocol(rng, min_start=0, max_start=1000, min_period=400, max_period=1e9):
    """deepx/generate.py:16-27 -- start in [min_start, max_start), duration 2,
    period in [min_period, max_period)."""
    g1, g2 = _split(rng, 2)
    return stimulus.Protocol(_randint(g1, min_start, max_start), 2,
                             _randint(g2, min_period, max_period))


def random_rectangular_stimulus(rng, shape, protocol, modulus=0.6):
    """deepx/generate.py:30-36 -- sizes in [max(shape[0] // 100, 10),
    shape[0] // 3), centre in [min(size), shape[0]) on both axes."""
    g1, g2 = _split(rng, 2)
    size = _randint(g2, max(shape[0] // 100, 10), shape[0] // 3, size=2)
    centre = _randint(g1, int(np.min(size)), shape[0], size=2)
    return stimulus.rectangular(shape, tuple(int(x) for x in centre),
                                tuple(int(x) for x in size), modulus, protocol)


def random_linear_stimulus(rng, shape, protocol, modulus=0.6):
    """deepx/generate.py:39-45 -- coverage |N(0, 1)| * 0.2, direction in
    {NORTH, EAST, SOUTH}."""
    g1, _ = _split(rng, 2)
    coverage = abs(float(g1.normal()))
    direction = stimulus.Direction(_randint(g1, 0, 3))
    return stimulus.linear(shape, direction, coverage * 0.2, modulus, protocol)


def random_triangular_stimulus(rng, shape, protocol, modulus=0.6):
    """deepx/generate.py:48-56 -- angle |N(0, 1)| * 45 degrees, coverage
    |N(0, 1)| * 0.2, direction in {NORTH, EAST, SOUTH}."""
    g1, _ = _split(rng, 2)
    angle, coverage = np.abs(g1.normal(size=2))
    direction = stimulus.Direction(_randint(g1, 0, 3))
    return stimulus.triangular(shape, direction, float(angle) * 45, float(coverage) * 0.2,
                               modulus, protocol)


def random_stimulus(rng, shape, min_start=0, max_start=0):
    """deepx/generate.py:59-76 -- a random protocol and one of the
    rectangular, triangular and linear builders (uniformly), modulus 20."""
    g1, g2, g3 = _split(rng, 3)
    protocol = random_protocol(g1, min_start=min_start, max_start=max_start)
    builders = (random_rectangular_stimulus, random_triangular_stimulus, random_linear_stimulus)
    return builders[_randint(g2, 0, len(builders))](g3, shape, protocol, 20.0)


def random_diffusivity(rng, shape, domain=(0.0001, 0.001)):
    """deepx/generate.py:79-83 -- the random spline scar, inverted and
    rescaled onto ``domain`` (cardiax/convert.py:23-26): its minimum becomes
    domain[0], its maximum domain[1]. A float32 tensor on the current GPU."""
    return scars.random_diffusivity(rng, tuple(int(x) for x in shape), domain)


def random_sequence(rng, params, filepath, shape=(1200, 1200), n_stimuli=2, start=0,
                    stop=1000, step=1, dt=0.01, dx=0.01, reshape=None, use_memory=False,
                    plot_while=True):
    """deepx/generate.py:86-132 -- n_stimuli random stimuli, the i-th starting
    at step 1 + i * ms_to_units(tau_d * 1000, dt), a random D-map, and
    ``sequence`` over [start, stop) ms with a checkpoint every ``step`` ms."""
    seqs = _seq(rng).spawn(n_stimuli)
    max_start = np.arange(1, convert.ms_to_units(stop, dt),
                          convert.ms_to_units(params.tau_d * 1000, dt))
    stimuli = [random_stimulus(np.random.default_rng(seqs[i]), shape,
                               min_start=max_start[i], max_start=max_start[i] + 1)
               for i in range(n_stimuli)]
    # the reference draws the D-map from the last stimulus key again
    diffusivity = random_diffusivity(np.random.default_rng(seqs[-1]), shape)
    return sequence(start=convert.ms_to_units(start, dt), stop=convert.ms_to_units(stop, dt),
                    step=convert.ms_to_units(step, dt), dt=dt, dx=dx, params=params,
                    diffusivity=diffusivity, stimuli=stimuli, filename=filepath,
                    reshape=reshape, use_memory=use_memory, plot_while=plot_while)


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
    try:
        if hdf5 is not None:
            io.add_params(hdf5, params, D, dt, dx, shape=out_shape)
            io.add_stimuli(hdf5, stim, shape=out_shape)
            io.add_diffusivity(hdf5, D, shape=out_shape)
        return _sequence_frames(state_args=(params, D, stim, dt, dx), checkpoints=checkpoints,
                                shape=shape, out_shape=out_shape, hdf5=hdf5,
                                use_memory=use_memory, dev=dev)
    finally:
        if hdf5 is not None:
            hdf5.close()


def _sequence_frames(state_args, checkpoints, shape, out_shape, hdf5, use_memory, dev):
    """The checkpoint loop of ``sequence``: frames to ``hdf5`` (or kept and
    returned), the file's closing left to the caller."""
    params, D, stim, dt, dx = state_args
    n = max(len(checkpoints) - 1, 0)
    h, w = out_shape
    keep = hdf5 is None or use_memory  # every frame stays on the host
    # Two pinned slots carry the frames off the device; a finished frame is
    # copied into `frames` (pageable) or written to the file while the GPU
    # runs the next interval.
    host = torch.empty((2, 3, h, w), dtype=torch.float32, pin_memory=True)
    frames = np.empty((n, 3, h, w), dtype=np.float32) if keep else None

    def retire(k):  # frame k has reached its pinned slot
        copied[k % 2].synchronize()
        if keep:
            frames[k] = host[k % 2].numpy()
        else:
            io.add_state(hdf5["states"], host[k % 2].numpy(), k)

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
        with torch.cuda.stream(copy):  # slot i % 2 was retired at the previous interval
            host[i % 2].copy_(buf, non_blocking=True)
        done = torch.cuda.Event()
        done.record(copy)
        copied[i % 2] = done
        if i > 0:  # retire the previous frame while the GPU runs this interval
            retire(i - 1)
    if n > 0:
        retire(n - 1)
    torch.cuda.synchronize()
    if hdf5 is not None and use_memory:
        io.add_states(hdf5["states"], frames, 0, n)
    print()
    return frames


def io_workspace(shape, out_shape) -> int:
    """Bytes of fk_resize workspace for one H x W -> h x w field."""
    from . import _lib

    return int(_lib.lib().fk_resize_workspace_bytes(1, shape[0], shape[1], out_shape[0],
                                                     out_shape[1]))
