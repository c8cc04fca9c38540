"""CPU tests of the dataset generator's output path.

* The oracle's restatement of imresize (cardiax/io.py:118-124, jax.image.resize
  bilinear with its default antialiasing) is pinned by torch's antialiased
  bilinear interpolate (the same PIL-style triangle filter, an independent
  implementation) and by closed forms: exact copies on unchanged axes,
  constants preserved, the [1, 3, 3, 1] / 8 taps of halving (renormalised to
  [3, 3, 1] / 7 at the edge), linear ramps reproduced by upsampling.
* The storage functions of cardiax_amd.io write the reference's HDF5 layout
  (cardiax/io.py:13-115): checked with real h5py under the image's conda
  Python when present, and with a dict-backed stand-in here.
* fk_resize rejects bad arguments before any device work.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from cardiax_amd import _lib
from cardiax_amd import io as cio
from cardiax_amd.params import PARAMSET_3
from cardiax_amd.stimulus import Protocol, Stimulus
from oracle import io_oracle

import fake_h5py

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONDA_PY = "/opt/conda/bin/python3.9"

SHAPES = [(600, 600, 128, 128), (37, 53, 11, 7), (10, 10, 4, 4), (8, 8, 16, 16),
          (13, 9, 13, 20), (20, 30, 20, 30), (5, 7, 3, 7), (1200, 900, 256, 192),
          (64, 64, 1, 1), (1, 9, 1, 4), (7, 7, 21, 3)]


@pytest.mark.parametrize("H,W,h,w", SHAPES)
def test_oracle_resize_matches_antialiased_bilinear(H, W, h, w):
    rng = np.random.default_rng(H * 1000 + W)
    a = rng.standard_normal((3, H, W))
    got = io_oracle.imresize(a, (h, w))
    ref = F.interpolate(torch.from_numpy(a)[None], size=(h, w), mode="bilinear",
                        antialias=True, align_corners=False)[0].numpy()
    assert got.shape == (3, h, w)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


def test_oracle_resize_closed_forms():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((12, 16))
    # unchanged size: exact copy; constant: preserved
    assert np.array_equal(io_oracle.imresize(a, (12, 16)), a)
    np.testing.assert_allclose(io_oracle.imresize(np.full((30, 50), 2.5), (7, 11)), 2.5,
                               rtol=0, atol=1e-14)
    # halving: taps [1, 3, 3, 1] / 8, and [3, 3, 1] / 7 at the first output
    W2 = io_oracle.weight_matrix(16, 8)
    # (output o samples input 2o + 1/2: taps 2o-1 .. 2o+2)
    np.testing.assert_allclose(W2[3:7, 2], [1 / 8, 3 / 8, 3 / 8, 1 / 8], atol=1e-15)
    np.testing.assert_allclose(W2[0:3, 0], [3 / 7, 3 / 7, 1 / 7], atol=1e-15)
    assert np.all(W2[7:, 2] == 0) and np.all(W2[:3, 2] == 0)
    # upsampling reproduces a linear ramp wherever both taps are inside
    x = np.arange(10, dtype=np.float64)
    ramp = np.tile(3.0 * x - 2.0, (4, 1))
    up = io_oracle.imresize(ramp, (4, 40))
    s = (np.arange(40) + 0.5) * 10 / 40 - 0.5
    inside = (s >= 0) & (s <= 9)
    np.testing.assert_allclose(up[:, inside], np.tile(3.0 * s[inside] - 2.0, (4, 1)),
                               atol=1e-12)
    # every output's weights sum to 1
    for n_in, n_out in [(600, 128), (7, 21), (64, 1)]:
        np.testing.assert_allclose(io_oracle.weight_matrix(n_in, n_out).sum(0), 1.0,
                                   atol=1e-14)


def test_oracle_layout_counts_the_unwritten_frame():
    lay = io_oracle.layout(5, 2, (8, 9), PARAMSET_3._fields)
    assert lay["states"] == (5, 3, 8, 9) and lay["stimuli"] == (2, 8, 9)
    assert lay["field"] == (2, 8, 9) and lay["params/D"] == (8, 9)
    assert all(lay["params/" + f] == () for f in PARAMSET_3._fields)


def test_resize_arguments_rejected_before_device_work():
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.fk_resize(fake, 1, 0, 8, fake, 4, 4, fake, 1 << 20, None) == _lib.FK_ERR_ARG
    assert L.fk_resize(fake, 1, 8, 8, fake, 4, 0, fake, 1 << 20, None) == _lib.FK_ERR_ARG
    assert L.fk_resize(fake, -1, 8, 8, fake, 4, 4, fake, 1 << 20, None) == _lib.FK_ERR_ARG
    assert L.fk_resize(None, 1, 8, 8, fake, 4, 4, fake, 1 << 20, None) == _lib.FK_ERR_ARG
    need = L.fk_resize_workspace_bytes(2, 8, 8, 4, 4)
    assert need == 2 * 4 * 8 * 8
    assert L.fk_resize(fake, 2, 8, 8, ctypes.c_void_p(0x9000), 4, 4, fake, need - 1,
                       None) == _lib.FK_ERR_ARG
    assert b"workspace" in L.fk_last_error()
    # dst inside src
    assert L.fk_resize(ctypes.c_void_p(0x1000), 2, 8, 8, ctypes.c_void_p(0x1100), 4, 4,
                       ctypes.c_void_p(0x100000), need, None) == _lib.FK_ERR_ARG
    assert b"overlap" in L.fk_last_error()
    assert L.fk_resize(fake, 0, 8, 8, fake, 4, 4, fake, 0, None) == _lib.FK_OK


def test_imresize_refuses_without_gpu_and_other_methods():
    with pytest.raises(NotImplementedError):
        cio.imresize(np.zeros((4, 4)), (2, 2), method="nearest")
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            cio.imresize(np.zeros((4, 4)), (2, 2))


def _write_and_read_fake(tmp_path):
    cio._h5py = lambda: fake_h5py  # the module's only door to h5py
    path = str(tmp_path / "a.hdf5")
    rng = np.random.default_rng(3)
    D = rng.uniform(1e-4, 1e-3, (6, 7)).astype(np.float32)
    stims = [Stimulus(Protocol(0, 2, 1e9), rng.uniform(0, 20, (6, 7)).astype(np.float32)),
             Stimulus(Protocol(500, 2, 400), np.zeros((6, 7), np.float32))]
    f = cio.init(path, (6, 7), n_iter=4, n_stimuli=2)
    cio.add_params(f, PARAMSET_3, D, 0.01, 0.02)
    cio.add_stimuli(f, stims)
    cio.add_diffusivity(f, D)
    frames = rng.standard_normal((3, 3, 6, 7)).astype(np.float32)
    cio.add_states(f["states"], list(frames[:2]), 0, 2)
    cio.add_state(f["states"], frames[2], 2)
    f.close()
    return path, D, stims, frames


def test_storage_layout_with_stand_in(tmp_path, monkeypatch):
    monkeypatch.setattr(cio, "_h5py", cio._h5py)
    path, D, stims, frames = _write_and_read_fake(tmp_path)
    f = fake_h5py.File(path, "r")
    flat = f.flat()
    lay = io_oracle.layout(4, 2, (6, 7), PARAMSET_3._fields)
    assert set(flat) == set(lay)
    for k, shp in lay.items():
        assert np.shape(flat[k]) == shp, k
    assert flat["states"].dtype == np.float32
    np.testing.assert_array_equal(flat["states"][:3], frames)
    assert not flat["states"][3].any()  # the frame the generator never writes
    assert not flat["stimuli"].any()    # created, never written
    np.testing.assert_array_equal(flat["params/D"], D)
    np.testing.assert_array_equal(flat["diffusivity"], D)
    assert float(flat["params/dx"]) == 0.02
    np.testing.assert_array_equal(flat["start"], [0, 500])
    np.testing.assert_array_equal(flat["period"], [1e9, 400])
    # loaders (cardiax/io.py:90-115)
    got = cio.load_stimuli(f)
    assert [float(s.protocol.start) for s in got] == [0.0, 500.0]
    np.testing.assert_array_equal(got[0].field, stims[0].field)
    prm, D2 = cio.load_params(path)
    np.testing.assert_array_equal(D2, D)
    assert float(prm[0]) == PARAMSET_3[0]
    np.testing.assert_array_equal(cio.load_diffusivity(path), D)
    np.testing.assert_array_equal(cio.load_state(f["states"], 0, 3, 2), frames[0:3:2])
    assert len(cio.load(path)) == 7  # the top-level datasets; the params group is skipped


_CONDA_SCRIPT = r"""
import importlib.util, json, sys
from collections import namedtuple
import numpy as np
spec = importlib.util.spec_from_file_location("cio", sys.argv[1])
cio = importlib.util.module_from_spec(spec); spec.loader.exec_module(cio)
import h5py
fields = json.loads(sys.argv[3])
P = namedtuple("Params", fields)(*[float(i + 1) for i in range(len(fields))])
class Protocol(tuple):
    start = property(lambda s: s[0]); duration = property(lambda s: s[1]); period = property(lambda s: s[2])
class Stim:
    def __init__(self, p, f): self.protocol, self.field = p, f
rng = np.random.default_rng(3)
D = rng.uniform(1e-4, 1e-3, (6, 7)).astype(np.float32)
stims = [Stim(Protocol((0, 2, 1e9)), rng.uniform(0, 20, (6, 7)).astype(np.float32))]
f = cio.init(sys.argv[2], (6, 7), n_iter=4, n_stimuli=1)
cio.add_params(f, P, D, 0.01, 0.02); cio.add_stimuli(f, stims); cio.add_diffusivity(f, D)
frames = rng.standard_normal((3, 3, 6, 7)).astype(np.float32)
cio.add_states(f["states"], list(frames[:2]), 0, 2); cio.add_state(f["states"], frames[2], 2)
f.close()
out = {}
with h5py.File(sys.argv[2], "r") as g:
    g.visititems(lambda k, v: out.__setitem__(k, list(v.shape)) if isinstance(v, h5py.Dataset) else None)
    ok = bool(np.array_equal(g["states"][:3], frames) and not g["states"][3].any()
              and np.array_equal(g["params/D"][...], D) and g["states"].dtype == np.float32
              and float(g["params/dx"][()]) == 0.02)
loaded = cio.load(sys.argv[2])
print(json.dumps({"shapes": out, "ok": ok, "n_loaded": len(loaded),
                  "diff": cio.load_diffusivity(sys.argv[2]).tolist() == D.tolist()}))
"""


@pytest.mark.skipif(not os.path.exists(CONDA_PY), reason="no conda interpreter with h5py")
def test_storage_layout_with_real_h5py(tmp_path):
    probe = subprocess.run([CONDA_PY, "-c", "import h5py, numpy"], capture_output=True)
    if probe.returncode != 0:
        pytest.skip("conda interpreter lacks h5py")
    script = tmp_path / "w.py"
    script.write_text(_CONDA_SCRIPT)
    fields = list(PARAMSET_3._fields)
    res = subprocess.run([CONDA_PY, str(script), os.path.join(ROOT, "cardiax_amd", "io.py"),
                          str(tmp_path / "real.hdf5"), json.dumps(fields)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = json.loads(res.stdout.strip().splitlines()[-1])
    assert got["ok"] and got["diff"]
    lay = io_oracle.layout(4, 1, (6, 7), fields)
    assert {k: tuple(v) for k, v in got["shapes"].items()} == lay
    # top-level datasets: states stimuli field start duration period diffusivity
    assert got["n_loaded"] == 7


def test_sequence_needs_gpu_and_refuses_plots():
    from cardiax_amd import generate

    with pytest.raises(NotImplementedError):
        generate.sequence(0, 10, 5, 0.01, 0.01, PARAMSET_3, np.ones((8, 8)), [], None,
                          plot_while=True)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            generate.sequence(0, 10, 5, 0.01, 0.01, PARAMSET_3, np.ones((8, 8)), [], None,
                              plot_while=False)
