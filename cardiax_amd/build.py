"""Compile ``libcardiax_amd.so`` for gfx950 in-tree with hipcc.

Usage: ``python -m cardiax_amd.build``. Each source is compiled to an object
file in parallel (no device code crosses a translation unit), then the
objects are linked into the library. The library is written next to this
file (the ctypes loader looks for it there) via a temporary file and an atomic
rename, so a process that loads it never sees a partial file.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcardiax_amd.so")

SOURCES = ["fk_common.hip", "fk_forward.hip", "fk_batch.hip", "fk_adjoint.hip", "fk_io.hip",
           "fk_stimulus.hip", "fk_noalloc.hip", "fk_scar.hip", "fk_rollout_batch.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Keep the compiler from forming packed fp32 (v_pk_*_f32) instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

# Per-file additions. The two-step streaming kernels fully unroll band loops of
# several thousand instructions, past LLVM's default size limit for
# `#pragma unroll`; a loop left rolled would index its register windows at run
# time and move them to scratch memory.
FILE_FLAGS = {"fk_forward.hip": ["-mllvm", "-pragma-unroll-threshold=1000000"],
              "fk_batch.hip": ["-mllvm", "-pragma-unroll-threshold=1000000"],
              "fk_rollout_batch.hip": ["-mllvm", "-pragma-unroll-threshold=1000000"]}

# The host pass does not know the packed-fp32 feature and says so.
_NOISE = "'-packed-fp32-ops' is not a recognized feature"


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if _NOISE not in l)
    return res.returncode, err


def build(verbose: bool = False) -> str:
    tmpdir = tempfile.mkdtemp(prefix=".objs.", dir=HERE)
    fd, tmp = tempfile.mkstemp(prefix=".libcardiax_amd.", suffix=".so.tmp", dir=HERE)
    os.close(fd)
    objs = [os.path.join(tmpdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), *HIPCC_FLAGS, *FILE_FLAGS.get(s, []), "-I", INCLUDE, "-I", CSRC, "-c",
             os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    try:
        workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
        with ThreadPoolExecutor(max_workers=workers) as ex:
            results = list(ex.map(_run, cmds))
        for src, (rc, err) in zip(SOURCES, results):
            if rc != 0:
                raise RuntimeError("hipcc failed on %s (%d):\n%s" % (src, rc, err))
            if verbose and err.strip():
                print(err, file=sys.stderr)
        rc, err = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
        if rc != 0:
            raise RuntimeError("link failed (%d):\n%s" % (rc, err))
        os.replace(tmp, OUT)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
        for o in objs:
            if os.path.exists(o):
                os.unlink(o)
        os.rmdir(tmpdir)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
