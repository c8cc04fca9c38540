"""Compile ``libcardiax_amd.so`` for gfx950 in-tree with hipcc.

Usage: ``python -m cardiax_amd.build``. The library is written next to this
file (the ctypes loader looks for it there) via a temporary file and an atomic
rename, so a process that loads it never sees a partial file.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcardiax_amd.so")

SOURCES = ["fk_common.hip", "fk_forward.hip", "fk_batch.hip", "fk_adjoint.hip", "fk_io.hip",
           "fk_stimulus.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Keep the compiler from forming packed fp32 (v_pk_*_f32) instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    fd, tmp = tempfile.mkstemp(prefix=".libcardiax_amd.", suffix=".so.tmp", dir=HERE)
    os.close(fd)
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-o", tmp]
    try:
        res = subprocess.run(cmd, capture_output=True, text=True)
        # The host pass does not know the packed-fp32 feature and says so.
        noise = "'-packed-fp32-ops' is not a recognized feature"
        err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
        if res.returncode != 0:
            raise RuntimeError("hipcc failed (%d):\n%s" % (res.returncode, err))
        if verbose and err.strip():
            print(err, file=sys.stderr)
        os.replace(tmp, OUT)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
