"""The C ABI from C itself: a C99 program that includes include/cardiax_amd.h,
links libcardiax_amd.so and calls the entry points that need no GPU (sizes,
version, argument errors and their messages). This is what a cgo / JNI / FFI
binding of the library compiles against, with no Python or torch involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cardiax_amd")

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "cardiax_amd.h"

int main(void) {
  int bad = 0;
  if (fk_abi_version() < 1) { puts("abi version"); bad = 1; }
  if (fk_forward_workspace_bytes(128, 64) != (size_t)(3 * 128 * 64 * 4 + 16 * 128 * 64)) {
    puts("workspace bytes"); bad = 1;
  }
  if (fk_resize_workspace_bytes(3, 600, 600, 128, 128) != (size_t)3 * 128 * 600 * 8) {
    puts("resize workspace"); bad = 1;
  }
  fk_params p;
  memset(&p, 0, sizeof p);
  p.tau_v_plus = 3.33f; p.tau_v1_minus = 19.6f; p.tau_v2_minus = 1250.f; p.tau_w_plus = 870.f;
  p.tau_w_minus = 41.f; p.tau_d = 0.25f; p.tau_0 = 12.5f; p.tau_r = 33.33f; p.tau_si = 29.f;
  p.k = 10.f; p.V_csi = 0.85f; p.V_c = 0.13f; p.V_v = 0.04f; p.Cm = 1.f;
  float* fake = (float*)0x1000;
  /* a 2-row grid is refused before any device work, with a message */
  if (fk_step(&p, fake, fake, fake, fake, 2, 8, NULL, 0, 0.0, 0.01f, fake, fake, fake, NULL)
      != FK_ERR_ARG || strlen(fk_last_error()) == 0) { puts("grid check"); bad = 1; }
  /* a successful query clears the thread's message */
  (void)fk_forward_variant_for(64, 64);
  if (strlen(fk_last_error()) != 0) { puts("error not cleared"); bad = 1; }
  if (fk_forward_variant_for(1024, 1024) != 15) { puts("default variant"); bad = 1; }
  if (fk_check_finite(NULL, 1, 10, fake, NULL) != FK_ERR_ARG) { puts("finite args"); bad = 1; }
  if (FK_ERR_NONFINITE != -4 || FK_MAX_STIMULI != 32) { puts("constants"); bad = 1; }
  printf("%s\n", bad ? "FAIL" : "OK");
  return bad;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_header_compiles_as_c99_and_links(tmp_path):
    lib = os.path.join(LIBDIR, "libcardiax_amd.so")
    if not os.path.exists(lib):
        pytest.fail("libcardiax_amd.so is not built")
    src = tmp_path / "abi.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "abi"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", LIBDIR, "-lcardiax_amd", "-Wl,-rpath," + LIBDIR,
           "-Wl,-rpath," + os.path.join(rocm, "lib")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip().endswith("OK"), run.stdout + run.stderr
