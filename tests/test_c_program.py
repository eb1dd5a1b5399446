"""CPU tests of the C boundary as a C or C++ caller sees it: include/minhashcuda.h
compiles as strict C99 and C++11, and a C program linked against libMHCUDA.so
(the link line of the C integration example) runs the reference's host-side
argument checks (minhashcuda.cc:217-225, 587-589) and the no-device path."""
import os
import shutil
import subprocess

import pytest

from conftest import LIB_PATH, ROOT

INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.dirname(LIB_PATH)

PROGRAM = r"""
#include <stdio.h>
#include "minhashcuda.h"

int main(void) {
  MHCUDAResult st = mhcudaRuntimeError;
  float f[4] = {0};
  uint32_t u[4] = {0};
  MinhashCudaGenerator *g = mhcuda_init(0, 128, 1, 0, 0, 0, &st);
  MinhashCudaGeneratorParameters p = mhcuda_get_parameters(NULL);
  if (g != NULL || st != mhcudaInvalidArguments) return 1;
  if (p.dim != 0 || p.samples != 0 || p.verbosity != 0) return 2;
  if (mhcuda_fini(NULL) != mhcudaSuccess) return 3;
  if (mhcuda_calc(NULL, f, u, u, 1, u) != mhcudaInvalidArguments) return 4;
  if (mhcuda_assign_random_vars(NULL, f, f, f) != mhcudaInvalidArguments) return 5;
  st = mhcudaSuccess;
  g = mhcuda_init(10, 128, 1, 0, 0, 0, &st);
  if (g != NULL) {
    printf("device\n");
    return mhcuda_fini(g) == mhcudaSuccess ? 0 : 6;
  }
  printf("status %d\n", (int)st);
  return st == mhcudaNoSuchDevice ? 0 : 7;
}
"""


def compiler(name):
    path = shutil.which(name)
    if path is None:
        pytest.skip("%s not found" % name)
    return path


@pytest.mark.parametrize("cc,std,lang", [("gcc", "-std=c99", "c"), ("g++", "-std=c++11", "c++")])
def test_header_compiles_strictly(tmp_path, cc, std, lang):
    src = tmp_path / ("h." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "minhashcuda.h"\nint main(void) { return 0; }\n')
    r = subprocess.run([compiler(cc), std, "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-fsyntax-only", "-I", INCLUDE, str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_c_program_links_and_runs(tmp_path):
    src = tmp_path / "app.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "app"
    r = subprocess.run([compiler("gcc"), "-std=c99", "-Wall", "-Werror", "-I", INCLUDE,
                        str(src), "-o", str(exe), "-L", LIB_DIR, "-lMHCUDA",
                        "-Wl,-rpath," + LIB_DIR],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
