"""Builds libMHCUDA.so (HIP kernels for gfx950 + host runtime + CPython module).

The library is built in-tree, next to this file, so that it travels with the
source tree.  Every translation unit is compiled with hipcc and

  -ffp-contract=off -fno-fast-math   the hash formula must stay unfused FP32
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                      no packed FP32 VALU instructions

Usage:  python -m minhashcuda_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.environ.get("MHCUDA_BUILD_DIR", os.path.join(HERE, "_build"))
LIB = os.environ.get("MHCUDA_LIB_OUT", os.path.join(HERE, "libMHCUDA.so"))
ARCH = os.environ.get("MHCUDA_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-command-line-argument",
    "-ffp-contract=off", "-fno-fast-math",
    "-I" + INCLUDE, "-I" + CSRC,
]
DEVICE = os.environ.get("MHCUDA_EXTRA_DEVICE_FLAGS", "").split() + [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# host-only translation units: plain C++ against the HIP runtime API
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-pthread"]

SOURCES = [
    ("wmh_kernels.hip", True),
    ("lsh.hip", True),
    ("generator.cpp", False),
    ("serve.cpp", False),
    ("autotune.cpp", False),
    ("params_file.cpp", False),
    ("python_module.cpp", False),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in res.stdout.splitlines()
                    if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (res.returncode, " ".join(cmd), out))
    if verbose and out.strip():
        print(out, flush=True)


def build(verbose=False):
    """Compiles every source and links libMHCUDA.so; returns its path."""
    import numpy
    os.makedirs(BUILD, exist_ok=True)
    py_inc = sysconfig.get_paths()["include"]
    objs = []
    for src, is_device in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), "-c"] + COMMON + (DEVICE if is_device else HOST)
        if src == "python_module.cpp":
            cmd += ["-I" + py_inc, "-I" + numpy.get_include()]
        cmd += [os.path.join(CSRC, src), "-o", obj]  # after -x: it applies to the file
        _run(cmd, verbose)
        objs.append(obj)
    libdir = sysconfig.get_config_var("LIBDIR")
    ldlib = sysconfig.get_config_var("LDLIBRARY") or "libpython3.so"
    pylib = ldlib[3:].split(".so")[0]
    tmp = LIB + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "-o", tmp] + objs + ["--offload-arch=" + ARCH] +
         ["-L" + libdir, "-l" + pylib, "-pthread", "-Wl,-rpath," + "/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
