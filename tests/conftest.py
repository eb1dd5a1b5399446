"""Shared fixtures of the test suite.

Tests marked ``gpu`` need an MI355X; everything else runs on the CPU.  The
engine is used exactly as a user of the reference would use it: ``import
libMHCUDA`` for the Python module, and ctypes on the same libMHCUDA.so for the
C ABI.  torch (when installed) is imported before the engine so that both
share one HIP runtime in the process.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "minhashcuda_amd")
LIB_PATH = os.path.join(PKG, "libMHCUDA.so")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

try:  # one HIP runtime per process: torch's, if torch is present
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    torch = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _ensure_built():
    if not os.path.exists(LIB_PATH):
        from minhashcuda_amd import build
        build.build()
    import oracle
    if not all(os.path.exists(p) for p in (oracle.LIB_PATH, oracle.LOGF_CHECK, oracle.BOUND_CHECK)):
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def libmhcuda():
    """The CPython module `libMHCUDA` (the reference's Python API)."""
    import libMHCUDA
    return libMHCUDA


@pytest.fixture(scope="session")
def capi():
    """libMHCUDA.so through ctypes with the C ABI prototypes of minhashcuda.h."""
    from minhashcuda_amd import capi as _capi
    return _capi.load(LIB_PATH)


@pytest.fixture(scope="session")
def gpu_available():
    """GPU tests are selected with -m gpu; on a machine without a usable GPU
    they fail loudly rather than skip."""
    if torch is None or not torch.cuda.is_available():
        pytest.fail("test marked gpu but no GPU is visible")
    return True
