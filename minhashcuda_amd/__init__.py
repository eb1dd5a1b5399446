"""Weighted MinHash on AMD MI355X (gfx950): a drop-in for the hashing path of
src-d/minhashcuda.

The engine is one shared library, ``libMHCUDA.so`` in this directory, holding
the HIP kernels, the host runtime with the C ABI of ``include/minhashcuda.h``
and the CPython module ``libMHCUDA`` (same functions as the reference's
module).  Build it with ``python -m minhashcuda_amd.build``.

    from minhashcuda_amd import libMHCUDA      # or: import libMHCUDA with this
                                               # directory on sys.path, or
                                               # after pip install (setup.py)
    gen = libMHCUDA.minhash_cuda_init(dim, 128, seed=1)
    hashes = libMHCUDA.minhash_cuda_calc(gen, csr)   # (rows, 128, 2) uint32
    libMHCUDA.minhash_cuda_fini(gen)

When torch is used in the same process, import torch first so that both load
the same HIP runtime.
"""
import importlib
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LOCAL = os.path.join(HERE, "libMHCUDA.so")


def _lib_path():
    """The engine library: the one built next to this file, else the
    top-level module libMHCUDA that pip installed (one copy of the library
    serves both the module and the ctypes C ABI, so generators are shared)."""
    if os.path.exists(_LOCAL):
        return _LOCAL
    spec = importlib.util.find_spec("libMHCUDA")
    return spec.origin if spec is not None and spec.origin else _LOCAL


LIB_PATH = _lib_path()


def __getattr__(name):
    if name == "libMHCUDA":
        if os.path.exists(_LOCAL):
            return importlib.import_module(".libMHCUDA", __name__)
        if LIB_PATH != _LOCAL:
            return importlib.import_module("libMHCUDA")
        raise ImportError("libMHCUDA.so is not built: run python -m minhashcuda_amd.build")
    raise AttributeError(name)
