"""Weighted MinHash on AMD MI355X (gfx950): a drop-in for the hashing path of
src-d/minhashcuda.

The engine is one shared library, ``libMHCUDA.so`` in this directory, holding
the HIP kernels, the host runtime with the C ABI of ``include/minhashcuda.h``
and the CPython module ``libMHCUDA`` (same functions as the reference's
module).  Build it with ``python -m minhashcuda_amd.build``.

    from minhashcuda_amd import libMHCUDA      # or: import libMHCUDA with this
                                               # directory on sys.path
    gen = libMHCUDA.minhash_cuda_init(dim, 128, seed=1)
    hashes = libMHCUDA.minhash_cuda_calc(gen, csr)   # (rows, 128, 2) uint32
    libMHCUDA.minhash_cuda_fini(gen)

When torch is used in the same process, import torch first so that both load
the same HIP runtime.
"""
import importlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libMHCUDA.so")


def __getattr__(name):
    if name == "libMHCUDA":
        if not os.path.exists(LIB_PATH):
            raise ImportError("libMHCUDA.so is not built: run python -m minhashcuda_amd.build")
        return importlib.import_module(".libMHCUDA", __name__)
    raise AttributeError(name)
