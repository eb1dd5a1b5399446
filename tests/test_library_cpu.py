"""CPU tests of the drop-in boundary: libMHCUDA.so loads, exports every symbol
include/minhashcuda.h declares plus the CPython module entry point, and the
host-side argument checks behave like the reference's (minhashcuda.cc:217-225,
587-589; python.cc:147-155,291-294) without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import LIB_PATH, ROOT, torch
from minhashcuda_amd import capi as capi_mod

HEADER = os.path.join(ROOT, "include", "minhashcuda.h")
HAS_GPU = torch is not None and torch.cuda.is_available()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhcuda_\w+)\s*\(", text)))


def test_header_declares_reference_abi():
    names = declared_functions()
    for ref in ("mhcuda_init", "mhcuda_get_parameters", "mhcuda_retrieve_random_vars",
                "mhcuda_assign_random_vars", "mhcuda_calc", "mhcuda_fini"):
        assert ref in names
    assert sorted(capi_mod.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared_functions() + ["PyInit_libMHCUDA"]:
        assert hasattr(lib, name), name


def test_result_codes_and_struct_layout():
    text = open(HEADER).read()
    enum = re.search(r"enum MHCUDAResult \{(.*?)\}", text, re.S).group(1)
    names = [x.strip() for x in enum.split(",") if x.strip()]
    assert names == ["mhcudaSuccess", "mhcudaInvalidArguments", "mhcudaNoSuchDevice",
                     "mhcudaMemoryAllocationFailure", "mhcudaRuntimeError",
                     "mhcudaMemoryCopyError"]
    P = capi_mod.GeneratorParameters
    assert ctypes.sizeof(P) == 12
    assert (P.dim.offset, P.samples.offset, P.verbosity.offset) == (0, 4, 8)


def test_init_argument_checks(capi):
    status = ctypes.c_int(-1)
    assert capi.mhcuda_init(0, 128, 1, 0, 0, 0, ctypes.byref(status)) is None
    assert status.value == 1
    status.value = -1
    assert capi.mhcuda_init(10, 0, 1, 0, 0, 0, ctypes.byref(status)) is None
    assert status.value == 1
    assert capi.mhcuda_init(0, 128, 1, 0, 0, 0, None) is None  # NULL status is fine


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device path")
def test_init_without_device(capi):
    status = ctypes.c_int(-1)
    assert capi.mhcuda_init(10, 128, 1, 0, 0, 0, ctypes.byref(status)) is None
    assert status.value == 2  # mhcudaNoSuchDevice


def test_null_generator_paths(capi):
    p = capi.mhcuda_get_parameters(None)
    assert (p.dim, p.samples, p.verbosity) == (0, 0, 0)
    buf = (ctypes.c_float * 4)()
    assert capi.mhcuda_retrieve_random_vars(None, buf, buf, buf) == 1
    assert capi.mhcuda_assign_random_vars(None, buf, buf, buf) == 1
    u = (ctypes.c_uint32 * 4)()
    assert capi.mhcuda_calc(None, buf, u, u, 1, u) == 1
    assert capi.mhcuda_calc_device(None, buf, u, u, 1, u, None) == 1
    assert capi.mhcuda_calc_batch(None, 1, None, 0) == 1
    assert capi.mhcuda_fini(None) == 0
    assert capi.mhcuda_set_tuning(None, 2, 0, 0) == 1
    assert capi.mhcuda_set_stripes(None, 512) == 1
    assert capi.mhcuda_get_stripes(None, None, None, None) == 1


def test_python_module_surface(libmhcuda):
    for fn in ("minhash_cuda_init", "minhash_cuda_retrieve_vars", "minhash_cuda_assign_vars",
               "minhash_cuda_calc", "minhash_cuda_fini"):
        assert callable(getattr(libmhcuda, fn))


def test_python_argument_errors(libmhcuda):
    import scipy.sparse as sp
    with pytest.raises(ValueError):
        libmhcuda.minhash_cuda_init(0, 128)
    with pytest.raises(TypeError):
        libmhcuda.minhash_cuda_init("x", 128)
    m = sp.csr_matrix(np.eye(3, dtype=np.float32))
    with pytest.raises(ValueError):
        libmhcuda.minhash_cuda_calc(0, m)
    with pytest.raises(ValueError):
        libmhcuda.minhash_cuda_retrieve_vars(0)
    with pytest.raises(ValueError):
        libmhcuda.minhash_cuda_assign_vars(0, [1.0], [1.0], [1.0])
    assert libmhcuda.minhash_cuda_fini(0) is None


def _corrupt_indptr_matrices():
    """CSR matrices whose indptr was changed after construction so that it
    points past data / indices, decreases, or starts negative."""
    import scipy.sparse as sp
    out = []
    m = sp.csr_matrix(np.eye(4, dtype=np.float32))
    m.indptr[-1] = 9          # past len(data) == 4
    out.append(m)
    m = sp.csr_matrix(np.eye(4, dtype=np.float32))
    m.indptr[2] = 0           # decreasing
    out.append(m)
    m = sp.csr_matrix(np.eye(4, dtype=np.float32))
    m.indptr = m.indptr.astype(np.int64)
    m.indptr[0] = -(1 << 32)  # wraps to 0 as uint32
    out.append(m)
    return out


def test_python_calc_entry_points_check_indptr_before_any_device_work(libmhcuda):
    # the checks run before the generator is touched: a pointer that is never
    # dereferenced stands in for one (no GPU on this host)
    for m in _corrupt_indptr_matrices():
        with pytest.raises(ValueError, match="indptr"):
            libmhcuda.minhash_cuda_calc(1, m)
        with pytest.raises(ValueError, match="indptr"):
            libmhcuda.minhash_cuda_calc_batch(1, [m])
