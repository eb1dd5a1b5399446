"""The allocation-free entry point (mhcuda_calc_device_noalloc): called
through the C ABI twice in a row on device-resident CSR input -- the second
call with a smaller input, after the first left its plan in the shared device
scratch -- each result bit-equal to mhcuda_calc_device's on the same input
and to the oracle; sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import torch
from minhashcuda_amd import workloads

pytestmark = pytest.mark.gpu
DIM = 20000


def to_device(m, samples):
    w = torch.from_numpy(m.data.astype(np.float32)).cuda()
    c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
    r = torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda()
    o = torch.full((m.shape[0], samples, 2), -7, dtype=torch.int32, device="cuda")
    return w, c, r, o


def test_noalloc_equals_calc_device(libmhcuda, capi, gpu_available):
    gen = libmhcuda.minhash_cuda_init(DIM, 128, seed=9, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        # 1 MiB column stripes over the 30.7 MB table: rows of 1024 nonzeros
        # or more are hashed stripe by stripe (the striped-row list in the
        # scratch is used), shorter ones through packed work items
        assert capi.mhcuda_set_stripes(g, 1) == 0
        rng = np.random.default_rng(4)
        big = workloads.batched_csr(rng.integers(1, 3000, 300).astype(np.int64), dim=DIM,
                                    sigma=1.5, seed=21)
        small = workloads.batched_csr(rng.integers(1, 1500, 37).astype(np.int64), dim=DIM,
                                      sigma=1.5, seed=22)
        sptr = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for m in (big, small):
            print("noalloc case: rows %d nnz %d" % (m.shape[0], m.nnz), flush=True)
            w, c, r, o1 = to_device(m, 128)
            o2 = o1.clone()
            assert capi.mhcuda_calc_device(g, w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                           m.shape[0], o1.data_ptr(), sptr) == 0
            assert capi.mhcuda_calc_device_noalloc(g, w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                                   m.shape[0], o2.data_ptr(), sptr) == 0
            torch.cuda.synchronize()
            h1 = o1.cpu().numpy().view(np.uint32)
            h2 = o2.cpu().numpy().view(np.uint32)
            assert np.array_equal(h2, h1)
            assert np.array_equal(h2, oracle.calc_omp(*params, m.data, m.indices, m.indptr,
                                                      16)[0])
        # NULL stream: the generator's stream, returns after completion
        w, c, r, o3 = to_device(small, 128)
        assert capi.mhcuda_calc_device_noalloc(g, w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                               small.shape[0], o3.data_ptr(), None) == 0
        assert np.array_equal(o3.cpu().numpy().view(np.uint32), h1)
        # beyond the scratch's 1,048,576 rows: refused before any device work
        assert capi.mhcuda_calc_device_noalloc(g, w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                               (1 << 20) + 1, o3.data_ptr(), sptr) == 1
    finally:
        libmhcuda.minhash_cuda_fini(gen)
