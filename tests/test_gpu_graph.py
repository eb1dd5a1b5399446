"""hipGraph replay of the device-pointer calls: every replay equals the eager
result, with unrelated eager work and synchronizations between replays, and a
changed call is captured anew."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import torch
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import workloads

pytestmark = pytest.mark.gpu
DIM = 1 << 14


def device_batch(mats):
    bufs, descs = [], (capi_mod.Matrix * len(mats))()
    for i, m in enumerate(mats):
        w = torch.from_numpy(m.data.astype(np.float32)).cuda()
        c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
        r = torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda()
        o = torch.zeros((m.shape[0], 128, 2), dtype=torch.int32, device="cuda")
        bufs.append((w, c, r, o))
        descs[i] = capi_mod.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), m.shape[0],
                                   o.data_ptr())
    return bufs, descs


@pytest.mark.parametrize("kernel", [capi_mod.KERNEL_PERSISTENT, capi_mod.KERNEL_PLAIN])
def test_graph_replays_equal_eager(libmhcuda, capi, gpu_available, kernel):
    layout = workloads.batched_layout(count=3, min_rows=30, max_rows=400, seed=8)
    mats = [workloads.batched_csr(n, dim=DIM, seed=40 + i) for i, n in enumerate(layout)]
    gen = libmhcuda.minhash_cuda_init(DIM, 128, seed=5)
    g = ctypes.c_void_p(gen)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    ref = [oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0] for m in mats]
    bufs, descs = device_batch(mats)
    stream = torch.cuda.current_stream()
    sptr = ctypes.c_void_p(stream.cuda_stream)
    assert capi.mhcuda_calc_batch_device(g, 3, descs, kernel, sptr) == 0  # eager
    torch.cuda.synchronize()
    eager = [b[3].cpu().numpy().view(np.uint32).copy() for b in bufs]
    for e, r in zip(eager, ref):
        assert np.array_equal(e, r)
    assert capi.mhcuda_set_graph(g, 1) == 0
    other = sp.csr_matrix(oracle.c1_data(3, shape=(50, 60)), dtype=np.float32)
    for replay in range(4):
        for b in bufs:
            b[3].zero_()  # eager fill kernel between replays
        assert capi.mhcuda_calc_batch_device(g, 3, descs, kernel, sptr) == 0
        torch.cuda.synchronize()
        for b, e in zip(bufs, eager):
            assert np.array_equal(b[3].cpu().numpy().view(np.uint32), e), replay
        # unrelated eager work on the same generator (host-memory path)
        libmhcuda.minhash_cuda_calc(gen, other)
    n = ctypes.c_uint64()
    assert capi.mhcuda_graph_captures(g, ctypes.byref(n)) == 0 and n.value == 1
    # a different call shape (two matrices) is captured anew and still exact
    for b in bufs:
        b[3].zero_()
    assert capi.mhcuda_calc_batch_device(g, 2, descs, kernel, sptr) == 0
    torch.cuda.synchronize()
    assert capi.mhcuda_graph_captures(g, ctypes.byref(n)) == 0 and n.value == 2
    for b, e in zip(bufs[:2], eager[:2]):
        assert np.array_equal(b[3].cpu().numpy().view(np.uint32), e)
    assert not bufs[2][3].any()
    libmhcuda.minhash_cuda_fini(gen)


def test_autotune_picks_and_caches(libmhcuda, capi, gpu_available):
    """The autotuner measures the batch's row-length histogram, times
    candidates once per batch shape, caches the choice and installs it; hashes
    under the tuned choice still equal the oracle's."""
    layout = workloads.batched_layout(count=4, min_rows=100, max_rows=800, seed=12)
    mats = [workloads.batched_csr(n, dim=DIM, seed=60 + i) for i, n in enumerate(layout)]
    gen = libmhcuda.minhash_cuda_init(DIM, 128, seed=6)
    g = ctypes.c_void_p(gen)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    bufs, descs = device_batch(mats)
    sptr = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    r1 = capi_mod.TuneResult()
    assert capi.mhcuda_autotune(g, 4, descs, sptr, ctypes.byref(r1)) == 0
    assert r1.cached == 0 and r1.histogram_rows == sum(m.shape[0] for m in mats)
    assert r1.histogram_nnz == sum(m.nnz for m in mats)
    assert r1.kernel in (capi_mod.KERNEL_PLAIN, capi_mod.KERNEL_PERSISTENT)
    assert r1.samples_per_lane in (1, 2, 4) and r1.sample_ms > 0
    r2 = capi_mod.TuneResult()
    assert capi.mhcuda_autotune(g, 4, descs, sptr, ctypes.byref(r2)) == 0
    assert r2.cached == 1 and r2.key == r1.key
    assert (r2.kernel, r2.samples_per_lane, r2.segment) == (r1.kernel, r1.samples_per_lane,
                                                            r1.segment)
    spl, seg, wpc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    capi.mhcuda_get_tuning(g, ctypes.byref(spl), ctypes.byref(seg), ctypes.byref(wpc))
    assert (spl.value, seg.value, wpc.value) == (r1.samples_per_lane, r1.segment,
                                                 r1.waves_per_cu)
    for b in bufs:
        b[3].zero_()
    assert capi.mhcuda_calc_batch_device(g, 4, descs, capi_mod.KERNEL_AUTO, sptr) == 0
    torch.cuda.synchronize()
    for m, b in zip(mats, bufs):
        ref = oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0]
        assert np.array_equal(b[3].cpu().numpy().view(np.uint32), ref)
    libmhcuda.minhash_cuda_fini(gen)
