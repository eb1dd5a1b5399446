"""GPU tests of the host-side paths around the kernels: calls on different
streams sharing one generator's launch staging, host-memory calls mixing
pinned and pageable arrays (pageable ones are staged through pinned buffers
by the copy pool), armed allocation and copy failures of the pipelined host
path and of serving slots (the error reaches the caller, and the generator
or server stays usable), admission of requests against the whole serving
budget, and the refusal of parameter or tuning changes while a server runs."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from conftest import torch
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import serving, workloads

pytestmark = pytest.mark.gpu
DIM, SAMPLES = 1 << 14, 128
FAIL_ALLOC, FAIL_COPY = 1, 2


def reference(params, m):
    return oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0]


@pytest.fixture(scope="module")
def gen_params(libmhcuda, gpu_available):
    gen = libmhcuda.minhash_cuda_init(DIM, SAMPLES, seed=41, devices=1)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    yield gen, params
    libmhcuda.minhash_cuda_fini(gen)


def device_matrix(m):
    w = torch.from_numpy(m.data.astype(np.float32)).cuda()
    c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
    r = torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda()
    o = torch.zeros((m.shape[0], SAMPLES, 2), dtype=torch.int32, device="cuda")
    return (w, c, r, o), capi_mod.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), m.shape[0],
                                         o.data_ptr())


def test_batches_on_two_streams_share_the_staging_safely(gen_params, capi):
    """Two different persistent batches enqueued back to back on two ordinary
    streams (the second call overwrites the generator's descriptor staging
    while the first may still run): both equal the oracle."""
    gen, params = gen_params
    g = ctypes.c_void_p(gen)
    rng = np.random.default_rng(5)
    a = [workloads.batched_csr(rng.integers(1, 4000, 200).astype(np.int64), dim=DIM, seed=s)
         for s in (1, 2)]
    b = [workloads.batched_csr(rng.integers(1, 50, 900).astype(np.int64), dim=DIM, seed=s)
         for s in (3, 4, 5)]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rnd in range(3):
        bufs_a, da = zip(*[device_matrix(m) for m in a])
        bufs_b, db = zip(*[device_matrix(m) for m in b])
        torch.cuda.synchronize()
        descs_a = (capi_mod.Matrix * 2)(*da)
        descs_b = (capi_mod.Matrix * 3)(*db)
        assert capi.mhcuda_calc_batch_device(g, 2, descs_a, capi_mod.KERNEL_PERSISTENT,
                                             ctypes.c_void_p(s1.cuda_stream)) == 0
        assert capi.mhcuda_calc_batch_device(g, 3, descs_b, capi_mod.KERNEL_PERSISTENT,
                                             ctypes.c_void_p(s2.cuda_stream)) == 0
        torch.cuda.synchronize()
        for m, buf in list(zip(a, bufs_a)) + list(zip(b, bufs_b)):
            assert np.array_equal(buf[3].cpu().numpy().view(np.uint32), reference(params, m)), rnd


def pinned_like(a):
    """A copy of numpy array `a` in pinned host memory (and the tensor that owns it)."""
    t = torch.empty(a.shape, dtype=torch.int32 if a.dtype != np.float32 else torch.float32,
                    pin_memory=True)
    v = t.numpy().view(a.dtype)
    v[...] = a
    return t, v


@pytest.mark.parametrize("pieces", ["auto", "many"])
def test_host_calls_mix_pinned_and_pageable_arrays(libmhcuda, capi, gpu_available, pieces):
    """mhcuda_calc_batch over matrices whose weights, columns, indptr and
    output are each pinned or pageable in every combination (pageable inputs
    go through the staging buffers, pinned ones straight to DMA; pageable
    outputs through the landing buffers), in one piece or in pieces of at
    most 3,000 nonzeros: every hash equals the oracle's."""
    if pieces == "many":
        os.environ["MHCUDA_PIECE_NNZ"] = "3000"
    try:
        gen = libmhcuda.minhash_cuda_init(DIM, SAMPLES, seed=43, devices=1)
    finally:
        os.environ.pop("MHCUDA_PIECE_NNZ", None)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        rng = np.random.default_rng(8)
        keep, mats, outs, refs = [], [], [], []
        for combo in range(16):
            m = workloads.batched_csr(rng.integers(1, 600, 40 + combo).astype(np.int64),
                                      dim=DIM, seed=50 + combo)
            arrays = [m.data.astype(np.float32), m.indices.astype(np.uint32),
                      m.indptr.astype(np.uint32),
                      np.zeros((m.shape[0], SAMPLES, 2), np.uint32)]
            for bit in range(4):
                if combo >> bit & 1:
                    t, arrays[bit] = pinned_like(arrays[bit])
                    keep.append(t)
            keep.append(arrays)
            mats.append(capi_mod.Matrix(arrays[0].ctypes.data, arrays[1].ctypes.data,
                                        arrays[2].ctypes.data, m.shape[0],
                                        arrays[3].ctypes.data))
            outs.append(arrays[3])
            refs.append(reference(params, m))
        descs = (capi_mod.Matrix * len(mats))(*mats)
        for kernel in (capi_mod.KERNEL_PLAIN, capi_mod.KERNEL_PERSISTENT):
            for o in outs:
                o[...] = 0
            assert capi.mhcuda_calc_batch(g, len(mats), descs, kernel) == 0
            for combo, (o, ref) in enumerate(zip(outs, refs)):
                assert np.array_equal(o, ref), (kernel, combo)
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_pipelined_path_failures_propagate_and_recover(libmhcuda, capi, gpu_available):
    """An armed device-buffer failure and an armed copy failure in the middle
    of a many-piece pipelined call reach the caller as MemoryError and
    RuntimeError("hipMemcpy failed"); the next call on the same generator is
    exact."""
    os.environ["MHCUDA_PIECE_NNZ"] = "3000"  # read at init: many pieces
    try:
        gen = libmhcuda.minhash_cuda_init(DIM, SAMPLES, seed=42, devices=1)
    finally:
        del os.environ["MHCUDA_PIECE_NNZ"]
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        m = workloads.batched_csr(np.random.default_rng(6).integers(1, 400, 300).astype(np.int64),
                                  dim=DIM, seed=7)
        ref = reference(params, m)
        for nth in (1, 5):
            assert capi.mhcuda_inject_fault(g, FAIL_ALLOC, nth) == 0
            with pytest.raises(MemoryError):
                libmhcuda.minhash_cuda_calc(gen, m)
            assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, m), ref)
        for nth in (1, 7, 40):
            assert capi.mhcuda_inject_fault(g, FAIL_COPY, nth) == 0
            with pytest.raises(RuntimeError, match="Memcpy failed"):
                libmhcuda.minhash_cuda_calc(gen, m)
            assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, m), ref)
        # batches too
        assert capi.mhcuda_inject_fault(g, FAIL_COPY, 9) == 0
        with pytest.raises(RuntimeError, match="Memcpy failed"):
            libmhcuda.minhash_cuda_calc_batch(gen, [m, m])
        for h in libmhcuda.minhash_cuda_calc_batch(gen, [m, m]):
            assert np.array_equal(h, ref)
        assert capi.mhcuda_inject_fault(g, 0, 0) == 0
        assert capi.mhcuda_inject_fault(g, 3, 1) == 1
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_serving_failures_propagate_and_recover(gen_params, capi):
    gen, params = gen_params
    g = ctypes.c_void_p(gen)
    # the device pool cannot be allocated: the server does not start
    assert capi.mhcuda_inject_fault(g, FAIL_ALLOC, 1) == 0
    with pytest.raises(RuntimeError, match="MemoryAllocationFailure"):
        serving.Server(gen, memory_budget=32 << 20, slots=2)
    server = serving.Server(gen, memory_budget=32 << 20, slots=2)
    try:
        m = workloads.batched_csr(np.full(40, 300), dim=DIM, seed=8)
        ref = reference(params, m)
        # a request is four copies: weights, columns, indptr up, hashes down
        for nth in (1, 3, 4):
            assert capi.mhcuda_inject_fault(g, FAIL_COPY, nth) == 0
            t = server.submit(m)
            with pytest.raises(RuntimeError, match="MemoryCopyError"):
                server.wait(t)
            assert np.array_equal(server.wait(server.submit(m)), ref)
    finally:
        assert capi.mhcuda_inject_fault(g, 0, 0) == 0
        server.stop()


def test_admission_uses_the_whole_budget(gen_params):
    """A request larger than budget / slots but within the budget is hashed
    (admission is against the whole pool, not a per-slot share), also while
    small requests occupy other slots."""
    gen, params = gen_params
    budget = 4 << 20
    server = serving.Server(gen, memory_budget=budget, slots=4, max_batch_requests=8)
    try:
        big = workloads.batched_csr(np.full(200, 1500), dim=DIM, seed=9)  # ~2.6 MiB in flight
        small = [workloads.batched_csr(np.full(5, 20), dim=DIM, seed=10 + i) for i in range(6)]
        tickets = [server.submit(x) for x in small[:3]] + [server.submit(big)] + \
                  [server.submit(x) for x in small[3:]]
        outs = [server.wait(t) for t in tickets]
        stats = server.stats()
    finally:
        server.stop()
    assert stats["rejected"] == 0
    for m, h in zip(small[:3] + [big] + small[3:], outs):
        assert np.array_equal(h, reference(params, m))


def test_changes_are_refused_while_serving(gen_params, capi, libmhcuda, tmp_path):
    gen, params = gen_params
    g = ctypes.c_void_p(gen)
    path = str(tmp_path / "p.wmhp")
    assert capi.mhcuda_save_random_vars(g, path.encode(), 0) == 0
    server = serving.Server(gen, memory_budget=8 << 20, slots=1)
    try:
        assert capi.mhcuda_set_tuning(g, 2, 0, 0) == 1
        assert capi.mhcuda_set_stripes(g, 4) == 1
        assert capi.mhcuda_load_random_vars(g, path.encode()) == 1
        rs, ln_cs, betas = params
        assert capi.mhcuda_assign_random_vars(g, rs.ctypes.data, ln_cs.ctypes.data,
                                              betas.ctypes.data) == 1
        m = workloads.batched_csr(np.full(4, 30), dim=DIM, seed=11)
        _, d = device_matrix(m)
        descs = (capi_mod.Matrix * 1)(d)
        r = capi_mod.TuneResult()
        assert capi.mhcuda_autotune(g, 1, descs, None, ctypes.byref(r)) == 1
        # hashing calls still run beside the server
        assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, m), reference(params, m))
    finally:
        server.stop()
    assert capi.mhcuda_set_tuning(g, 0, 0, 0) == 0
    assert capi.mhcuda_set_stripes(g, 4) == 0
