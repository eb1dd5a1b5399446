"""GPU parity at the full shapes of the uneven batch, the large single-GPU
share and the serving workload.  Where the parameter tables are too large for
a host copy the oracle recomputes only the table columns the checked rows
touch (oracle.calc_touched).

* uneven batch: 64 matrices, dim 2^20, rows per matrix log-uniform in
  [1k, 200k], nnz/row Zipf(1.2) clipped to [1, 20000], lognormal(0, 1.5)
  weights, 128 samples (2.53M rows, 7.8e9 nonzeros), generated in device
  memory: one persistent launch at the tuning the autotuner picks for it
  equals the 64 plain launches in every hash, and sampled rows of three
  matrices (the longest, striped and skip-bounded, the first and last, and
  random ones) equal the oracle.

* 1,000,000 x 12,000,000 share: rows of 16,800 sorted uniform columns,
  uniform (0, 100] weights, 128 samples, the 18.4 GB generated table, 1,024
  column stripes of 18 MiB in one shared queue.  65,600 rows (1.1e9 nonzeros)
  are hashed device-resident by the persistent kernel and through
  mhcuda_calc_batch's pipelined host path in 65,536-row chunks (the
  reference's chunked uploads, minhashcuda.cc:385-455): the two outputs are
  equal in full, and 64 rows spread over both chunks equal the oracle.
* serving: dim 2^22, 256 samples (12.9 GB table), requests of 1-4,096 rows
  with lognormal(median 200, sigma 1) nonzeros clipped to 50,000, plus one
  request of rows long enough to be striped, hashed concurrently by the
  serving loop; every request equals the oracle."""
import ctypes
import os
import threading

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import torch
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import serving, workloads

pytestmark = pytest.mark.gpu


def host_rows(w, c, r, rows, dim):
    """scipy CSR of the selected rows of a matrix given as host arrays."""
    ip = r.astype(np.int64)
    parts_w, parts_c, indptr = [], [], [0]
    for i in rows:
        s, e = ip[i], ip[i + 1]
        parts_w.append(w[s:e])
        parts_c.append(c[s:e])
        indptr.append(indptr[-1] + (e - s))
    return sp.csr_matrix((np.concatenate(parts_w), np.concatenate(parts_c),
                          np.asarray(indptr, np.int64)), shape=(len(rows), dim))


def device_rows(w, c, r, rows, dim):
    """scipy CSR of the selected rows of a matrix held in device tensors."""
    ip = r.cpu().numpy().view(np.uint32).astype(np.int64)
    parts_w, parts_c, indptr = [], [], [0]
    for i in rows:
        s, e = int(ip[i]), int(ip[i + 1])
        parts_w.append(w[s:e].cpu().numpy())
        parts_c.append(c[s:e].cpu().numpy().view(np.uint32))
        indptr.append(indptr[-1] + (e - s))
    return sp.csr_matrix((np.concatenate(parts_w), np.concatenate(parts_c),
                          np.asarray(indptr, np.int64)), shape=(len(rows), dim))


def test_uneven_batch_full_shape(libmhcuda, capi, gpu_available):
    dim, samples = 1 << 20, 128
    layout = workloads.batched_layout(count=64, seed=1)
    print("uneven batch: generating %d rows, %d nonzeros" % (
        sum(len(n) for n in layout), sum(int(n.sum()) for n in layout)), flush=True)
    tensors = [workloads.batched_csr_torch(n, dim, 1.5, seed=100 + i)
               for i, n in enumerate(layout)]
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=1, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        assert capi.mhcuda_set_tuning(g, 2, 4096, 16) == 0
        assert capi.mhcuda_set_stripes(g, 32) == 0
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        outs = {}
        for kernel in (capi_mod.KERNEL_PERSISTENT, capi_mod.KERNEL_PLAIN):
            o = [torch.zeros((len(n), samples, 2), dtype=torch.int32, device="cuda")
                 for n in layout]
            desc = (capi_mod.Matrix * len(layout))()
            for i, ((w, c, r), oi) in enumerate(zip(tensors, o)):
                desc[i] = capi_mod.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                          len(layout[i]), oi.data_ptr())
            print("uneven batch: kernel %d" % kernel, flush=True)
            assert capi.mhcuda_calc_batch_device(g, len(layout), desc, kernel, stream) == 0
            torch.cuda.synchronize()
            outs[kernel] = o
        pers, plain = outs[capi_mod.KERNEL_PERSISTENT], outs[capi_mod.KERNEL_PLAIN]
        differ = [i for i, (a, b) in enumerate(zip(pers, plain)) if not torch.equal(a, b)]
        assert differ == [], "matrices whose hashes differ between the kernels: %s" % differ
        del outs, plain
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        rng = np.random.default_rng(3)
        for m in (0, 31, 63):
            nnz = layout[m]
            longest = np.argsort(nnz)[-5:]
            pick = np.unique(np.concatenate([[0, 1, 2, len(nnz) - 1], longest,
                                             rng.integers(0, len(nnz), 40)]))
            sub = device_rows(*tensors[m], pick, dim)
            print("uneven batch: oracle on matrix %d, %d rows (%d nnz, longest %d)" % (
                m, len(pick), sub.nnz, int(nnz.max())), flush=True)
            ref = oracle.calc_omp(*params, sub.data, sub.indices, sub.indptr, 16)[0]
            got = pers[m][torch.as_tensor(pick, device="cuda")].cpu().numpy().view(np.uint32)
            assert np.array_equal(got, ref), m
    finally:
        libmhcuda.minhash_cuda_fini(gen)
        del tensors
        torch.cuda.empty_cache()


def test_large_share_full_shape(libmhcuda, capi, gpu_available):
    dim, samples, per_row, chunk = 12_000_000, 128, 16_800, 65_536
    rows = chunk + 64
    os.environ["MHCUDA_PIECE_NNZ"] = str(1 << 31)
    os.environ["MHCUDA_PIECE_ROWS"] = str(chunk)
    try:
        gen = libmhcuda.minhash_cuda_init(dim, samples, seed=1, devices=1)
    finally:
        del os.environ["MHCUDA_PIECE_NNZ"], os.environ["MHCUDA_PIECE_ROWS"]
    g = ctypes.c_void_p(gen)
    try:
        k, n = ctypes.c_uint32(), ctypes.c_uint32()
        assert capi.mhcuda_get_stripes(g, None, ctypes.byref(k), ctypes.byref(n)) == 0
        assert (k.value, n.value) == (1024, 8192)  # every row is striped
        assert capi.mhcuda_set_tuning(g, 2, 4096, 12) == 0
        print("large share: generating %d rows x %d nonzeros" % (rows, per_row), flush=True)
        w, c, r = workloads.batched_csr_torch(np.full(rows, per_row), dim, seed=7,
                                              weights="uniform100")
        nnz = int(w.numel())
        out = torch.zeros((rows, samples, 2), dtype=torch.int32, device="cuda")
        desc = (capi_mod.Matrix * 1)(capi_mod.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(),
                                                     rows, out.data_ptr()))
        # torch's current stream (the legacy default stream unless changed):
        # the call must be ordered after the generation above without a sync
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert capi.mhcuda_calc_batch_device(g, 1, desc, capi_mod.KERNEL_PERSISTENT,
                                             stream) == 0
        torch.cuda.synchronize()
        dev_h = out.cpu().numpy().view(np.uint32)
        del out
        print("large share: device-resident pass done (%d nnz); pipelined pass" % nnz,
              flush=True)
        hw = torch.empty(nnz, dtype=torch.float32, pin_memory=True)
        hc = torch.empty(nnz, dtype=torch.int32, pin_memory=True)
        hw.copy_(w)
        hc.copy_(c)
        hr = r.cpu().numpy().view(np.uint32).copy()
        del w, c, r
        torch.cuda.synchronize()
        host_out = np.zeros((rows, samples, 2), np.uint32)
        mats = (capi_mod.Matrix * 1)(capi_mod.Matrix(hw.data_ptr(), hc.data_ptr(),
                                                     hr.ctypes.data, rows, host_out.ctypes.data))
        assert capi.mhcuda_calc_batch(g, 1, mats, capi_mod.KERNEL_PERSISTENT) == 0
        diff = np.flatnonzero(np.any(host_out != dev_h, axis=(1, 2)))
        assert len(diff) == 0, "rows differing between the two passes: %s" % diff[:8]
        sample = list(range(16)) + list(range(chunk - 16, chunk)) + list(range(chunk, rows, 2))
        sub = host_rows(hw.numpy(), hc.numpy().view(np.uint32), hr, sample, dim)
        print("large share: oracle on %d rows" % len(sample), flush=True)
        ref = oracle.calc_touched(1, dim, samples, sub)
        assert np.array_equal(dev_h[sample], ref)
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_serving_full_shape(libmhcuda, capi, gpu_available):
    dim, samples = 1 << 22, 256
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=3, devices=1)
    try:
        rng = np.random.default_rng(17)
        sizes = [1, 4096, 37, 600, 2, 150, 1200, 9, 3000, 64]
        reqs = []
        for i, rows in enumerate(sizes):
            nnz = np.clip(np.rint(rng.lognormal(np.log(200), 1.0, rows)), 1, 50000)
            reqs.append(workloads.batched_csr(nnz.astype(np.int64), dim=dim, sigma=1.0,
                                              seed=500 + i))
        # rows long enough to be hashed in column-stripe pieces
        reqs.append(workloads.batched_csr(np.array([50000, 8192, 3, 20000], np.int64), dim=dim,
                                          sigma=1.0, seed=600))
        server = serving.Server(gen, memory_budget=4 << 30, slots=4, max_batch_requests=256)
        results = {}

        def client(idx):
            for i in idx:
                results[i] = server.wait(server.submit(reqs[i]))

        threads = [threading.Thread(target=client, args=(range(k, len(reqs), 3),))
                   for k in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        server.stop()
        for i, m in enumerate(reqs):
            print("serving shape: request %d (%d rows, %d nnz)" % (i, m.shape[0], m.nnz),
                  flush=True)
            assert np.array_equal(results[i], oracle.calc_touched(3, dim, samples, m)), i
    finally:
        libmhcuda.minhash_cuda_fini(gen)
