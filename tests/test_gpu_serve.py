"""GPU tests of the continuous-batching serving loop: every request's hashes
equal the oracle's for that request alone, whatever it was batched with and
however the submissions interleave."""
import ctypes
import threading

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import serving, workloads

pytestmark = pytest.mark.gpu
DIM, SAMPLES = 1 << 15, 128


def make_requests(n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        rows = int(np.exp(rng.uniform(0, np.log(300))))
        nnz = np.clip(rng.lognormal(np.log(50), 1.0, rows), 0, 3000).astype(np.int64)
        nnz[rng.random(rows) < 0.05] = 0  # some empty rows
        reqs.append(workloads.batched_csr(nnz, dim=DIM, sigma=1.0, seed=seed * 1000 + i))
    return reqs


@pytest.fixture(scope="module")
def gen_params(libmhcuda, gpu_available):
    gen = libmhcuda.minhash_cuda_init(DIM, SAMPLES, seed=77)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    yield gen, params
    libmhcuda.minhash_cuda_fini(gen)


@pytest.mark.parametrize("slots,budget", [(1, 64 << 20), (3, 96 << 20), (4, 24 << 20)])
def test_concurrent_requests_equal_alone(gen_params, slots, budget):
    gen, params = gen_params
    reqs = make_requests(60, seed=slots)
    server = serving.Server(gen, memory_budget=budget, slots=slots, max_batch_requests=16)
    results = {}

    def client(idx):
        for i in idx:
            t = server.submit(reqs[i])
            results[i] = server.wait(t)

    threads = [threading.Thread(target=client, args=(range(k, 60, 4),)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    stats = server.stats()
    server.stop()
    assert stats["completed"] == 60 and stats["rejected"] == 0
    assert stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    # phase means: a request's queueing is part of its latency, and one
    # launch's phases add up to at most the longest latency
    assert 0 <= stats["queue_ms"] <= stats["mean_ms"]
    phases = stats["stage_in_ms"] + stats["device_ms"] + stats["stage_out_ms"]
    assert min(stats["stage_in_ms"], stats["device_ms"], stats["stage_out_ms"]) >= 0
    assert 0 < phases <= stats["max_ms"]
    for i, m in enumerate(reqs):
        ref = oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0]
        assert np.array_equal(results[i], ref), "request %d" % i


def test_burst_is_coalesced_and_drained_on_stop(gen_params):
    """A burst submitted at once is coalesced into few launches; stop() hashes
    everything still queued before returning."""
    gen, params = gen_params
    reqs = make_requests(40, seed=9)
    server = serving.Server(gen, memory_budget=256 << 20, slots=2, max_batch_requests=64)
    tickets = [server.submit(m) for m in reqs]
    outs = [server.wait(t) for t in tickets[:5]]
    stats = server.stats()
    # the remaining 35 are waited for after stop(): their arrays are kept alive
    keep = [server._keep[t][3] for t in tickets[5:]]
    server.stop()
    assert stats["batches"] < 40
    for i, m in enumerate(reqs):
        h = outs[i] if i < 5 else keep[i - 5]
        assert np.array_equal(h, oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0])


def test_oversize_request_is_rejected(gen_params):
    gen, _ = gen_params
    server = serving.Server(gen, memory_budget=1 << 20, slots=1)
    big = workloads.batched_csr(np.full(100, 3000), dim=DIM, seed=1)  # > 1 MiB in flight
    t = server.submit(big)
    with pytest.raises(RuntimeError, match="MemoryAllocationFailure"):
        server.wait(t)
    small = workloads.batched_csr(np.full(3, 5), dim=DIM, seed=2)
    assert server.wait(server.submit(small)).shape == (3, SAMPLES, 2)
    assert server.stats()["rejected"] == 1
    server.stop()


def pinned_copy(a):
    """A copy of `a` in pinned host memory (torch pin_memory), as numpy."""
    import torch
    t = torch.empty(a.shape, dtype={np.float32: torch.float32, np.uint32: torch.int32}[a.dtype.type],
                    pin_memory=True)
    view = t.numpy().view(a.dtype)
    view[...] = a
    return t, view


def test_pinned_requests_skip_staging(gen_params):
    """Requests whose arrays live in pinned host memory are DMA'd directly;
    mixed with pageable ones in the same launches, all equal the oracle."""
    gen, params = gen_params
    reqs = make_requests(24, seed=31)
    server = serving.Server(gen, memory_budget=64 << 20, slots=2, max_batch_requests=8)
    keep, tickets = [], []
    for i, m in enumerate(reqs):
        if i % 2:
            tickets.append(server.submit(m))
            continue
        arrs = []
        for a in (m.data.astype(np.float32), m.indices.astype(np.uint32),
                  m.indptr.astype(np.uint32), np.zeros((m.shape[0], SAMPLES, 2), np.uint32)):
            t, v = pinned_copy(np.ascontiguousarray(a))
            keep.append(t)
            arrs.append(v)
        tickets.append(server.submit_arrays(*arrs))
    outs = [server.wait(t) for t in tickets]
    server.stop()
    for m, h in zip(reqs, outs):
        assert np.array_equal(h, oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0])
