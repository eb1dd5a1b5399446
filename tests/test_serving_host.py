"""CPU tests of the Python serving client (minhashcuda_amd/serving.py) against
a stand-in for the library's mhcuda_serve_* entry points: the arrays handed
over (dtypes, contiguity, row count, output shape), their lifetime until the
request is waited on, error mapping, and the refusal of requests whose
nonzeros do not fit the ABI's uint32 indptr.  The GPU tests run the same
client against the real library."""
import ctypes
import types

import numpy as np
import pytest
import scipy.sparse as sp

from minhashcuda_amd import serving


class StubLib:
    SAMPLES = 4

    def __init__(self):
        self.next_ticket = 7
        self.requests = {}
        self.fail_wait = False
        self.stopped = False

    def mhcuda_serve_start(self, gen, cfg, status):
        self.cfg = (cfg._obj.memory_budget, cfg._obj.slots, cfg._obj.max_batch_requests)
        return 0xABC

    def mhcuda_get_parameters(self, gen):
        return types.SimpleNamespace(dim=10, samples=self.SAMPLES, verbosity=0)

    def mhcuda_serve_submit(self, handle, w, c, r, rows, out, ticket):
        assert handle == 0xABC
        t = self.next_ticket
        self.next_ticket += 1
        self.requests[t] = (w, c, r, rows, out)
        ticket._obj.value = t
        return 0

    def mhcuda_serve_wait(self, handle, ticket, latency):
        if self.fail_wait:
            return 3
        w, c, r, rows, out = self.requests.pop(ticket)
        hashes = (ctypes.c_uint32 * (rows * self.SAMPLES * 2)).from_address(out)
        for i in range(len(hashes)):
            hashes[i] = i + ticket
        latency._obj.value = 1.5
        return 0

    def mhcuda_serve_stop(self, handle):
        self.stopped = True


def test_submit_hands_over_converted_arrays_and_keeps_them():
    lib = StubLib()
    srv = serving.Server(1, memory_budget=1 << 20, slots=3, max_batch_requests=5, lib=lib)
    assert lib.cfg == (1 << 20, 3, 5)
    m = sp.random(6, 10, density=0.4, format="csr", random_state=1, dtype=np.float64)
    m.indices = m.indices.astype(np.int64)
    t = srv.submit(m)
    w, c, r, out = srv._keep[t]
    assert (w.dtype, c.dtype, r.dtype, out.dtype) == (np.float32, np.uint32, np.uint32, np.uint32)
    assert all(a.flags["C_CONTIGUOUS"] for a in (w, c, r, out))
    assert out.shape == (6, StubLib.SAMPLES, 2)
    addr = lib.requests[t]
    assert addr == (w.ctypes.data, c.ctypes.data, r.ctypes.data, 6, out.ctypes.data)
    assert np.array_equal(w, m.data.astype(np.float32)) and np.array_equal(r, m.indptr)
    got, lat = srv.wait(t, with_latency=True)
    assert lat == 1.5 and got is out and not srv._keep
    assert np.array_equal(got.ravel(), np.arange(got.size, dtype=np.uint32) + t)
    srv.stop()
    assert lib.stopped and srv.handle is None


def test_wait_maps_errors_and_releases_arrays():
    lib = StubLib()
    srv = serving.Server(1, lib=lib)
    t = srv.submit(sp.csr_matrix(np.eye(3, dtype=np.float32)))
    lib.fail_wait = True
    with pytest.raises(RuntimeError, match="mhcudaMemoryAllocationFailure"):
        srv.wait(t)
    assert not srv._keep


def test_requests_beyond_uint32_indptr_are_refused():
    lib = StubLib()
    srv = serving.Server(1, lib=lib)
    huge = types.SimpleNamespace(nnz=1 << 32, shape=(2, 10), data=np.zeros(1, np.float32),
                                 indices=np.zeros(1, np.uint32),
                                 indptr=np.zeros(3, np.uint32))
    with pytest.raises(ValueError, match="2\\^32"):
        srv.submit(huge)
    assert not lib.requests


def test_requests_with_indptr_outside_the_arrays_are_refused():
    lib = StubLib()
    srv = serving.Server(1, lib=lib)
    bad = []
    m = sp.csr_matrix(np.eye(4, dtype=np.float32))
    m.indptr[-1] = 9          # past data and indices
    bad.append(m)
    m = sp.csr_matrix(np.eye(4, dtype=np.float32))
    m.indptr[2] = 0           # decreasing
    bad.append(m)
    for m in bad:
        with pytest.raises(ValueError, match="indptr"):
            srv.submit(m)
    w = np.ones(4, np.float32)
    c = np.zeros(4, np.uint32)
    out = np.empty((2, StubLib.SAMPLES, 2), np.uint32)
    with pytest.raises(ValueError, match="indptr"):
        srv.submit_arrays(w, c, np.array([0, 3, 7], np.uint32), out)
    assert not lib.requests
