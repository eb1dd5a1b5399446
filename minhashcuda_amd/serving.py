"""Python client of the engine's serving loop (mhcuda_serve_* in minhashcuda.h).

    server = Server(gen, memory_budget=1 << 30, slots=4)
    ticket = server.submit(csr)          # scipy csr_matrix, float32 data
    hashes = server.wait(ticket)         # (rows, samples, 2) uint32
    print(server.stats())
    server.stop()

Requests are hashed by continuous batching: whatever is pending when a slot
frees up goes into one persistent launch; slots overlap on the device.
"""
import ctypes
import threading

import numpy as np

from . import LIB_PATH
from . import capi as capi_mod


class Server:
    def __init__(self, gen, memory_budget=1 << 30, slots=4, max_batch_requests=256, lib=None):
        self.lib = lib or capi_mod.load(LIB_PATH)
        cfg = capi_mod.ServeConfig(memory_budget, slots, max_batch_requests)
        status = ctypes.c_int(0)
        self.handle = self.lib.mhcuda_serve_start(ctypes.c_void_p(gen), ctypes.byref(cfg),
                                                  ctypes.byref(status))
        if not self.handle:
            raise RuntimeError("mhcuda_serve_start failed: %s"
                               % capi_mod.RESULT_NAMES.get(status.value, status.value))
        params = self.lib.mhcuda_get_parameters(ctypes.c_void_p(gen))
        self.samples = params.samples
        self._keep = {}  # ticket -> arrays that must outlive the request
        self._lock = threading.Lock()

    def submit(self, m):
        """Queues a scipy CSR matrix; returns its ticket.  The C ABI's indptr
        is uint32, so a request holds fewer than 2^32 nonzeros."""
        ip = np.asarray(m.indptr, dtype=np.int64)
        if m.nnz > 0xFFFFFFFF or (len(ip) and ip[-1] > 0xFFFFFFFF):
            raise ValueError("a request must hold fewer than 2^32 nonzeros (uint32 indptr)")
        if len(ip) < 2:
            raise ValueError("a request needs at least one row")
        # the positions are read on the host and uploaded as they are: they
        # must stay inside data and indices
        if ip[0] < 0 or np.any(np.diff(ip) < 0) or ip[-1] > min(len(m.data), len(m.indices)):
            raise ValueError("csr_matrix.indptr must be non-decreasing and within data and indices")
        w = np.ascontiguousarray(m.data, dtype=np.float32)
        c = np.ascontiguousarray(m.indices, dtype=np.uint32)
        r = ip.astype(np.uint32)
        out = np.empty((m.shape[0], self.samples, 2), np.uint32)
        t = ctypes.c_uint64(0)
        capi_mod.check(self.lib.mhcuda_serve_submit(
            self.handle, w.ctypes.data, c.ctypes.data, r.ctypes.data, m.shape[0],
            out.ctypes.data, ctypes.byref(t)), "mhcuda_serve_submit")
        with self._lock:
            self._keep[t.value] = (w, c, r, out)
        return t.value

    def submit_arrays(self, weights, cols, indptr, output):
        """Queues a request given as contiguous arrays (float32 weights, uint32
        columns and indptr, uint32 output of shape (rows, samples, 2)); pinned
        arrays are uploaded and downloaded without staging copies."""
        for a in (weights, cols, indptr, output):
            assert a.flags["C_CONTIGUOUS"]
        if len(indptr) < 2 or output.size != (len(indptr) - 1) * self.samples * 2:
            raise ValueError("indptr must hold rows + 1 entries and output rows * samples * 2")
        ip = np.asarray(indptr).view(np.uint32)
        if np.any(ip[1:] < ip[:-1]) or ip[-1] > min(len(weights), len(cols)):
            raise ValueError("indptr must be non-decreasing and within weights and cols")
        t = ctypes.c_uint64(0)
        capi_mod.check(self.lib.mhcuda_serve_submit(
            self.handle, weights.ctypes.data, cols.ctypes.data, indptr.ctypes.data,
            len(indptr) - 1, output.ctypes.data, ctypes.byref(t)), "mhcuda_serve_submit")
        with self._lock:
            self._keep[t.value] = (weights, cols, indptr, output)
        return t.value

    def wait(self, ticket, with_latency=False):
        """Blocks until the request is hashed; returns its hashes (and latency ms)."""
        lat = ctypes.c_double(0)
        rc = self.lib.mhcuda_serve_wait(self.handle, ticket, ctypes.byref(lat))
        with self._lock:
            keep = self._keep.pop(ticket, None)
        capi_mod.check(rc, "request %d" % ticket)
        return (keep[3], lat.value) if with_latency else keep[3]

    def stats(self, reset=False):
        s = capi_mod.ServeStats()
        capi_mod.check(self.lib.mhcuda_serve_stats(self.handle, ctypes.byref(s), int(reset)),
                       "mhcuda_serve_stats")
        return s.as_dict()

    def stop(self):
        if self.handle:
            self.lib.mhcuda_serve_stop(self.handle)
            self.handle = None
            self._keep.clear()
