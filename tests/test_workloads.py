"""The synthetic benchmark inputs on the CPU: distinct sorted columns per row,
no shared column pattern between rows of equal length, and the same layout
from the host (numpy) and device (torch) generators."""
import numpy as np
import pytest

from minhashcuda_amd import workloads


def check_rows(indptr, cols, dim):
    for r in range(len(indptr) - 1):
        c = cols[indptr[r]:indptr[r + 1]].astype(np.int64)
        assert np.all(np.diff(c) > 0) and (c.size == 0 or (c[0] >= 0 and c[-1] < dim))


def test_host_columns():
    m = workloads.batched_csr(np.array([0, 1, 5, 50, 50, 50, 1000, 1000]), dim=1000, seed=3)
    check_rows(m.indptr, m.indices, 1000)
    full = m.indices[m.indptr[6]:m.indptr[7]]
    assert np.array_equal(full, np.arange(1000))  # n == dim: every column once
    a, b = (m.indices[m.indptr[i]:m.indptr[i + 1]] for i in (3, 4))
    assert not np.array_equal(a, b)


def test_equal_length_rows_do_not_align():
    """The i-th columns of equal-length rows spread over the whole range, not
    over one stratum of dim / n columns."""
    dim, n, rows = 1 << 20, 1000, 400
    m = workloads.batched_csr(np.full(rows, n), dim=dim, seed=4)
    ith = m.indices.reshape(rows, n)[:, n // 2].astype(np.int64)
    assert ith.std() > 5 * dim / n  # one stratum would give dim / n / sqrt(12)
    assert abs(ith.mean() - dim / 2) < 0.05 * dim
    check_rows(m.indptr, m.indices, dim)


def test_plain_csr_shape():
    m = workloads.plain_csr(rows=300, dim=5000, mean=20.0, seed=2)
    assert m.shape == (300, 5000)
    check_rows(m.indptr, m.indices, 5000)


def test_torch_generator_on_cpu():
    torch = pytest.importorskip("torch")
    lens = np.array([3, 0, 17, 200, 200, 1, 4096], np.int64)
    w, c, r = workloads.batched_csr_torch(lens, dim=4096, seed=5, device="cpu")
    indptr = r.numpy().view(np.uint32).astype(np.int64)
    cols = c.numpy().view(np.uint32)
    assert np.array_equal(np.diff(indptr), lens)
    check_rows(indptr, cols, 4096)
    assert np.array_equal(cols[indptr[6]:indptr[7]], np.arange(4096))
    assert (w.numpy() > 0).all()


def test_rows_longer_than_dim():
    m = workloads.batched_csr(np.array([5, 300, 40]), dim=100, seed=6)
    c = m.indices[m.indptr[1]:m.indptr[2]].astype(np.int64)
    assert np.all(np.diff(c) >= 0) and c.min() >= 0 and c.max() < 100
    check_rows(m.indptr[[0, 1]], m.indices, 100)
