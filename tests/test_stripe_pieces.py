"""CPU test of the persistent kernel's stripe-piece search (wmh_stripes.h,
built as host code into tests/native/libstripe_check.so): for every column
order the K pieces of a row partition its CSR positions, and for a sorted row
piece k holds exactly the row's columns in [k * width, (k + 1) * width)
(columns past the last bound fall into the last piece)."""
import ctypes
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
LIB = os.path.join(NATIVE, "libstripe_check.so")


@pytest.fixture(scope="module")
def pieces_fn():
    if not os.path.exists(LIB):
        sys.path.insert(0, NATIVE)
        import build_helpers
        build_helpers.build_stripe_check()
    lib = ctypes.CDLL(LIB)
    fn = lib.wmh_stripe_pieces
    p = ctypes.POINTER(ctypes.c_uint32)
    u = ctypes.c_uint32
    fn.argtypes = [p, u, u, u, u, p]
    fn.restype = None

    def pieces(cols, s, e, K, width):
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        out = np.empty(2 * K, np.uint32)
        fn(cols.ctypes.data_as(p), s, e, K, width, out.ctypes.data_as(p))
        return out.reshape(K, 2).astype(np.int64)
    return pieces


def check_partition(b, s, e):
    assert b[0, 0] == s and b[-1, 1] == e
    assert (b[:, 0] <= b[:, 1]).all()
    assert (b[1:, 0] == b[:-1, 1]).all()


def padded(row, rng, before, after):
    """The row inside a larger column array, with extreme values around it."""
    pad = lambda n: rng.choice(np.array([0, 0xFFFFFFFF], np.uint32), n)
    return np.concatenate([pad(before), row.astype(np.uint32), pad(after)]), before


@pytest.mark.parametrize("K", [1, 2, 3, 7, 64, 384, 1024])
def test_sorted_rows_follow_column_bounds(pieces_fn, K):
    rng = np.random.default_rng(K)
    dim = int(rng.integers(K, 3_000_000))
    width = -(-dim // K)
    for n in (0, 1, 2, 9, 100, 3072, 20000):
        n = min(n, dim)
        row = np.sort(rng.choice(dim, n, replace=False))
        if n > 4:
            row[-2:] = [dim + 3, 0xFFFFFFFE]  # columns >= dim sort last
        cols, s = padded(row, rng, int(rng.integers(0, 50)), int(rng.integers(0, 50)))
        b = pieces_fn(cols, s, s + n, K, width)
        check_partition(b, s, s + n)
        for k in range(K):
            a = s + np.searchsorted(row, k * width, "left")
            z = s + n if k == K - 1 else s + np.searchsorted(row, (k + 1) * width, "left")
            assert (b[k, 0], b[k, 1]) == (a, z), (n, k)


@pytest.mark.parametrize("K", [2, 5, 384, 1024])
def test_unsorted_rows_are_partitioned(pieces_fn, K):
    rng = np.random.default_rng(100 + K)
    dim = 1 << 20
    width = -(-dim // K)
    for n in (1, 2, 3, 17, 500, 8192):
        for row in (rng.integers(0, dim, n),                       # duplicates possible
                    rng.permutation(dim)[:n],
                    np.sort(rng.integers(0, dim, n))[::-1]):        # descending
            cols, s = padded(row, rng, int(rng.integers(0, 9)), int(rng.integers(0, 9)))
            b = pieces_fn(cols, s, s + n, K, width)
            check_partition(b, s, s + n)


def test_sorted_rows_with_repeated_columns(pieces_fn):
    rng = np.random.default_rng(7)
    K, dim = 16, 1000
    width = -(-dim // K)
    row = np.sort(rng.integers(0, dim, 400))
    b = pieces_fn(row, 0, len(row), K, width)
    check_partition(b, 0, len(row))
    for k in range(K):
        cols = row[b[k, 0]:b[k, 1]]
        assert ((cols >= k * width) & (cols < (k + 1) * width)).all()
