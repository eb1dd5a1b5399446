"""CSR ingest on the GPU: memory-mapped .npz matrices, row blocks, and the row
slices minhash_cuda_calc cuts for matrices beyond the engine's uint32 indptr
(2^32 or more nonzeros), each checked against a plain call and the oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from minhashcuda_amd import ingest

pytestmark = pytest.mark.gpu
DIM, SAMPLES = 3000, 64


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gen(libmhcuda, gpu_available):
    g = libmhcuda.minhash_cuda_init(DIM, SAMPLES, seed=21, devices=1)
    yield g
    libmhcuda.minhash_cuda_fini(g)


@pytest.fixture(scope="module")
def tables(libmhcuda, gen):
    return libmhcuda.minhash_cuda_retrieve_vars(gen)


def matrix(rows=700, seed=0, long_row=True):
    rng = np.random.default_rng(seed)
    nnz = rng.integers(0, 60, rows)
    nnz[::9] = 0
    if long_row:
        nnz[rows // 2] = 2500  # longer than the slice limits used below
    indptr = np.concatenate([[0], np.cumsum(nnz)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(DIM, n, replace=False)) for n in nnz]).astype(np.int32)
    data = rng.lognormal(0, 1, indptr[-1]).astype(np.float32)
    return sp.csr_matrix((data, cols, indptr), shape=(rows, DIM))


def test_npz_memmap(libmhcuda, gen, tables, tmp_path):
    say("npz memmap")
    m = matrix()
    path = str(tmp_path / "m.npz")
    sp.save_npz(path, m, compressed=False)
    mm = ingest.load_npz_csr(path)
    h = libmhcuda.minhash_cuda_calc(gen, mm)
    assert np.array_equal(h, libmhcuda.minhash_cuda_calc(gen, m))
    assert np.array_equal(h, oracle.calc_csr(*tables, m))


@pytest.mark.parametrize("limit", ["1", "500", "4000"])
def test_forced_slices_equal_one_call(libmhcuda, gen, monkeypatch, limit):
    say("forced slices", limit)
    m = matrix(seed=1)
    whole = libmhcuda.minhash_cuda_calc(gen, m)
    part = libmhcuda.minhash_cuda_calc(gen, m, row_start=33, row_finish=650)
    monkeypatch.setenv("MHCUDA_PY_SLICE_NNZ", limit)
    assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, m), whole)
    assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, m, row_start=33, row_finish=650), part)
    assert np.array_equal(part, whole[33:650])


def test_row_blocks(libmhcuda, gen, tmp_path):
    say("row blocks")
    m = matrix(seed=2)
    path = str(tmp_path / "b.npz")
    sp.save_npz(path, m, compressed=False)
    mm = ingest.load_npz_csr(path)
    whole = libmhcuda.minhash_cuda_calc(gen, m)
    got = np.concatenate([libmhcuda.minhash_cuda_calc(gen, blk)
                          for _, _, blk in ingest.iter_row_blocks(mm, 777)])
    assert np.array_equal(got, whole)


def test_indptr_outside_arrays_is_refused(libmhcuda, gen):
    m = matrix(seed=3, long_row=False)
    bad = ingest.csr_from_arrays(m.data[:-5], m.indices, m.indptr, m.shape)
    with pytest.raises(ValueError):
        libmhcuda.minhash_cuda_calc(gen, bad)
    # rows that stay inside the arrays can still be hashed
    ok = libmhcuda.minhash_cuda_calc(gen, bad, row_finish=10)
    assert np.array_equal(ok, libmhcuda.minhash_cuda_calc(gen, m, row_finish=10))


def test_beyond_2p32_positions(libmhcuda, gen, tables, tmp_path):
    """Rows whose CSR positions lie beyond 2^32, in sparse-file memmaps (the
    arrays span 2^32 + 400 entries, only the last rows hold data)."""
    tail = [0, 5, 17, 1, 64, 65, 3, 0, 200, 9]
    p0 = 2 ** 32 + 13
    n = p0 + sum(tail) + 1
    dpath, cpath = str(tmp_path / "data.bin"), str(tmp_path / "cols.bin")
    data = np.memmap(dpath, dtype=np.float32, mode="w+", shape=(n,))
    cols = np.memmap(cpath, dtype=np.int32, mode="w+", shape=(n,))
    for p in (dpath, cpath):
        st = os.stat(p)
        if st.st_blocks * 512 >= st.st_size // 2:
            pytest.skip("the file system does not keep sparse files")
    say("beyond 2^32: sparse arrays of", n, "entries")
    rng = np.random.default_rng(4)
    m_tail = sum(tail)
    data[p0:p0 + m_tail] = rng.lognormal(0, 1, m_tail).astype(np.float32)
    cols[p0:p0 + m_tail] = np.concatenate(
        [np.sort(rng.choice(DIM, k, replace=False)) for k in tail]).astype(np.int32)
    data.flush()
    cols.flush()
    indptr = np.concatenate([[0, 2 ** 31, p0], p0 + np.cumsum(tail)]).astype(np.int64)
    big = ingest.csr_from_arrays(data, cols, indptr, (len(indptr) - 1, DIM))
    small = sp.csr_matrix((np.array(data[p0:p0 + m_tail]), np.array(cols[p0:p0 + m_tail]),
                           indptr[2:] - p0), shape=(len(tail), DIM))
    want = oracle.calc_csr(*tables, small)
    assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, big, row_start=2), want)
    assert np.array_equal(libmhcuda.minhash_cuda_calc(gen, big, row_start=4, row_finish=9),
                          want[2:7])
    with pytest.raises(ValueError):  # a batch does not slice
        libmhcuda.minhash_cuda_calc_batch(gen, [big])
    del data, cols
