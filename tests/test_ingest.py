"""CSR ingest from disk on the CPU: memory-mapped .npz matrices and bounded
row blocks (the GPU tests hash them)."""
import numpy as np
import pytest
import scipy.sparse as sp

from minhashcuda_amd import ingest


def matrix(rows=500, dim=3000, seed=0, empty_every=7):
    rng = np.random.default_rng(seed)
    nnz = rng.integers(0, 40, rows)
    nnz[::empty_every] = 0
    indptr = np.concatenate([[0], np.cumsum(nnz)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(dim, n, replace=False)) for n in nnz]).astype(np.int32)
    data = rng.lognormal(0, 1, indptr[-1]).astype(np.float32)
    return sp.csr_matrix((data, cols, indptr), shape=(rows, dim))


def same(a, b):
    return (a.shape == b.shape and np.array_equal(np.asarray(a.data), np.asarray(b.data))
            and np.array_equal(np.asarray(a.indices), np.asarray(b.indices))
            and np.array_equal(np.asarray(a.indptr), np.asarray(b.indptr)))


def test_load_npz_is_memory_mapped(tmp_path):
    m = matrix()
    path = str(tmp_path / "m.npz")
    sp.save_npz(path, m, compressed=False)
    got = ingest.load_npz_csr(path)
    assert isinstance(got, sp.csr_matrix)
    assert same(got, m)
    for arr in (got.data, got.indices, got.indptr):
        assert isinstance(arr, np.memmap) and not arr.flags.writeable
    assert (got != m).nnz == 0  # scipy operations still work


def test_load_npz_int64_indices(tmp_path):
    m = matrix(seed=1)
    m64 = sp.csr_matrix((m.data, m.indices.astype(np.int64), m.indptr.astype(np.int64)),
                        shape=m.shape)
    m64.indices = m64.indices.astype(np.int64)
    m64.indptr = m64.indptr.astype(np.int64)
    path = str(tmp_path / "m64.npz")
    sp.save_npz(path, m64, compressed=False)
    got = ingest.load_npz_csr(path)
    assert got.indices.dtype == np.int64 and got.indptr.dtype == np.int64
    assert same(got, m)


def test_load_npz_refusals(tmp_path):
    m = matrix(seed=2)
    comp = str(tmp_path / "c.npz")
    sp.save_npz(comp, m, compressed=True)
    with pytest.raises(ValueError, match="compressed"):
        ingest.load_npz_csr(comp)
    csc = str(tmp_path / "csc.npz")
    sp.save_npz(csc, m.tocsc(), compressed=False)
    with pytest.raises(ValueError, match="csc"):
        ingest.load_npz_csr(csc)
    other = str(tmp_path / "o.npz")
    np.savez(other, a=np.arange(3))
    with pytest.raises(ValueError):
        ingest.load_npz_csr(other)


def test_empty_matrix(tmp_path):
    m = sp.csr_matrix((4, 10), dtype=np.float32)
    path = str(tmp_path / "e.npz")
    sp.save_npz(path, m, compressed=False)
    got = ingest.load_npz_csr(path)
    assert got.shape == (4, 10) and got.nnz == 0
    assert [(r0, r1) for r0, r1, _ in ingest.iter_row_blocks(got, 5)] == [(0, 4)]


@pytest.mark.parametrize("max_nnz", [1, 17, 64, 1000, 10 ** 9])
def test_row_blocks_cover_the_matrix(tmp_path, max_nnz):
    m = matrix(seed=3)
    path = str(tmp_path / "m.npz")
    sp.save_npz(path, m, compressed=False)
    src = ingest.load_npz_csr(path)
    blocks = list(ingest.iter_row_blocks(src, max_nnz))
    assert blocks[0][0] == 0 and blocks[-1][1] == m.shape[0]
    for (a0, a1, _), (b0, b1, _) in zip(blocks, blocks[1:]):
        assert a1 == b0 and a1 > a0
    for r0, r1, blk in blocks:
        assert blk.shape == (r1 - r0, m.shape[1])
        assert blk.indptr[0] == 0
        assert blk.nnz <= max_nnz or r1 - r0 == 1  # only a single long row may exceed it
        assert same(blk, m[r0:r1])
        assert np.shares_memory(blk.data, src.data) or blk.nnz == 0
    stacked = sp.vstack([b for _, _, b in blocks]).tocsr()
    assert (stacked != m).nnz == 0
