#!/usr/bin/env python3
"""Writes tests/golden/*.npz: small input/expected-output vectors.

Inputs are the fixtures of the reference's tests (test.py:13-14 v1/v2,
test.py:185-189 float vector, test.py:35-38 / README.md:85-90 6400x130 recipe)
plus hand-made edge rows.  Expected hashes come from the C oracle and are
only written after they match the independent numpy restatement of the
reference tests' oracle bit for bit.  Parameters come either from the seeded
datasketch-style recipe or from the engine's own Philox generator (the C
restatement of it), so the fixtures also pin the device parameter generator.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402


def checked(rs, ln_cs, betas, m):
    h = oracle.calc_csr(rs, ln_cs, betas, m)
    ref = oracle.numpy_minhash(m.toarray(), rs, ln_cs, betas)
    assert np.array_equal(h, ref), "C oracle and numpy restatement disagree"
    return h


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(path, {k: v.shape for k, v in arrays.items()})


def csr_arrays(m):
    return dict(data=m.data.astype(np.float32), indices=m.indices.astype(np.uint32),
                indptr=m.indptr.astype(np.uint32), shape=np.array(m.shape, np.int64))


def main():
    # test_calc_tiny (test.py:12-32): v1, v2 with a 23-dim parameter recipe
    rs, ln_cs, betas = oracle.datasketch_style_params(23, 128, seed=1)
    m = sp.csr_matrix(np.array([oracle.V1, oracle.V2], dtype=np.float32))
    save("tiny.npz", rs=rs, ln_cs=ln_cs, betas=betas, hashes=checked(rs, ln_cs, betas, m),
         **csr_arrays(m))

    # test_float (test.py:184-210): weights < 1 (negative t), Philox params, seed 7
    rs, ln_cs, betas = oracle.gen_params(7, 20, 128)
    m = sp.csr_matrix(np.array([oracle.VFLOAT], dtype=np.float32))
    save("float.npz", seed=np.array(7), rs=rs, ln_cs=ln_cs, betas=betas,
         hashes=checked(rs, ln_cs, betas, m), **csr_arrays(m))

    # test_calc_big recipe (test.py:35-38, seed 0): first 256 rows
    data = oracle.c1_data(seed=0)[:256]
    rs, ln_cs, betas = oracle.datasketch_style_params(130, 128, seed=1)
    m = sp.csr_matrix(data, dtype=np.float32)
    save("c1_rows256.npz", rs=rs, ln_cs=ln_cs, betas=betas, hashes=checked(rs, ln_cs, betas, m),
         **csr_arrays(m))

    # Philox parameter generator: seed 1, 1000 x 128 (test_random_vars shape);
    # a strided sample of each table keeps the file small
    rs, ln_cs, betas = oracle.gen_params(1, 1000, 128)
    idx = np.arange(0, 128 * 1000, 97)
    save("philox_params.npz", seed=np.array(1), dim=np.array(1000), samples=np.array(128),
         index=idx, rs=rs.ravel()[idx], ln_cs=ln_cs.ravel()[idx], betas=betas.ravel()[idx])

    # edge rows (semantics fixed where the reference is undefined)
    dim, samples = 16, 64
    rs, ln_cs, betas = oracle.datasketch_style_params(dim, samples, seed=3)
    tiny = np.float32(1e-40)  # subnormal weight
    rows = [
        ([], []),                                    # empty row -> sentinel
        ([3], [5.0]),                                # single nonzero
        ([0, 1, 2], [0.0, 0.0, 0.0]),                # explicit zeros only -> sentinel
        ([4, 5], [-1.0, np.nan]),                    # negative / NaN only -> sentinel
        ([1, 7, 9], [1.0, 1.0, 1.0]),                # w = 1: t = floor(beta) = 0
        ([2, 6], [np.inf, 2.0]),                     # +inf weight: ln_a = -inf wins
        ([0, 15], [tiny, 3.0]),                      # subnormal weight
        ([5, 5, 5], [2.0, 2.0, 2.0]),                # duplicate column, ties -> first
        ([3, 8, 12], [1e30, 1e-30, 7.5]),            # extreme weights
        ([0, 3, 99], [2.0, 4.0, 8.0]),               # column 99 >= dim is ignored
    ]
    indptr = np.zeros(len(rows) + 1, np.uint32)
    cols, data = [], []
    for i, (c, w) in enumerate(rows):
        cols += c
        data += w
        indptr[i + 1] = indptr[i] + len(c)
    cols = np.array(cols, np.uint32)
    data = np.array(data, np.float32)
    h = oracle.calc(rs, ln_cs, betas, data, cols, indptr)
    # checks that follow from the semantics alone
    S = oracle.SENTINEL
    assert (h[0] == S).all() and (h[2] == S).all() and (h[3] == S).all()
    assert (h[1, :, 0] == 3).all()
    assert (h[4, :, 1] == 0).all()
    assert (h[5, :, 0] == 2).all() and (h[5, :, 1] == 0x7FFFFFFF).all()
    assert (h[7, :, 0] == 5).all()
    assert set(np.unique(h[9, :, 0])) <= {0, 3}
    save("edges.npz", rs=rs, ln_cs=ln_cs, betas=betas, data=data, indices=cols, indptr=indptr,
         dim=np.array(dim), hashes=h)


if __name__ == "__main__":
    main()
