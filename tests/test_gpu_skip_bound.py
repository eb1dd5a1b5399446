"""GPU parity of the skip bound (wmh_bound.h): the persistent kernel's row
walks and its column-stripe pieces (once their row holds a minimum for every
sample of the tile) skip the exact evaluation of (nonzero, sample) pairs
whose lower bound is above the running minimum.  The hashes must be the same
with the skip on and off and equal the oracle's: for sample counts that fill
one or two 128-sample tiles or leave part of a tile empty, for a count that
is not a multiple of 8 (the stripe pieces do not use the bound, the row walks
do), for every lane width, for rows split across work items, for tables
outside the bound's range (not used), and for heavy-tailed, zero, negative,
NaN, subnormal, infinite and unit weights and columns >= dim inside long
rows."""
import ctypes

import numpy as np
import pytest

import oracle
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import workloads

pytestmark = pytest.mark.gpu
PERSISTENT = capi_mod.KERNEL_PERSISTENT


def skip_state(capi, g):
    en, act = ctypes.c_int(), ctypes.c_int()
    assert capi.mhcuda_get_skip(g, ctypes.byref(en), ctypes.byref(act)) == 0
    return en.value, act.value


def long_row_batch(dim, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(3):
        nnz = np.concatenate([rng.integers(1500, 9000, 10), rng.integers(1, 200, 20)])
        rng.shuffle(nnz)
        mats.append(workloads.batched_csr(nnz.astype(np.int64), dim=dim, sigma=sigma,
                                          seed=seed * 10 + i))
    return mats


@pytest.mark.parametrize("samples", [64, 128, 200, 256, 100])
def test_skip_on_and_off_equal_the_oracle(libmhcuda, capi, gpu_available, samples):
    print("samples", samples, flush=True)
    dim = 1 << 15
    mats = long_row_batch(dim, samples)
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=samples, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        assert capi.mhcuda_set_stripes(g, 1) == 0
        assert capi.mhcuda_set_tuning(g, 2, 1024, 0) == 0
        assert skip_state(capi, g) == (1, 1)
        on = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
        assert capi.mhcuda_set_skip(g, 0) == 0
        assert skip_state(capi, g) == (0, 0)
        off = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
        assert capi.mhcuda_set_skip(g, 1) == 0
    finally:
        libmhcuda.minhash_cuda_fini(gen)
    for m, a, b in zip(mats, on, off):
        ref = oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0]
        assert np.array_equal(a, ref)
        assert np.array_equal(b, ref)


def test_tables_outside_the_bound_range(libmhcuda, capi, gpu_available):
    """Assigned tables with a beta of 1, an infinite ln_c, a huge ln_c or an
    ln_c - r below the fp16 range (a bound entry of -inf: the codes have no
    finite lower end) turn the bound off (and the hashes stay exact); valid
    tables turn it on, also with a range of bound entries so wide that the
    code step is 256 (the codes then prune little but stay lower bounds)."""
    dim, samples = 20000, 128
    mats = long_row_batch(dim, 3)
    rs, ln_cs, betas = oracle.datasketch_style_params(dim, samples, seed=4)
    gen = libmhcuda.minhash_cuda_init(dim, samples, deferred=True, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        assert skip_state(capi, g) == (1, 0)  # nothing assigned yet
        assert capi.mhcuda_set_stripes(g, 1) == 0
        assert capi.mhcuda_set_tuning(g, 2, 0, 0) == 0
        cases = []
        b2 = betas.copy()
        b2[3, 11] = 1.0
        cases.append((rs, ln_cs, b2))
        c2 = ln_cs.copy()
        c2[7, 100] = -np.inf
        cases.append((rs, c2, betas))
        c3 = ln_cs.copy()
        c3[0, 5] = np.float32(2.0 ** 61)
        cases.append((rs, c3, betas))
        c4 = ln_cs.copy()
        c4[5, 9] = np.float32(-1e5)
        cases.append((rs, c4, betas))
        for params in cases:
            libmhcuda.minhash_cuda_assign_vars(gen, *params)
            assert skip_state(capi, g) == (1, 0)
            got = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
            for m, h in zip(mats, got):
                assert np.array_equal(h, oracle.calc_omp(*params, m.data, m.indices,
                                                         m.indptr, 16)[0])
        c5 = ln_cs.copy()
        c5[1, 17], c5[2, 40] = np.float32(-3e4), np.float32(3e4)
        for lc in (ln_cs, c5):
            libmhcuda.minhash_cuda_assign_vars(gen, rs, lc, betas)
            assert skip_state(capi, g) == (1, 1)
            got = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
            for m, h in zip(mats, got):
                assert np.array_equal(h, oracle.calc_omp(rs, lc, betas, m.data, m.indices,
                                                         m.indptr, 16)[0])
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_extreme_weights_in_striped_rows(libmhcuda, capi, gpu_available):
    """Heavy-tailed weights (lognormal sigma 6), zeros, subnormals, the
    largest floats, unit weights (ln w = 0) and an infinite weight early in a
    long row (its minimum is -inf from then on, so the bound skips everything
    after it) -- with the skip on, exactly the oracle's hashes."""
    dim, samples = 1 << 15, 128
    rng = np.random.default_rng(8)
    nnz = np.array([12000, 3000, 9000, 5000, 40, 2500, 7000], np.int64)
    m = workloads.batched_csr(nnz, dim=dim, sigma=6.0, seed=8)
    special = np.array([0.0, 1e-42, 3e38, 1.0, 1.0, -1.0], np.float32)
    pick = rng.choice(m.nnz, 3000, replace=False)
    m.data[pick] = special[rng.integers(0, len(special), 3000)]
    m.data[m.indptr[0] + 700] = np.inf
    m.data[m.indptr[3]:m.indptr[4]] = 1.0   # every ln w = 0: t = floor(beta) = 0
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=8, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        assert capi.mhcuda_set_stripes(g, 1) == 0
        assert skip_state(capi, g) == (1, 1)
        for seg in (256, 4096):
            assert capi.mhcuda_set_tuning(g, 2, seg, 0) == 0
            got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
            assert np.array_equal(got, oracle.calc(*params, m.data, m.indices, m.indptr)), seg
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def special_weights(m, rng, count):
    """Puts zeros, negatives, NaN, subnormals, the largest floats and unit
    weights at random positions of m, and one infinite weight early in its
    first long row (that row's minimum is -inf from then on)."""
    special = np.array([0.0, -2.0, np.nan, 1e-42, 3e38, 1.0], np.float32)
    pick = rng.choice(m.nnz, count, replace=False)
    m.data[pick] = special[rng.integers(0, len(special), count)]
    lens = np.diff(m.indptr)
    long_row = int(np.flatnonzero(lens >= 400)[0])
    m.data[m.indptr[long_row] + 37] = np.inf
    return m


@pytest.mark.parametrize("samples,spl", [(2, 1), (34, 1), (64, 1), (100, 2), (128, 2),
                                         (130, 2), (128, 4), (200, 4), (202, 4), (256, 4),
                                         (256, 2)])
def test_row_walk_bound_equals_the_oracle(libmhcuda, capi, gpu_available, samples, spl):
    """Unstriped rows of 1-5,000 nonzeros (most of them hashed by the row
    walk's bounded blocks), some columns >= dim, special weights; items of
    64, 512 and 4096 nonzeros (many rows split across items); bound on and
    off."""
    print("samples", samples, "spl", spl, flush=True)
    dim = 1 << 15
    rng = np.random.default_rng(samples * 10 + spl)
    nnz = np.concatenate([rng.integers(1, 64, 40), rng.integers(64, 1500, 20), [3000, 5000]])
    rng.shuffle(nnz)
    m = workloads.batched_csr(nnz.astype(np.int64), dim=dim + 9, sigma=2.0, seed=samples + spl)
    m = special_weights(m, rng, 500)
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=samples + 1, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        ref = oracle.calc(*params, m.data, m.indices, m.indptr)
        assert capi.mhcuda_set_stripes(g, 0) == 0
        assert skip_state(capi, g) == (1, 1)
        for seg in (64, 512, 4096):
            assert capi.mhcuda_set_tuning(g, spl, seg, 0) == 0
            got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
            assert np.array_equal(got, ref), seg
        assert capi.mhcuda_set_skip(g, 0) == 0
        off = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
        assert np.array_equal(off, ref)
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_row_walk_bound_with_stripes_and_short_rows(libmhcuda, capi, gpu_available):
    """Striped long rows beside bounded row walks of the other rows, at 4
    samples per lane (stripe pieces through the row walk) and at 2 (paired
    and bounded stripe pieces), 256 samples."""
    dim, samples = 1 << 15, 256
    rng = np.random.default_rng(5)
    nnz = np.concatenate([rng.integers(1, 900, 60), rng.integers(1100, 6000, 8)])
    rng.shuffle(nnz)
    m = workloads.batched_csr(nnz.astype(np.int64), dim=dim, sigma=1.5, seed=55)
    m = special_weights(m, rng, 300)
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=6, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        ref = oracle.calc(*params, m.data, m.indices, m.indptr)
        assert capi.mhcuda_set_stripes(g, 1) == 0
        for spl in (4, 2):
            assert capi.mhcuda_set_tuning(g, spl, 1024, 0) == 0
            got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
            assert np.array_equal(got, ref), spl
    finally:
        libmhcuda.minhash_cuda_fini(gen)


@pytest.mark.parametrize("item_table", ["1", "0"])
@pytest.mark.parametrize("samples", [128, 256, 64, 200])
def test_heavy_seeds_equal_the_oracle(libmhcuda, capi, gpu_available, monkeypatch, samples,
                                      item_table):
    """Heavy seeds of the striped rows (mhcuda_set_seeds): the same hashes as
    without them and as the oracle, for one and two 128-sample tiles and part
    tiles, with special weights, columns >= dim, a striped row of zeros only
    (no seed has a finite hash), a striped row whose every nonzero has a
    column >= dim, duplicate (column, weight) pairs of the heaviest weight in
    one lane and in neighbouring lanes (equal hashes at two positions: the
    earlier one wins), and the same rows in random column order; the bounded
    pieces reading the interleaved item table and (MHCUDA_ITEM_TABLE=0) the
    column-block parameter table, and the item table rebuilt when new
    parameters are assigned."""
    print("samples", samples, "item table", item_table, flush=True)
    monkeypatch.setenv("MHCUDA_ITEM_TABLE", item_table)
    dim = 1 << 15
    rng = np.random.default_rng(samples + 3)
    nnz = np.concatenate([rng.integers(1100, 9000, 10), rng.integers(1, 300, 20),
                          [20000, 1500, 1300]])
    m = workloads.batched_csr(nnz.astype(np.int64), dim=dim + 9, sigma=2.0, seed=samples + 4)
    m = special_weights(m, rng, 500)
    lens = np.diff(m.indptr)
    zeros_row, far_row = len(nnz) - 2, len(nnz) - 1
    m.data[m.indptr[zeros_row]:m.indptr[zeros_row + 1]] = 0.0
    m.indices[m.indptr[far_row]:m.indptr[far_row + 1]] = dim + 3
    dup_row = int(np.flatnonzero(lens >= 1100)[1])  # (row 0 holds the infinite weight)
    s = int(m.indptr[dup_row])
    for a, b in ((5, 69), (300, 301)):
        m.indices[s + b] = m.indices[s + a]
        m.data[s + a] = m.data[s + b] = np.float32(3.3e38)
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=samples + 5, devices=1)
    g = ctypes.c_void_p(gen)
    try:
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
        ref = oracle.calc(*params, m.data, m.indices, m.indptr)
        assert capi.mhcuda_set_stripes(g, 1) == 0
        assert skip_state(capi, g) == (1, 1)
        sd = ctypes.c_int()
        assert capi.mhcuda_get_seeds(g, ctypes.byref(sd)) == 0 and sd.value == 0
        assert capi.mhcuda_set_seeds(g, 1) == 0
        assert capi.mhcuda_get_seeds(g, ctypes.byref(sd)) == 0 and sd.value == 1
        for seg in (512, 4096):
            assert capi.mhcuda_set_tuning(g, 2, seg, 0) == 0
            got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
            assert np.array_equal(got, ref), seg
        # the same rows with their nonzeros in random order (the stripe pieces
        # then gather from anywhere in the table)
        rng2 = np.random.default_rng(samples + 6)
        mu = m.copy()
        for r in range(mu.shape[0]):
            s0, e0 = mu.indptr[r], mu.indptr[r + 1]
            perm = rng2.permutation(e0 - s0)
            mu.indices[s0:e0] = mu.indices[s0:e0][perm]
            mu.data[s0:e0] = mu.data[s0:e0][perm]
        got_u = libmhcuda.minhash_cuda_calc_batch(gen, [mu], kernel=PERSISTENT)[0]
        assert np.array_equal(got_u, oracle.calc(*params, mu.data, mu.indices, mu.indptr))
        assert capi.mhcuda_set_seeds(g, 0) == 0
        off = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
        assert np.array_equal(off, ref)
        # other parameters (those of another seed): the bound and item tables
        # are rebuilt from them
        other = libmhcuda.minhash_cuda_init(dim, samples, seed=samples + 7, devices=1)
        try:
            params2 = libmhcuda.minhash_cuda_retrieve_vars(other)
        finally:
            libmhcuda.minhash_cuda_fini(other)
        libmhcuda.minhash_cuda_assign_vars(gen, *params2)
        assert skip_state(capi, g) == (1, 1)
        assert capi.mhcuda_set_seeds(g, 1) == 0
        got2 = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
        assert np.array_equal(got2, oracle.calc(*params2, m.data, m.indices, m.indptr))
    finally:
        libmhcuda.minhash_cuda_fini(gen)
