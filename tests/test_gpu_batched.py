"""GPU parity of the batched persistent path (and the plain path on batches).

The uneven batch is scaled down so the oracle can check every hash: Zipf(1.2)
row lengths clipped to [1, 20000], log-uniform row counts, empty rows and an
empty matrix, lognormal(0, 1.5) weights.  Each row's result must equal the
oracle's bit for bit whatever the work-item size, lane width, queue size or
kernel -- which is also the "alone == in the batch" property, since the
oracle hashes every row on its own.
"""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from minhashcuda_amd import capi as capi_mod
from minhashcuda_amd import workloads

pytestmark = pytest.mark.gpu
PLAIN, PERSISTENT = capi_mod.KERNEL_PLAIN, capi_mod.KERNEL_PERSISTENT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ref_hashes(params, m):
    return oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)[0]


@pytest.fixture(scope="module")
def batch():
    layout = workloads.batched_layout(count=6, min_rows=20, max_rows=600, zipf_s=1.2,
                                      clip=20000, seed=3)
    mats = [workloads.batched_csr(n, dim=1 << 14, sigma=1.5, seed=i) for i, n in enumerate(layout)]
    # an all-empty matrix, an empty row at the start and a run of empty rows
    mats.append(sp.csr_matrix((7, 1 << 14), dtype=np.float32))
    m = mats[0].tolil()
    m[0, :] = 0
    m[5:9, :] = 0
    mats[0] = m.tocsr()
    mats[0].eliminate_zeros()
    return mats


@pytest.fixture(scope="module")
def expected(batch, libmhcuda, gpu_available):
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, seed=21)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    libmhcuda.minhash_cuda_fini(gen)
    return params, [ref_hashes(params, m) for m in batch]


@pytest.mark.parametrize("seg,spl,wpc", [(64, 2, 16), (256, 1, 8), (2048, 2, 16),
                                         (4096, 4, 24), (128, 2, 1)])
def test_persistent_matches_oracle(libmhcuda, capi, batch, expected, seg, spl, wpc):
    say("seg", seg, "spl", spl, "waves/CU", wpc)
    params, ref = expected
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, *params)
    assert capi.mhcuda_set_tuning(ctypes.c_void_p(gen), spl, seg, wpc) == 0
    got = libmhcuda.minhash_cuda_calc_batch(gen, batch, kernel=PERSISTENT)
    # a second call reuses every buffer and must give the same answer
    again = libmhcuda.minhash_cuda_calc_batch(gen, batch[::-1], kernel=PERSISTENT)[::-1]
    libmhcuda.minhash_cuda_fini(gen)
    for i, (g, a, r) in enumerate(zip(got, again, ref)):
        assert np.array_equal(g, r), "matrix %d" % i
        assert np.array_equal(a, r), "matrix %d (second call)" % i


def shuffled_rows(m, seed):
    """The same rows with their nonzeros in random order."""
    rng = np.random.default_rng(seed)
    m = m.copy()
    for r in range(m.shape[0]):
        s, e = m.indptr[r], m.indptr[r + 1]
        perm = rng.permutation(e - s)
        m.indices[s:e] = m.indices[s:e][perm]
        m.data[s:e] = m.data[s:e][perm]
    return m


@pytest.mark.parametrize("stripe_mib,seg,spl", [(8, 512, 2), (2, 256, 1), (3, 2048, 4)])
@pytest.mark.parametrize("shuffled", [False, True])
def test_column_stripes(libmhcuda, capi, batch, expected, stripe_mib, seg, spl, shuffled):
    """Rows of many nonzeros cut into column-stripe pieces (sorted columns,
    and the same rows in random order): every hash equals the oracle's."""
    say("stripes", stripe_mib, "MiB seg", seg, "spl", spl, "shuffled", shuffled)
    params, ref = expected
    mats = [shuffled_rows(m, i) for i, m in enumerate(batch)] if shuffled else batch
    if shuffled:
        ref = [ref_hashes(params, m) for m in mats]
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, *params)
    g = ctypes.c_void_p(gen)
    assert capi.mhcuda_set_tuning(g, spl, seg, 0) == 0
    assert capi.mhcuda_set_stripes(g, stripe_mib) == 0
    mib, k, min_nnz = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32()
    assert capi.mhcuda_get_stripes(g, ctypes.byref(mib), ctypes.byref(k), ctypes.byref(min_nnz)) == 0
    assert mib.value == stripe_mib and k.value > 1
    striped = sum(int((np.diff(m.indptr) >= min_nnz.value).sum()) for m in mats)
    assert striped > 0
    got = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
    again = libmhcuda.minhash_cuda_calc_batch(gen, mats[::-1], kernel=PERSISTENT)[::-1]
    libmhcuda.minhash_cuda_fini(gen)
    for i, (a, b, r) in enumerate(zip(got, again, ref)):
        assert np.array_equal(a, r), "matrix %d" % i
        assert np.array_equal(b, r), "matrix %d (second call)" % i


def test_plain_batch_matches_oracle(libmhcuda, batch, expected):
    params, ref = expected
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, *params)
    got = libmhcuda.minhash_cuda_calc_batch(gen, batch, kernel=PLAIN)
    libmhcuda.minhash_cuda_fini(gen)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


@pytest.mark.parametrize("samples", [2, 100, 256])
def test_persistent_sample_counts(libmhcuda, capi, batch, samples):
    say("samples", samples)
    gen = libmhcuda.minhash_cuda_init(1 << 14, samples, seed=samples)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    assert capi.mhcuda_set_tuning(ctypes.c_void_p(gen), 0, 512, 0) == 0
    got = libmhcuda.minhash_cuda_calc_batch(gen, batch[:3], kernel=PERSISTENT)
    libmhcuda.minhash_cuda_fini(gen)
    for g, m in zip(got, batch[:3]):
        assert np.array_equal(g, ref_hashes(params, m))


def test_single_matrix_default_path(libmhcuda, capi, batch, expected):
    """minhash_cuda_calc (the reference entry point) on a matrix with
    20000-nonzero rows, split across many work items."""
    params, ref = expected
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, *params)
    assert capi.mhcuda_set_tuning(ctypes.c_void_p(gen), 0, 1024, 0) == 0
    big = int(np.argmax([m.nnz for m in batch]))
    h = libmhcuda.minhash_cuda_calc(gen, batch[big])
    h2 = libmhcuda.minhash_cuda_calc(gen, batch[big], row_start=3, row_finish=batch[big].shape[0] - 2)
    libmhcuda.minhash_cuda_fini(gen)
    assert np.array_equal(h, ref[big])
    assert np.array_equal(h2, ref[big][3:-2])


def test_host_rejects_malformed_indptr(libmhcuda, capi):
    gen = libmhcuda.minhash_cuda_init(16, 8)
    w = np.ones(4, np.float32)
    c = np.arange(4, dtype=np.uint32)
    rows = np.array([0, 3, 1, 4], np.uint32)  # decreasing
    out = np.zeros((3, 8, 2), np.uint32)
    rc = capi.mhcuda_calc(ctypes.c_void_p(gen), w.ctypes.data, c.ctypes.data, rows.ctypes.data,
                          3, out.ctypes.data)
    libmhcuda.minhash_cuda_fini(gen)
    assert rc == 1


@pytest.mark.parametrize("kernel", [PLAIN, PERSISTENT])
def test_pipelined_host_path_many_pieces(libmhcuda, batch, expected, monkeypatch, kernel):
    """Host-memory calls are cut into pieces that alternate between two device
    buffer sets (upload of one overlapping the hashing of the other).  Tiny
    piece limits force dozens of pieces, rows larger than the nonzero budget
    and spans that start mid-matrix; every hash must still equal the oracle's."""
    params, ref = expected
    monkeypatch.setenv("MHCUDA_PIECE_NNZ", "20000")
    monkeypatch.setenv("MHCUDA_PIECE_ROWS", "37")
    gen = libmhcuda.minhash_cuda_init(1 << 14, 128, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, *params)
    got = libmhcuda.minhash_cuda_calc_batch(gen, batch, kernel=kernel)
    h = libmhcuda.minhash_cuda_calc(gen, batch[2], row_start=11)
    libmhcuda.minhash_cuda_fini(gen)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    assert np.array_equal(h, ref[2][11:])


def test_stripes_edge_values(libmhcuda, capi):
    """Striped rows holding the edge cases: zero, negative, NaN, infinite and
    subnormal weights, columns >= dim, a row of exactly the striping threshold,
    an all-zero striped row, rows whose columns all fall in one stripe (the
    first, a middle one, only columns >= dim), unsorted rows whose pieces are
    cut by binary search over unordered columns, and a tie between two columns
    with identical parameters on either side of a stripe boundary (the earlier
    position must win, as in the sequential walk)."""
    dim, samples = 4000, 64
    rng = np.random.default_rng(17)
    rs = rng.gamma(2.0, 1.0, (samples, dim)).astype(np.float32)
    ln_cs = np.log(rng.gamma(2.0, 1.0, (samples, dim))).astype(np.float32)
    betas = rng.uniform(0, 1, (samples, dim)).astype(np.float32)
    width = -(-dim // 3)  # 3 stripes of 1334 columns
    j1, j2 = width - 1, width
    for t in (rs, ln_cs, betas):
        t[:, j2] = t[:, j1]

    def row_sorted(n, hi=dim, lo=0):
        return np.sort(rng.choice(np.arange(lo, hi), n, replace=False)).astype(np.uint32)

    rows_cols, rows_w = [], []
    # A: 3000 nonzeros; the tied pair sits at positions 999 | 1000, the last column of
    # stripe 0 and the first of stripe 1, and the last columns are >= dim
    below = np.sort(rng.choice(j1, 999, replace=False))
    above = np.sort(rng.choice(np.arange(j2 + 1, dim + 10), 1999, replace=False))
    c = np.concatenate([below, [j1, j2], above]).astype(np.uint32)
    assert c[999] == j1 and c[1000] == j2 and c[-1] >= dim
    w = rng.lognormal(0, 1, 3000).astype(np.float32)
    w[999] = w[1000] = np.float32(1e30)
    w[10], w[11], w[12], w[13] = 0.0, -1.0, np.nan, np.float32(1e-42)
    rows_cols.append(c)
    rows_w.append(w)
    # E: striped, an infinite weight in its last piece and a subnormal one
    c = row_sorted(1500)
    w = rng.lognormal(0, 1, 1500).astype(np.float32)
    w[1400], w[3] = np.inf, np.float32(1e-41)
    rows_cols.append(c)
    rows_w.append(w)
    rows_cols.append(np.array([], np.uint32)); rows_w.append(np.array([], np.float32))
    # B: exactly the threshold, unsorted columns
    c = rng.permutation(row_sorted(1024)).astype(np.uint32)
    rows_cols.append(c); rows_w.append(rng.lognormal(0, 1, 1024).astype(np.float32))
    # C: one below the threshold
    rows_cols.append(row_sorted(1023)); rows_w.append(rng.lognormal(0, 1, 1023).astype(np.float32))
    # D: striped, every weight zero
    rows_cols.append(row_sorted(2000)); rows_w.append(np.zeros(2000, np.float32))
    # F, G, H: every column in stripe 0, in stripe 1, or >= dim (one piece holds the row)
    for lo, hi in ((0, width), (width, 2 * width), (dim, dim + 10)):
        n = min(1100, hi - lo) if hi - lo >= 1024 else 0
        c = row_sorted(n, hi, lo) if n else np.repeat(np.arange(lo, hi, dtype=np.uint32), 120)[:1200]
        rows_cols.append(c); rows_w.append(rng.lognormal(0, 1, len(c)).astype(np.float32))
    # I: unsorted, columns alternating between the two ends of the table
    c = row_sorted(2400)
    c = np.stack([c[:1200], c[1200:][::-1]], 1).ravel().astype(np.uint32)
    rows_cols.append(c); rows_w.append(rng.lognormal(0, 1, 2400).astype(np.float32))
    for n in (1, 5, 60):
        rows_cols.append(row_sorted(n)); rows_w.append(rng.lognormal(0, 1, n).astype(np.float32))
    indptr = np.concatenate([[0], np.cumsum([len(c) for c in rows_cols])]).astype(np.uint32)
    m = sp.csr_matrix((np.concatenate(rows_w), np.concatenate(rows_cols), indptr),
                      shape=(len(rows_cols), dim + 10))
    m.indices = np.concatenate(rows_cols).astype(np.uint32)  # keep out-of-range columns
    ref = oracle.calc(rs, ln_cs, betas, m.data, m.indices, m.indptr)
    # the tie resolves to the earlier column wherever the pair wins
    assert (ref[0, :, 0] == j1).any() and not (ref[0, :, 0] == j2).any()
    gen = libmhcuda.minhash_cuda_init(dim, samples, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, rs, ln_cs, betas)
    g = ctypes.c_void_p(gen)
    assert capi.mhcuda_set_stripes(g, 1) == 0
    k, min_nnz = ctypes.c_uint32(), ctypes.c_uint32()
    assert capi.mhcuda_get_stripes(g, None, ctypes.byref(k), ctypes.byref(min_nnz)) == 0
    assert k.value == 3 and min_nnz.value == 1024
    for seg in (64, 1024):
        assert capi.mhcuda_set_tuning(g, 0, seg, 0) == 0
        got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
        assert np.array_equal(got, ref), seg
    libmhcuda.minhash_cuda_fini(gen)


@pytest.mark.parametrize("seed", range(12))
def test_random_batches_and_tunings(libmhcuda, capi, seed):
    """Seeded random batches (row lengths from empty to several thousand,
    sorted or shuffled columns, some columns >= dim, zero weights) under
    random tunings, including column stripes: every hash equals the oracle's."""
    rng = np.random.default_rng(1000 + seed)
    dim = int(rng.choice([700, 3000, 20000]))
    samples = int(rng.choice([2, 64, 100, 128, 256]))
    count = int(rng.integers(1, 6))
    mats = []
    for i in range(count):
        rows = int(rng.integers(0, 300))
        kind = rng.integers(0, 4, rows)
        nnz = np.where(kind == 0, 0, np.where(kind == 1, rng.integers(1, 40, rows),
                       np.where(kind == 2, rng.integers(40, 600, rows),
                                rng.integers(600, 4000, rows))))
        nnz = np.minimum(nnz, dim + 5)
        m = workloads.batched_csr(nnz.astype(np.int64), dim=dim + 5, sigma=1.5,
                                  seed=seed * 10 + i)
        if rng.random() < 0.5:
            m = shuffled_rows(m, seed * 10 + i)
        m.data[rng.random(m.nnz) < 0.03] = 0.0
        mats.append(m)
    spl = int(rng.choice([1, 2, 4]))
    seg = int(rng.choice([64, 256, 1024, 4096]))
    wpc = int(rng.choice([1, 8, 16, 32]))
    stripe = int(rng.choice([0, 1, 2]))
    say("seed", seed, "dim", dim, "samples", samples, "count", count, "spl", spl, "seg", seg,
        "wpc", wpc, "stripe MiB", stripe)
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=seed)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    g = ctypes.c_void_p(gen)
    assert capi.mhcuda_set_tuning(g, spl, seg, wpc) == 0
    assert capi.mhcuda_set_stripes(g, stripe) == 0
    try:
        got = libmhcuda.minhash_cuda_calc_batch(gen, mats, kernel=PERSISTENT)
    finally:
        libmhcuda.minhash_cuda_fini(gen)
    for i, (h, m) in enumerate(zip(got, mats)):
        ref = oracle.calc(*params, m.data, m.indices.astype(np.uint32), m.indptr)
        assert np.array_equal(h, ref), (seed, i)


@pytest.mark.parametrize("case", ["near_range_ends", "below_range", "above_range", "nan_r"])
@pytest.mark.parametrize("kernel", [PLAIN, PERSISTENT])
def test_division_modes(libmhcuda, capi, case, kernel):
    """ln(w) / r is computed with a short correctly rounded sequence when every
    r of the table lies in [2^-60, 2^60] and with the general IEEE sequence
    otherwise; both must give the oracle's hashes.  Tables with every r near
    one end of that range (short sequence), and tables with one r just below
    it, just above it or NaN (general sequence).  Weights include 1 (ln w =
    0), 0, subnormals, the largest floats, NaN, negatives, and +inf, whose
    chunk of nonzeros takes the IEEE sequence in either mode."""
    say("case", case, "kernel", kernel)
    dim, samples = 3000, 64
    rng = np.random.default_rng(41)
    rs = rng.gamma(2.0, 1.0, (samples, dim)).astype(np.float32)
    if case == "near_range_ends":
        e = rng.choice([-60, -59, -58, 57, 58, 59], (samples, dim))
        rs = np.ldexp(rng.uniform(1.0, 2.0, (samples, dim)), e).astype(np.float32)
        assert rs.min() >= 2.0 ** -60 and rs.max() <= 2.0 ** 60
    elif case == "below_range":
        rs[5, 17] = np.float32(2.0 ** -61)
    elif case == "above_range":
        rs[5, 17] = np.float32(2.0 ** 61)
    else:
        rs[5, 17] = np.nan
    ln_cs = np.log(rng.gamma(2.0, 1.0, (samples, dim))).astype(np.float32)
    betas = rng.uniform(0, 1, (samples, dim)).astype(np.float32)
    nnz = rng.integers(1, 3000, 40)
    m = workloads.batched_csr(nnz.astype(np.int64), dim=dim, sigma=3.0, seed=7)
    w = m.data
    special = np.array([1.0, 0.0, 1e-42, 3e38, np.nan, -2.0, 1.0, np.float32(1 - 2 ** -24)],
                       np.float32)
    pick = rng.choice(m.nnz, 400, replace=False)
    w[pick] = special[rng.integers(0, len(special), 400)]
    w[m.indptr[3] + 5] = np.inf   # row 3: a +inf weight (selected in every sample)
    w[m.indptr[30]:m.indptr[31]] = 1.0  # row 30: every ln w = 0
    ref = oracle.calc(rs, ln_cs, betas, m.data, m.indices, m.indptr)
    gen = libmhcuda.minhash_cuda_init(dim, samples, deferred=True)
    libmhcuda.minhash_cuda_assign_vars(gen, rs, ln_cs, betas)
    g = ctypes.c_void_p(gen)
    assert capi.mhcuda_set_stripes(g, 1) == 0
    try:
        got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=kernel)[0]
    finally:
        libmhcuda.minhash_cuda_fini(gen)
    c_inf = m.indices[m.indptr[3] + 5]
    assert (ref[3, np.isfinite(rs[:, c_inf]), 0] == c_inf).all()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("samples", [2, 100, 200, 256])
def test_paired_piece_shapes(libmhcuda, capi, samples):
    """Stripe pieces at 2 samples per lane are hashed two nonzeros per step
    (a half-wave per nonzero, 4 samples per lane, 128-sample tiles): sample
    counts that leave a tile partly empty or use several tiles, sorted and
    unsorted striped rows of odd and even lengths, columns >= dim, and short
    rows between them."""
    say("samples", samples)
    dim = max(20000, (3 << 20) // (12 * ((samples + 3) // 4 * 4)))
    rng = np.random.default_rng(samples)
    nnz = np.concatenate([rng.integers(1100, 5000, 12), rng.integers(1, 300, 30)])
    rng.shuffle(nnz)
    m = workloads.batched_csr(nnz.astype(np.int64), dim=dim + 7, sigma=1.5, seed=samples)
    m = shuffled_rows(m, 1) if samples in (100, 256) else m
    m.data[rng.random(m.nnz) < 0.02] = 0.0
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=samples + 1)
    params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    g = ctypes.c_void_p(gen)
    assert capi.mhcuda_set_stripes(g, 1) == 0
    k = ctypes.c_uint32()
    assert capi.mhcuda_get_stripes(g, None, ctypes.byref(k), None) == 0
    assert k.value > 1
    ref = oracle.calc(*params, m.data, m.indices.astype(np.uint32), m.indptr)
    try:
        for seg in (256, 1024):
            assert capi.mhcuda_set_tuning(g, 2, seg, 0) == 0
            got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
            assert np.array_equal(got, ref), seg
    finally:
        libmhcuda.minhash_cuda_fini(gen)


def test_paired_pieces_wide_table(libmhcuda, capi):
    """A 4.6 GB parameter table (3M columns x 128 samples): the stripe pieces
    of an unsorted row hold columns more than 2^32 bytes of parameters away
    from their stripe, so their chunks gather with 64-bit addresses, while the
    sorted rows gather through 32-bit offsets from the stripe's start."""
    dim, samples = 3_000_000, 128
    rng = np.random.default_rng(5)
    nnz = np.array([5000, 1200, 40, 4097, 3001], np.int64)
    m = workloads.batched_csr(nnz, dim=dim, sigma=1.0, seed=5)
    s, e = m.indptr[0], m.indptr[1]
    assert e - s == 5000
    perm = rng.permutation(e - s)
    m.indices[s:e] = m.indices[s:e][perm]       # row 0 unsorted: far columns in every piece
    m.data[s:e] = m.data[s:e][perm]
    gen = libmhcuda.minhash_cuda_init(dim, samples, seed=9)
    g = ctypes.c_void_p(gen)
    try:
        assert capi.mhcuda_set_stripes(g, 4) == 0
        assert capi.mhcuda_set_tuning(g, 2, 1024, 0) == 0
        got = libmhcuda.minhash_cuda_calc_batch(gen, [m], kernel=PERSISTENT)[0]
        params = libmhcuda.minhash_cuda_retrieve_vars(gen)
    finally:
        libmhcuda.minhash_cuda_fini(gen)
    ref = oracle.calc(*params, m.data, m.indices.astype(np.uint32), m.indptr)
    assert np.array_equal(got, ref)
