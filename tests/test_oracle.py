"""CPU tests that pin the oracle (no GPU needed).

Each check is chosen so that a plausible mistake in the oracle -- a dropped
term, a wrong operation order, a transposed parameter table, the wrong
tie-break -- fails at least one of them:

* the logarithm: exhaustively against an 80-bit logl over every binary32
  input, plus special values;
* Philox4x32-10: the Random123 known-answer vectors;
* the formula and the argmin: against an independent numpy float32
  restatement of the reference tests' oracle (datasketch's minhash), against
  closed forms (single-nonzero rows, w = 1), against FP64 brute force, and on
  the statistical property the hash exists for (P[equal] = weighted Jaccard);
* the committed golden vectors.
"""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.stats import gamma, uniform

import oracle
from minhashcuda_amd import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
S = oracle.SENTINEL


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


# ---------------------------------------------------------------- logarithm

def test_logf_exhaustive_correctly_rounded():
    """wmh_logf(x) == RN(ln x) for every non-negative binary32 x."""
    res = subprocess.run([oracle.LOGF_CHECK], stdout=subprocess.PIPE, text=True, check=True)
    r = json.loads(res.stdout)
    assert r["n"] == 0x7F800001
    assert r["misrounded"] == 0, r
    assert r["specials_bad"] == 0
    assert r["hard"] == 0  # the 80-bit reference itself rounds unambiguously
    assert r["max_ulp"] <= 0.5 + 1e-6


def test_oracle_clean_under_sanitizers():
    """The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer on
    edge inputs (oracle/sanitize_check.c): no invalid access, no undefined
    operation, OpenMP walk equal to the single-thread walk."""
    oracle_dir = os.path.dirname(oracle.LOGF_CHECK)
    res = subprocess.run(["make", "-s", "-C", oracle_dir, "sanitize"], capture_output=True,
                         text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "sanitize_check: ok" in res.stdout


def test_logf_specials():
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, -1e-40, 1.0, 2.0, 1e-45],
                 dtype=np.float32)
    y = oracle.logf(x)
    assert y[0] == -np.inf and y[1] == -np.inf
    assert y[2] == np.inf
    assert np.all(np.isnan(y[3:7]))
    assert y[7] == 0.0 and not np.signbit(y[7])
    assert y[8] == np.float32(np.log(2.0))
    assert y[9] == np.float32(np.log(np.float64(np.float32(1e-45))))


def test_logf_matches_float64_rounding_on_random_inputs():
    rng = np.random.default_rng(5)
    bits = rng.integers(1, 0x7F800000, size=200000, dtype=np.uint32)
    x = bits.view(np.float32)
    ref = np.log(x.astype(np.float64)).astype(np.float32)
    got = oracle.logf(x)
    # double rounding through float64 can differ only on near-ties; none here
    assert np.array_equal(got, ref)


# ---------------------------------------------------------------- Philox

@pytest.mark.parametrize("ctr,key,expected", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expected):
    assert list(oracle.philox(ctr, key)) == expected


def test_generated_parameter_distributions():
    """test.py:72-90: Gamma(2,1) fits of rs and exp(ln_cs), uniform betas."""
    rs, ln_cs, betas = oracle.gen_params(1, 1000, 128)
    assert rs.shape == ln_cs.shape == betas.shape == (128, 1000)
    a, loc, scale = gamma.fit(rs.ravel())
    assert 1.97 < a < 2.03 and -0.01 < loc < 0.01 and 0.98 < scale < 1.02
    a, loc, scale = gamma.fit(np.exp(ln_cs.ravel()))
    assert 1.97 < a < 2.03 and -0.01 < loc < 0.01 and 0.98 < scale < 1.02
    bmin, bmax = uniform.fit(betas.ravel())
    assert 0 <= bmin < 0.001 and 0.999 <= bmax <= 1
    assert np.all(rs > 0) and np.all(np.isfinite(ln_cs))
    assert np.all((betas >= 0) & (betas < 1))


def test_generated_parameters_are_a_pure_function_of_seed():
    a = oracle.gen_params(9, 50, 16)
    b = oracle.gen_params(9, 50, 16)
    c = oracle.gen_params(10, 50, 16)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    # element (k, j) does not depend on the table shape (counter = k*dim + j)
    small = oracle.gen_params(9, 50, 8)
    assert np.array_equal(small[0], a[0][:8])


# ---------------------------------------------------------------- formula

def test_tiny_matches_numpy_restatement():
    g = load_golden("tiny.npz")
    m = sp.csr_matrix((g["data"], g["indices"], g["indptr"]), shape=tuple(g["shape"]))
    h = oracle.calc_csr(g["rs"], g["ln_cs"], g["betas"], m)
    assert h.shape == (2, 128, 2)
    assert np.array_equal(h, oracle.numpy_minhash(m.toarray(), g["rs"], g["ln_cs"], g["betas"]))
    assert np.array_equal(h, g["hashes"])


@pytest.mark.parametrize("seed", [0, 1])
def test_c1_matches_numpy_restatement(seed):
    """The 6400x130 recipe (test.py:35-38, README.md:85-90), all rows."""
    data = oracle.c1_data(seed)
    m = sp.csr_matrix(data, dtype=np.float32)
    assert seed != 1 or m.nnz == 164594
    rs, ln_cs, betas = oracle.datasketch_style_params(130, 128, seed=1)
    h = oracle.calc_csr(rs, ln_cs, betas, m)
    assert np.array_equal(h, oracle.numpy_minhash(data, rs, ln_cs, betas))


def test_float_vector_negative_t():
    g = load_golden("float.npz")
    m = sp.csr_matrix((g["data"], g["indices"], g["indptr"]), shape=tuple(g["shape"]))
    h = oracle.calc_csr(g["rs"], g["ln_cs"], g["betas"], m)
    assert np.array_equal(h, g["hashes"])
    assert (h[0, :, 1].view(np.int32) < 0).any()  # weights < 1 give negative t
    assert np.array_equal(h, oracle.numpy_minhash(m.toarray(), g["rs"], g["ln_cs"], g["betas"]))


def test_closed_form_single_nonzero():
    """A one-nonzero row selects its column with t = floor(ln w / r + beta)."""
    rng = np.random.default_rng(1)
    dim, samples = 40, 32
    rs, ln_cs, betas = oracle.datasketch_style_params(dim, samples, seed=4)
    cols = rng.integers(0, dim, 50).astype(np.uint32)
    w = rng.uniform(0.01, 100, 50).astype(np.float32)
    indptr = np.arange(51, dtype=np.uint32)
    h = oracle.calc(rs, ln_cs, betas, w, cols, indptr)
    lw = oracle.logf(w)
    for i in range(50):
        j = cols[i]
        t = np.floor(lw[i] / rs[:, j] + betas[:, j]).astype(np.int32)
        assert np.array_equal(h[i, :, 0], np.full(samples, j))
        assert np.array_equal(h[i, :, 1].view(np.int32), t)


def test_closed_form_unit_weights():
    """w = 1 gives ln w = 0 exactly, so t = floor(beta) = 0 and the argmin is
    over ln_c + beta*r - r -- independent of any logarithm implementation."""
    dim, samples = 30, 64
    rs, ln_cs, betas = oracle.datasketch_style_params(dim, samples, seed=5)
    cols = np.array([2, 5, 11, 17, 29], np.uint32)
    w = np.ones(5, np.float32)
    h = oracle.calc(rs, ln_cs, betas, w, cols, np.array([0, 5], np.uint32))
    assert (h[0, :, 1] == 0).all()
    with np.errstate(all="ignore"):
        ln_a = ln_cs[:, cols] - (np.float32(0) - betas[:, cols]) * rs[:, cols] - rs[:, cols]
    assert np.array_equal(h[0, :, 0], cols[np.argmin(ln_a, axis=1)])


def test_fp64_brute_force_agreement():
    """FP32 rounding rarely flips the argmin relative to FP64: a wrong term or
    sign in the FP32 path would disagree on most samples."""
    data = oracle.c1_data(2, shape=(500, 60))
    m = sp.csr_matrix(data, dtype=np.float32)
    rs, ln_cs, betas = oracle.datasketch_style_params(60, 128, seed=2)
    h32 = oracle.calc_csr(rs, ln_cs, betas, m)
    h64 = oracle.calc_f64(rs, ln_cs, betas, m.data, m.indices, m.indptr)
    flips = np.mean(np.any(h32 != h64, axis=2))
    assert flips < 1e-3


def test_edge_rows_golden():
    g = load_golden("edges.npz")
    h = oracle.calc(g["rs"], g["ln_cs"], g["betas"], g["data"], g["indices"], g["indptr"])
    assert np.array_equal(h, g["hashes"])
    assert (h[0] == S).all()          # empty row
    assert (h[2] == S).all()          # explicit zeros only
    assert (h[3] == S).all()          # negative / NaN only
    assert (h[7, :, 0] == 5).all()    # duplicate column: earliest position wins


def test_ties_keep_earliest_position_and_signed_zero():
    """Two nonzeros with bit-identical parameters tie on every sample: the
    earlier CSR position must win (strict <).  -0 and +0 weights' logs are
    -inf (never selected)."""
    dim, samples = 4, 16
    rs, ln_cs, betas = oracle.datasketch_style_params(dim, samples, seed=6)
    for p in (rs, ln_cs, betas):
        p[:, 2] = p[:, 1]
    w = np.array([3.0, 3.0], np.float32)
    h = oracle.calc(rs, ln_cs, betas, w, np.array([2, 1], np.uint32), np.array([0, 2], np.uint32))
    assert (h[0, :, 0] == 2).all()
    h = oracle.calc(rs, ln_cs, betas, np.array([-0.0, 0.0], np.float32),
                    np.array([0, 1], np.uint32), np.array([0, 2], np.uint32))
    assert (h == S).all()


def test_t_saturation():
    """(int32)t saturates: a huge weight over a tiny r overflows int32."""
    rs = np.full((8, 2), 1e-30, np.float32)
    ln_cs = np.zeros((8, 2), np.float32)
    betas = np.zeros((8, 2), np.float32)
    h = oracle.calc(rs, ln_cs, betas, np.array([1e30], np.float32), np.array([1], np.uint32),
                    np.array([0, 1], np.uint32))
    assert (h[0, :, 1] == 0x7FFFFFFF).all()


def test_row_independence_and_permutation():
    """Each row's hash depends only on that row: slicing and permuting rows
    permute the output (test.py:112-124)."""
    data = oracle.c1_data(0, shape=(300, 50))
    m = sp.csr_matrix(data, dtype=np.float32)
    rs, ln_cs, betas = oracle.datasketch_style_params(50, 64, seed=8)
    h = oracle.calc_csr(rs, ln_cs, betas, m)
    perm = np.random.default_rng(0).permutation(300)
    assert np.array_equal(oracle.calc_csr(rs, ln_cs, betas, m[perm]), h[perm])
    assert np.array_equal(oracle.calc_csr(rs, ln_cs, betas, m[100:200]), h[100:200])


def test_omp_equals_single_thread():
    data = oracle.c1_data(1)
    m = sp.csr_matrix(data, dtype=np.float32)
    rs, ln_cs, betas = oracle.datasketch_style_params(130, 128, seed=1)
    h = oracle.calc_csr(rs, ln_cs, betas, m)
    h2, used = oracle.calc_omp(rs, ln_cs, betas, m.data, m.indices, m.indptr, 4)
    assert used >= 1 and np.array_equal(h, h2)


def test_jaccard_estimator():
    """P[(k,t) equal] = weighted Jaccard (consistent weighted sampling); the
    mean error over many pairs at 128 samples is small (test.py:92-110 checks
    a single pair)."""
    rng = np.random.default_rng(11)
    dim, samples, pairs = 200, 128, 200
    rs, ln_cs, betas = oracle.gen_params(1, dim, samples)
    base = rng.uniform(0, 10, (pairs, dim)) * (rng.uniform(size=(pairs, dim)) < 0.3)
    other = base * rng.uniform(0.5, 1.5, (pairs, dim))
    other *= rng.uniform(size=(pairs, dim)) < 0.8
    rows = np.vstack([base, other]).astype(np.float32)
    m = sp.csr_matrix(rows)
    h = oracle.calc_csr(rs, ln_cs, betas, m)
    est = np.array([oracle.hash_jaccard(h[i], h[pairs + i]) for i in range(pairs)])
    true = np.array([oracle.weighted_jaccard(rows[i], rows[pairs + i]) for i in range(pairs)])
    err = est - true
    assert abs(err.mean()) < 0.01          # unbiased
    assert np.sqrt(np.mean(err ** 2)) < 0.06  # ~ sqrt(J(1-J)/128)


def test_mix64_known_answer():
    """splitmix64's first output for seed 0 is mix64(0x9e3779b97f4a7c15) =
    0xe220a8397b1dcdaf (the published reference sequence)."""
    assert int(oracle.mix64(np.uint64(0x9E3779B97F4A7C15))) == 0xE220A8397B1DCDAF
    assert int(oracle.mix64(np.uint64(0))) == 0


def test_band_keys_and_jaccard_properties():
    rng = np.random.default_rng(3)
    h = rng.integers(0, 1000, (6, 16, 2)).astype(np.uint32)
    h[1] = h[0]
    h[2, :8] = h[0, :8]                    # first half equal
    h[3] = oracle.SENTINEL                 # an empty row
    k = oracle.band_keys(h, 4)
    assert (k[1] == k[0]).all()
    assert (k[2, :2] == k[0, :2]).all() and (k[2, 2:] != k[0, 2:]).all()
    assert (k[3] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    j = oracle.jaccard_pairs(h, h, np.array([[0, 1], [0, 2], [3, 3], [0, 4]]))
    assert j[0] == 1.0 and j[1] == 0.5 and j[2] == 0.0 and j[3] < 0.2


def test_lsh_candidate_pairs():
    from minhashcuda_amd import lsh
    keys = np.array([[1, 7], [2, 7], [1, 9], [3, 9], [lsh.RESERVED, 4], [5, 4], [5, 8]],
                    dtype=np.uint64)
    pairs = lsh.candidate_pairs(keys)
    assert pairs.tolist() == [[0, 1], [0, 2], [2, 3], [4, 5], [5, 6]]
    # oversized buckets are skipped
    big = np.zeros((100, 1), np.uint64)
    assert len(lsh.candidate_pairs(big, max_bucket=64)) == 0
    assert len(lsh.candidate_pairs(big[:10], max_bucket=64)) == 45


def test_touched_columns_oracle_equals_full_table():
    """calc_touched (the generated table's columns a matrix touches, renumbered
    in order) gives the hashes of calc on the whole generated table, and
    gen_params_cols gives exactly those columns of gen_params."""
    dim, samples = 7000, 64
    full = oracle.gen_params(11, dim, samples)
    cols = np.array([6999, 0, 17, 3500, 17], np.uint32)
    part = oracle.gen_params_cols(11, dim, samples, cols)
    for a, b in zip(part, full):
        assert np.array_equal(a, b[:, cols])
    rng = np.random.default_rng(3)
    m = sp.random(80, dim, density=0.02, format="csr", random_state=4, dtype=np.float32)
    m.data[:] = rng.lognormal(0, 1.5, m.nnz).astype(np.float32)
    m.data[::7] = 0.0  # rows with unselectable entries, some with none at all
    assert np.array_equal(oracle.calc_touched(11, dim, samples, m), oracle.calc_csr(*full, m))


# ---------------------------------------------------------------- skip bound

def test_f16_conversions_exhaustive():
    """fp16 -> f32 equals numpy's (exact) conversion for every pattern, and
    rounding an fp16 value down returns it unchanged."""
    h = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    x = oracle.f16_to_f32(h)
    ref = h.view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(x[~nan], ref[~nan]) and np.all(np.isnan(x[nan]))
    fin = ~nan
    back = oracle.f16_round_down(x[fin])
    # -0 rounds down to -0 and +0 to +0; every other value is its own pattern
    assert np.array_equal(back, h[fin])


def test_f16_round_down_is_the_largest_fp16_not_above():
    rng = np.random.default_rng(12)
    bits = rng.integers(0, 1 << 32, 2_000_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x)]
    x = np.concatenate([x, np.float32([0.0, -0.0, 65504, 65505, 65520, -65504, -65505,
                                       -65520, 1e30, -1e30, 2.0**-24, 2.0**-25, -2.0**-25,
                                       2.0**-14, -2.0**-14 * 1.0001, np.inf, -np.inf])])
    h = oracle.f16_round_down(x)
    v = oracle.f16_to_f32(h).astype(np.float64)
    xd = x.astype(np.float64)
    assert np.all(v <= xd)
    # the next fp16 above (toward +inf) is above x
    hv = h.view(np.float16)
    with np.errstate(over="ignore"):
        up = np.nextafter(hv, np.float16(np.inf)).astype(np.float64)
    ok = (up > xd) | (np.isposinf(xd) & np.isinf(v))
    assert np.all(ok | (hv == np.float16(65504)) & (xd > 65504))


def test_bound_table_and_lw_terms():
    """Entries of generated-table parameters sit below lc - r by at least the
    stated margin; the per-nonzero term handles the non-finite cases."""
    rs, ln_cs, _ = oracle.gen_params(3, 4000, 64)
    x = oracle.f16_to_f32(oracle.bound_entry(rs, ln_cs)).astype(np.float64)
    r, lc = rs.ravel().astype(np.float64), ln_cs.ravel().astype(np.float64)
    assert np.all(x <= lc - r - 2.0**-17 * (np.abs(lc) + r))
    assert np.all(x >= lc - r - 2.0**-9 * (np.abs(lc) + r) - 2.0**-14)
    lw = np.float32([0.0, 1.5, -2.25, np.inf, -np.inf, np.nan, 88.0, -103.0])
    b = oracle.bound_lw(lw)
    assert b[0] == 0.0 and b[1] > 1.5 and b[2] > -2.25 and b[3] == np.inf
    assert b[4] == -np.inf and b[5] == -np.inf
    assert np.all(b[[1, 2, 6, 7]] >= lw[[1, 2, 6, 7]] + 2.0**-17 * np.abs(lw[[1, 2, 6, 7]]))


@pytest.mark.parametrize("variant,expect_violations", [(0, False), (1, True), (2, True),
                                                       (3, True), (4, True), (5, False),
                                                       (6, True), (7, True), (8, True)])
def test_skip_bound_brute_force(variant, expect_violations):
    """The skip bound never exceeds the computed ln_a over 2e7 adversarial
    trials (oracle/bound_check.c); each planted mistake -- the r term dropped,
    the sign of lw flipped, a nearest-rounded entry, no binary32 margin --
    is caught.  Variants 5-8: the byte-code test never skips a (nonzero,
    sample) whose ln_a is at or below the running minimum (ties included);
    T without its + 1, cB rounded down and codes rounded up are caught."""
    import subprocess
    res = subprocess.run([oracle.BOUND_CHECK, str(variant), "20000000"], capture_output=True,
                         text=True, timeout=300)
    violations = int(res.stdout.split("violations")[1].split()[0])
    if expect_violations:
        assert violations > 0, res.stdout
    else:
        assert res.returncode == 0 and violations == 0, res.stdout


def test_code_sum_bit15_is_the_skip_test():
    """Bit 15 of c + (0x4800 - T') + (0x4000 - cB) is set exactly when
    c >= T + cB (T = T' - 2048) over every code, every T' and the whole
    clamped cB range, and the packed form (two 16-bit halves in one 32-bit
    add) gives each half the same answer: no carry crosses the halves."""
    c = np.arange(256, dtype=np.int64)[:, None, None]
    tp = np.arange(0, 4352, dtype=np.int64)[None, :, None]
    cb = np.concatenate([np.arange(-2048, 2305, 37), [-2048, -1653, 0, 1420, 2304]])
    cb = cb.astype(np.int64)[None, None, :]
    h = c + (0x4800 - tp) + (0x4000 - cb)
    assert h.min() >= 0x6E01 and h.max() <= 0x90FF
    assert np.array_equal((h >> 15) & 1, (c >= tp - 2048 + cb).astype(np.int64))
    rng = np.random.default_rng(5)
    n = 1_000_000
    c2 = rng.integers(0, 256, (2, n))
    a2 = 0x4800 - rng.integers(0, 4352, (2, n))
    b = 0x4000 - rng.integers(-2048, 2305, n)
    word = (c2[0] | c2[1] << 16) + (a2[0] | a2[1] << 16) + (b | b << 16)
    for half in (0, 1):
        single = c2[half] + a2[half] + b
        assert np.array_equal((word >> (16 * half + 15)) & 1, (single >> 15) & 1)


def test_coarse_skip_implies_every_fine_skip():
    """The bounded stripe pieces first test the smallest code of an 8-sample
    group against the group's largest T' and skip the group when that sum's
    bit 15 is set.  The skip bit only turns on as c grows or T' falls, so
    then every sample's own (code, T') pair is skipped too: checked over
    random groups of 8 codes and thresholds (clamped ranges, ties and the
    extremes included) against the per-sample test."""
    rng = np.random.default_rng(11)
    n = 200_000
    codes = rng.integers(0, 256, (n, 8))
    codes[: n // 4] = rng.integers(250, 256, (n // 4, 8))           # near the top
    tp = rng.integers(0, 4352, (n, 8))
    tp[n // 4: n // 2] = rng.integers(2040, 2060, (n // 4, 8))       # around T' = 2048
    cb = rng.integers(-2048, 2305, n)
    cb[::7] = rng.choice([-2048, 2304, 0], n // 7 + 1)
    skip = lambda c, t: ((c + (0x4800 - t) + (0x4000 - cb[:, None])) >> 15) & 1
    coarse = skip(codes.min(axis=1, keepdims=True), tp.max(axis=1, keepdims=True))[:, 0]
    fine = skip(codes, tp)
    assert coarse.sum() > n // 10                   # the coarse test does skip groups
    assert np.all(fine[coarse == 1] == 1)
    # planted mistake: the largest code in place of the smallest is not safe
    wrong = skip(codes.max(axis=1, keepdims=True), tp.max(axis=1, keepdims=True))[:, 0]
    assert np.any(fine[wrong == 1] == 0)
    # the packed form: steps m and m + 4 share one add, skip bits 15 and 31
    lo, hi = codes[:, 0], codes[:, 1]
    tmax = tp.max(axis=1)
    word = (lo | hi << 16) + (0x4800 - tmax) * 0x10001 + ((0x4000 - cb) * 0x10001)
    assert np.array_equal((word >> 15) & 1, skip(lo[:, None], tmax[:, None])[:, 0])
    assert np.array_equal((word >> 31) & 1, skip(hi[:, None], tmax[:, None])[:, 0])


def test_codes_are_lower_bounds_of_the_entries():
    """Codes of generated-table entries and of every fp16 pattern decode to
    values at or below the entry, within one step of it unless clamped at
    255; the range covers the table with lo a multiple of delta >= 1/16."""
    rs, ln_cs, _ = oracle.gen_params(4, 3000, 128)
    h = oracle.bound_entry(rs, ln_cs)
    x = oracle.f16_to_f32(h).astype(np.float64)
    lo, delta = oracle.code_range(x.min(), x.max())
    assert delta >= 1 / 16 and lo <= x.min() and float(lo / delta).is_integer()
    assert 255 * delta >= 0.9 * (x.max() - x.min())
    c = oracle.bound_code(h, lo, delta).astype(np.float64)
    v = float(lo) + c * float(delta)
    assert np.all(v <= x)
    assert np.all((v + float(delta) > x) | (c == 255))
    # every finite fp16 value above a range's lower end
    allh = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    allx = oracle.f16_to_f32(allh).astype(np.float64)
    for mn, mx in ((-30.0, 3.0), (-70000.0, 70000.0), (-1.0, -1.0), (5.5, 9.25)):
        lo, delta = oracle.code_range(mn, mx)
        ok = np.isfinite(allx) & (allx >= lo)
        cc = oracle.bound_code(allh[ok], lo, delta).astype(np.float64)
        vv = float(lo) + cc * float(delta)
        assert np.all(vv <= allx[ok])
        assert np.all((vv + float(delta) > allx[ok]) | (cc == 255))


def test_code_thresholds_and_lw_terms():
    """T' is never below the exact floor((thr - lo) / delta) + 2049 (clamped
    to [0, 4351]) and at most one above it (the fma's rounding may reach the
    next integer: more evaluation, never less), exact on the code grid;
    0x4000 - cB equals 0x4000 - ceil(lwb / delta) for finite lwb; the
    infinities take the clamps that evaluate them."""
    rng = np.random.default_rng(6)
    for mn, mx in ((-30.0, 3.0), (-500.0, 40.0)):
        lo, delta = oracle.code_range(mn, mx)
        scale, t0 = np.float32(1 / delta), np.float32(2049 - lo / delta)
        thr = np.concatenate([rng.uniform(-600, 600, 400_000),
                              float(lo) + rng.integers(-3000, 3000, 100_000) * float(delta),
                              [np.inf, -np.inf, 3.4e38, -3.4e38]]).astype(np.float32)
        t = oracle.code_threshold(thr, scale, t0).astype(np.float64)
        with np.errstate(invalid="ignore", over="ignore"):
            exact = np.floor((thr.astype(np.float64) - float(lo)) / float(delta)) + 2049
        want = np.clip(exact, 0, 4351)
        assert np.all(t >= want) and np.all(t <= want + 1)
        grid = slice(400_000, 500_000)
        assert np.array_equal(t[grid], want[grid])
        lw = np.concatenate([rng.uniform(-103.3, 88.7, 400_000), [0.0, 88.72, -103.28]])
        lwb = oracle.bound_lw(lw.astype(np.float32))
        b = oracle.code_lw(lwb, scale).astype(np.int64)
        assert np.all(b == 0x4000 - np.ceil(lwb.astype(np.float64) / float(delta)))
        inf = oracle.code_lw(np.float32([np.inf, -np.inf]), scale)
        assert list(inf) == [0x4000 - 2304, 0x4000 + 2048]


def test_fast_cpu_baseline_equals_oracle():
    """The timed CPU baseline (transposed tables, vectorised, OpenMP) gives
    the oracle's hashes bit for bit, edge weights included."""
    dim, samples = 3000, 128
    params = oracle.gen_params(7, dim, samples)
    nnz = np.random.default_rng(2).integers(0, 400, 300).astype(np.int64)
    m = workloads.batched_csr(nnz, dim=dim, sigma=2.0, seed=3)
    m.data[::13] = 0.0
    m.data[5] = np.inf
    m.data[7] = -1.0
    m.data[9] = np.float32(1e-42)
    ref = oracle.calc_csr(*params, m)
    for threads in (1, 4):
        got, used = oracle.calc_fast(oracle.transpose_params(*params), m.data, m.indices,
                                     m.indptr, threads)
        assert used == threads and np.array_equal(got, ref)
