"""TEST INFRASTRUCTURE -- the CPU oracle of the Weighted MinHash hashing path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker (or the timed CPU baseline), never as
part of the product.

Two independent restatements of the reference's hashing path live here:

* ``calc`` / ``calc_omp`` / ``calc_f64`` -- ctypes bindings of the C oracle in
  wmh_oracle.c, which follows kernel.cu:76-111 (CSR walk, strict-< argmin from
  FLT_MAX, output (column, (uint32)(int32)t)).
* ``numpy_minhash`` -- a vectorised numpy float32 restatement of the test
  oracle the reference's tests compare against (datasketch's
  WeightedMinHashGenerator.minhash as called at test.py:22-23,51-52): zeros
  become NaN, ``t = floor(log(v)/r + beta)``, ``ln_a = ln_c - (t-beta)*r - r``,
  ``nanargmin`` per sample.  It shares no code with the C oracle except the
  logarithm it is handed.

Plus the input recipes of the reference's tests (test.py:13-14,35-38,185-189)
and a seeded parameter recipe in the style of datasketch's generator.
"""
import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwmh_oracle.so")
LOGF_CHECK = os.path.join(HERE, "logf_check")
BOUND_CHECK = os.path.join(HERE, "bound_check")
FAST_LIB_PATH = os.path.join(HERE, "libwmh_cpu_fast.so")
SENTINEL = 0xFFFFFFFF

_lock = threading.Lock()
_lib = None
_fast = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compiles the oracle library and the logf checker (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    """The loaded oracle library (built on first use if missing)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                build()
            L = ctypes.CDLL(LIB_PATH)
            L.wmh_oracle_logf.restype = ctypes.c_float
            L.wmh_oracle_logf.argtypes = [ctypes.c_float]
            L.wmh_oracle_logf_vec.argtypes = [_f32p, _f32p, ctypes.c_uint64]
            calc_args = [_f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                         _f32p, _u32p, _u64p, ctypes.c_uint64, _u32p]
            L.wmh_oracle_calc.argtypes = calc_args
            L.wmh_oracle_calc_f64.argtypes = calc_args
            L.wmh_oracle_calc_omp.argtypes = calc_args + [ctypes.c_int]
            L.wmh_oracle_calc_omp.restype = ctypes.c_int
            L.wmh_oracle_philox.argtypes = [_u32p, _u32p, _u32p]
            L.wmh_oracle_f16_round_down.argtypes = [_f32p, _u16p, ctypes.c_uint64]
            L.wmh_oracle_f16_to_f32.argtypes = [_u16p, _f32p, ctypes.c_uint64]
            L.wmh_oracle_bound_entry.argtypes = [_f32p, _f32p, _u16p, ctypes.c_uint64]
            L.wmh_oracle_bound_lw.argtypes = [_f32p, _f32p, ctypes.c_uint64]
            _f32ref = ctypes.POINTER(ctypes.c_float)
            L.wmh_oracle_code_range.argtypes = [ctypes.c_float, ctypes.c_float, _f32ref,
                                                _f32ref]
            L.wmh_oracle_bound_code.argtypes = [_u16p, _u8p, ctypes.c_uint64, ctypes.c_float,
                                                ctypes.c_float]
            L.wmh_oracle_code_threshold.argtypes = [_f32p, _u32p, ctypes.c_uint64,
                                                    ctypes.c_float, ctypes.c_float]
            L.wmh_oracle_code_lw.argtypes = [_f32p, _u32p, ctypes.c_uint64, ctypes.c_float]
            L.wmh_oracle_gen_params.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint32, _f32p, _f32p, _f32p]
            L.wmh_oracle_gen_params_cols.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                                     ctypes.c_uint32, _u32p, ctypes.c_uint64,
                                                     _f32p, _f32p, _f32p]
            _lib = L
    return _lib


def logf(x):
    """The engine's binary32 natural logarithm, evaluated on the CPU."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().wmh_oracle_logf_vec(x.ravel(), y.ravel(), x.size)
    return y


def _params(rs, ln_cs, betas):
    rs = np.ascontiguousarray(rs, dtype=np.float32)
    ln_cs = np.ascontiguousarray(ln_cs, dtype=np.float32)
    betas = np.ascontiguousarray(betas, dtype=np.float32)
    assert rs.shape == ln_cs.shape == betas.shape and rs.ndim == 2
    return rs, ln_cs, betas


def _csr_arrays(weights, cols, indptr):
    w = np.ascontiguousarray(weights, dtype=np.float32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    r = np.ascontiguousarray(indptr, dtype=np.uint64)
    return w, c, r


def _run(fn, rs, ln_cs, betas, weights, cols, indptr, *extra):
    rs, ln_cs, betas = _params(rs, ln_cs, betas)
    samples, dim = rs.shape
    w, c, r = _csr_arrays(weights, cols, indptr)
    length = len(r) - 1
    out = np.empty((length, samples, 2), dtype=np.uint32)
    if w.size == 0:  # ndpointer rejects empty arrays only by address; keep 1 elem
        w = np.zeros(1, np.float32)
        c = np.zeros(1, np.uint32)
    res = fn(rs, ln_cs, betas, dim, samples, w, c, r, length, out.reshape(-1), *extra)
    return out, res


def calc(rs, ln_cs, betas, weights, cols, indptr):
    """Hashes of every CSR row, single thread.  Parameters are [samples][dim]
    (the reference layout); indptr holds absolute positions into weights/cols."""
    return _run(lib().wmh_oracle_calc, rs, ln_cs, betas, weights, cols, indptr)[0]


def calc_omp(rs, ln_cs, betas, weights, cols, indptr, threads):
    """Same as calc() with rows spread over `threads` OpenMP threads.
    Returns (hashes, threads actually used)."""
    return _run(lib().wmh_oracle_calc_omp, rs, ln_cs, betas, weights, cols, indptr, threads)


def calc_f64(rs, ln_cs, betas, weights, cols, indptr):
    """FP64 twin of calc() (information only: how often FP32 flips the argmin)."""
    return _run(lib().wmh_oracle_calc_f64, rs, ln_cs, betas, weights, cols, indptr)[0]


def transpose_params(rs, ln_cs, betas):
    """The three [samples][dim] tables as contiguous [dim][samples] copies
    (the layout of the timed CPU baseline, calc_fast)."""
    return tuple(np.ascontiguousarray(np.asarray(t, np.float32).T) for t in (rs, ln_cs, betas))


def calc_fast(params_t, weights, cols, indptr, threads):
    """The timed CPU baseline (wmh_cpu_fast.c): the same hashes as calc(),
    from transposed tables (transpose_params), vectorised over samples and
    spread over `threads` OpenMP threads.  Returns (hashes, threads used)."""
    global _fast
    with _lock:
        if _fast is None:
            if not os.path.exists(FAST_LIB_PATH):
                build()
            F = ctypes.CDLL(FAST_LIB_PATH)
            F.wmh_cpu_calc_fast.restype = ctypes.c_int
            F.wmh_cpu_calc_fast.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32,
                                            ctypes.c_uint32, _f32p, _u32p, _u64p,
                                            ctypes.c_uint64, _u32p, ctypes.c_int]
            _fast = F
    rs_t, lc_t, b_t = params_t
    dim, samples = rs_t.shape
    w, c, r = _csr_arrays(weights, cols, indptr)
    length = len(r) - 1
    out = np.empty((length, samples, 2), dtype=np.uint32)
    if w.size == 0:
        w = np.zeros(1, np.float32)
        c = np.zeros(1, np.uint32)
    used = _fast.wmh_cpu_calc_fast(rs_t, lc_t, b_t, dim, samples, w, c, r, length,
                                   out.reshape(-1), threads)
    return out, used


def calc_csr(rs, ln_cs, betas, m):
    """calc() on a scipy.sparse CSR matrix."""
    return calc(rs, ln_cs, betas, m.data, m.indices, m.indptr)


def f16_round_down(x):
    """fp16 bit patterns of float32 values rounded toward -inf (the bound
    table's rounding, wmh_bound.h)."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    h = np.empty(x.size, np.uint16)
    lib().wmh_oracle_f16_round_down(x, h, x.size)
    return h


def f16_to_f32(h):
    h = np.ascontiguousarray(h, dtype=np.uint16).ravel()
    x = np.empty(h.size, np.float32)
    lib().wmh_oracle_f16_to_f32(h, x, h.size)
    return x


def bound_entry(r, lc):
    """Bound-table entries (fp16 bits) of parameters r, lc (wmh_bound.h)."""
    r = np.ascontiguousarray(r, dtype=np.float32).ravel()
    lc = np.ascontiguousarray(lc, dtype=np.float32).ravel()
    h = np.empty(r.size, np.uint16)
    lib().wmh_oracle_bound_entry(r, lc, h, r.size)
    return h


def bound_lw(lw):
    lw = np.ascontiguousarray(lw, dtype=np.float32).ravel()
    out = np.empty_like(lw)
    lib().wmh_oracle_bound_lw(lw, out, lw.size)
    return out


def code_range(mn, mx):
    """(lo, delta) of the byte codes of a table with entries in [mn, mx]
    (wmh_bound.h)."""
    lo, delta = ctypes.c_float(0.0), ctypes.c_float(0.0)
    lib().wmh_oracle_code_range(float(mn), float(mx), ctypes.byref(lo), ctypes.byref(delta))
    return np.float32(lo.value), np.float32(delta.value)


def bound_code(h, lo, delta):
    """Byte codes of bound entries (fp16 bits) for (lo, delta)."""
    h = np.ascontiguousarray(h, dtype=np.uint16).ravel()
    c = np.empty(h.size, np.uint8)
    lib().wmh_oracle_bound_code(h, c, h.size, float(lo), float(delta))
    return c


def code_threshold(thr, scale, t0):
    """T' (wmh_code_threshold) of running minima thr."""
    thr = np.ascontiguousarray(thr, dtype=np.float32).ravel()
    t = np.empty(thr.size, np.uint32)
    lib().wmh_oracle_code_threshold(thr, t, thr.size, float(scale), float(t0))
    return t


def code_lw(lwb, scale):
    """0x4000 - cB (wmh_code_lw) of nonzero terms lwb."""
    lwb = np.ascontiguousarray(lwb, dtype=np.float32).ravel()
    b = np.empty(lwb.size, np.uint32)
    lib().wmh_oracle_code_lw(lwb, b, lwb.size, float(scale))
    return b


def philox(ctr, key):
    """Philox4x32-10 of a 4-word counter under a 2-word key."""
    out = np.zeros(4, np.uint32)
    lib().wmh_oracle_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def gen_params(seed, dim, samples):
    """(rs, ln_cs, betas) exactly as the engine's init generates them, [samples][dim]."""
    rs = np.empty((samples, dim), np.float32)
    ln_cs = np.empty_like(rs)
    betas = np.empty_like(rs)
    lib().wmh_oracle_gen_params(seed, dim, samples, rs, ln_cs, betas)
    return rs, ln_cs, betas


def gen_params_cols(seed, dim, samples, cols):
    """Columns `cols` of gen_params(seed, dim, samples), [samples][len(cols)]
    (only the columns a sample of rows touches of a table too large to copy)."""
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    n = len(cols)
    rs = np.empty((samples, n), np.float32)
    ln_cs = np.empty_like(rs)
    betas = np.empty_like(rs)
    lib().wmh_oracle_gen_params_cols(seed, dim, samples, cols, n, rs, ln_cs, betas)
    return rs, ln_cs, betas


def calc_touched(seed, dim, samples, m, threads=16):
    """Hashes of CSR matrix `m` under the generated table gen_params(seed, dim,
    samples), computing only the table columns m touches: its columns are
    renumbered in increasing order (which keeps every nonzero's parameters
    and the CSR order, so every hash's arithmetic is unchanged) and mapped
    back in the output."""
    idx = np.asarray(m.indices, dtype=np.int64)
    uniq, local = np.unique(idx, return_inverse=True)
    params = gen_params_cols(seed, dim, samples, uniq)
    h, _ = calc_omp(*params, m.data, local.astype(np.uint32), m.indptr, threads)
    sel = h[..., 0] != SENTINEL
    h[..., 0] = np.where(sel, uniq.astype(np.uint32)[np.where(sel, h[..., 0], 0)], h[..., 0])
    return h


# --------------------------------------------------------------------------
# numpy restatement of the reference tests' oracle (datasketch minhash)

def numpy_minhash(dense, rs, ln_cs, betas, log=None):
    """Weighted MinHash of dense float32 rows, vectorised over rows.

    dense: (R, dim) array.  Returns (R, samples, 2) uint32 where [..., 0] is
    the argmin column and [..., 1] is (uint32)(int32)t.  Rows with no
    selectable entry get the sentinel pair.  `log` maps a float32 array to its
    logarithm (default: the engine's logf).
    """
    if log is None:
        log = logf
    v = np.array(dense, dtype=np.float32, ndmin=2)
    R, dim = v.shape
    samples = rs.shape[0]
    v = np.where(v == 0, np.float32(np.nan), v).astype(np.float32)
    vlog = log(v).astype(np.float32)
    out = np.empty((R, samples, 2), dtype=np.uint32)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        for i in range(samples):
            r = rs[i].astype(np.float32)
            b = betas[i].astype(np.float32)
            t = np.floor(vlog / r + b)
            ln_y = (t - b) * r
            ln_a = ln_cs[i].astype(np.float32) - ln_y - r
            # kernel.cu:79,106: only values below FLT_MAX can be selected
            sel = np.where(ln_a < np.finfo(np.float32).max, ln_a, np.float32(np.nan))
            allnan = np.all(np.isnan(sel), axis=1)
            k = np.argmin(np.where(np.isnan(sel), np.float32(np.inf), sel), axis=1)
            tk = t[np.arange(R), k]
            tk_i = np.clip(np.nan_to_num(tk, nan=0.0), -2.0**31, 2.0**31 - 1).astype(np.int64)
            out[:, i, 0] = np.where(allnan, SENTINEL, k).astype(np.uint32)
            out[:, i, 1] = np.where(allnan, SENTINEL, tk_i.astype(np.int32).view(np.uint32))
    return out


# --------------------------------------------------------------------------
# input recipes of the reference's tests

V1 = [1, 0, 0, 0, 3, 4, 5, 0, 0, 0, 0, 6, 7, 8, 0, 0, 0, 0, 0, 0, 9, 10, 4]   # test.py:13
V2 = [2, 0, 0, 0, 4, 3, 8, 0, 0, 0, 0, 4, 7, 10, 0, 0, 0, 0, 0, 0, 9, 0, 0]   # test.py:14
VFLOAT = [                                                                       # test.py:185-189
    0, 1.0497366, 0.8494359, 0.66231006, 0.66231006, 0.8494359,
    0, 0.66231006, 0.33652836, 0, 0, 0.5359344,
    0.8494359, 0.66231006, 1.0497366, 0.33652836, 0.66231006, 0.8494359,
    0.6800841, 0.33652836]


def c1_data(seed=1, shape=(6400, 130)):
    """The 6400x130 recipe of test.py:35-38 / README.md:85-90 (int64 dense)."""
    rng = np.random.RandomState(seed)
    data = rng.randint(0, 100, shape)
    mask = rng.randint(0, 5, data.shape)
    data *= (mask >= 4)
    return data


def datasketch_style_params(dim, samples=128, seed=1):
    """A seeded parameter recipe in the style of datasketch's
    WeightedMinHashGenerator: rs ~ Gamma(2,1), ln_cs = ln Gamma(2,1),
    betas ~ U[0,1), drawn from numpy's RandomState and cast to float32.
    It only has to be a fixed, reproducible recipe: parity never depends on
    it matching datasketch's own draws."""
    rng = np.random.RandomState(seed)
    rs = rng.gamma(2, 1, (samples, dim)).astype(np.float32)
    ln_cs = np.log(rng.gamma(2, 1, (samples, dim))).astype(np.float32)
    betas = rng.uniform(0, 1, (samples, dim)).astype(np.float32)
    return rs, ln_cs, betas


def weighted_jaccard(a, b):
    """sum(min) / sum(max) of two non-negative dense vectors."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(a, b).sum()
    return np.minimum(a, b).sum() / den if den > 0 else 1.0


def hash_jaccard(h1, h2):
    """Fraction of samples whose (column, t) pairs agree (the consistent
    weighted sampling estimator of weighted Jaccard similarity)."""
    h1 = np.asarray(h1).reshape(-1, 2)
    h2 = np.asarray(h2).reshape(-1, 2)
    return float(np.mean(np.all(h1 == h2, axis=1)))


# --------------------------------------------------------------------------
# consumers of the hashes (restated for the GPU kernels of the same name)

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(z):
    """splitmix64 finalizer (Steele, Lea, Flood, OOPSLA 2014) on uint64 arrays."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def band_keys(hashes, bands):
    """LSH band keys [rows][bands] of hashes [rows][samples][2] (uint32):
    h = mix64(g); h = mix64((h ^ (column << 32 | t)) + g) over the band's
    samples in order; 2^64 - 1 for a band holding a sentinel sample."""
    h = np.asarray(hashes, dtype=np.uint32)
    rows, samples, _ = h.shape
    r = samples // bands
    words = (h[..., 0].astype(np.uint64) << np.uint64(32)) | h[..., 1].astype(np.uint64)
    sentinel = (h[..., 0] == SENTINEL) & (h[..., 1] == SENTINEL)
    keys = np.empty((rows, bands), np.uint64)
    with np.errstate(over="ignore"):
        for g in range(bands):
            acc = np.full(rows, mix64(np.uint64(g)), np.uint64)
            for k in range(g * r, (g + 1) * r):
                acc = mix64((acc ^ words[:, k]) + np.uint64(g))
            keys[:, g] = np.where(sentinel[:, g * r:(g + 1) * r].any(axis=1),
                                  np.uint64(0xFFFFFFFFFFFFFFFF), acc)
    return keys


def jaccard_pairs(hashes_a, hashes_b, pairs):
    """Fraction of samples with equal (column, t) pairs for each row pair
    (sentinel samples never count as equal)."""
    a = np.asarray(hashes_a, np.uint32)[np.asarray(pairs)[:, 0]]
    b = np.asarray(hashes_b, np.uint32)[np.asarray(pairs)[:, 1]]
    eq = np.all(a == b, axis=2) & ~((a[..., 0] == SENTINEL) & (a[..., 1] == SENTINEL))
    return (eq.sum(axis=1) / a.shape[1]).astype(np.float32)
