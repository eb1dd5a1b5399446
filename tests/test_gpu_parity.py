"""GPU parity tests: the HIP path against the CPU oracle, bit for bit.

They mirror the reference's test.py (test_calc_tiny, test_calc_big, test_slice,
test_backwards, test_deferred, test_float, test_random_vars,
test_integration), with the live datasketch oracle replaced by this
repository's oracle and golden vectors, and add edge rows, sample counts,
tuning variants and the device/batch entry points of the C ABI.
"""
import ctypes
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.stats import gamma, uniform

import oracle
from conftest import torch
from minhashcuda_amd import capi as capi_mod

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
S = oracle.SENTINEL


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


@pytest.fixture
def L(libmhcuda, gpu_available):
    return libmhcuda


def csr_of(g):
    return sp.csr_matrix((g["data"], g["indices"], g["indptr"]), shape=tuple(g["shape"]))


# ------------------------------------------------------------ reference tests

def test_calc_tiny(L):
    """test.py:12-32 with the golden parameters assigned."""
    g = golden("tiny.npz")
    gen = L.minhash_cuda_init(23, 128, devices=1, verbosity=2)
    L.minhash_cuda_assign_vars(gen, g["rs"], g["ln_cs"], g["betas"])
    h = L.minhash_cuda_calc(gen, csr_of(g))
    L.minhash_cuda_fini(gen)
    assert h.shape == (2, 128, 2) and h.dtype == np.uint32
    assert np.array_equal(h, g["hashes"])


@pytest.mark.parametrize("seed", [0, 1])
def test_calc_big(L, seed):
    """test.py:34-67: the 6400x130 recipe, full output equal to the oracle."""
    data = oracle.c1_data(seed)
    rs, ln_cs, betas = oracle.datasketch_style_params(130, 128, seed=1)
    gen = L.minhash_cuda_init(130, 128, devices=1)
    L.minhash_cuda_assign_vars(gen, rs, ln_cs, betas)
    m = sp.csr_matrix(data, dtype=np.float32)
    h = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    assert h.shape == (6400, 128, 2)
    assert np.array_equal(h, oracle.calc_csr(rs, ln_cs, betas, m))


def test_golden_c1_rows(L):
    g = golden("c1_rows256.npz")
    gen = L.minhash_cuda_init(130, 128)
    L.minhash_cuda_assign_vars(gen, g["rs"], g["ln_cs"], g["betas"])
    h = L.minhash_cuda_calc(gen, csr_of(g))
    L.minhash_cuda_fini(gen)
    assert np.array_equal(h, g["hashes"])


def test_slice(L):
    """test.py:112-124."""
    data = oracle.c1_data(0)
    gen = L.minhash_cuda_init(130, 128, seed=3)
    m = sp.csr_matrix(data, dtype=np.float32)
    h = L.minhash_cuda_calc(gen, m)
    h2 = L.minhash_cuda_calc(gen, m, row_start=3200, row_finish=4800)
    h3 = L.minhash_cuda_calc(gen, m, row_start=6399)
    L.minhash_cuda_fini(gen)
    assert np.array_equal(h[3200:4800], h2)
    assert np.array_equal(h[6399:], h3)


def test_backwards(L):
    """test.py:126-152: device-generated parameters, retrieved, reproduce the
    device hashes on the CPU -- and equal the CPU restatement of the Philox
    generator bit for bit."""
    g = golden("tiny.npz")
    gen = L.minhash_cuda_init(23, 128, seed=12345, devices=1)
    rs, ln_cs, betas = L.minhash_cuda_retrieve_vars(gen)
    assert rs.shape == ln_cs.shape == betas.shape == (128, 23)
    for a, b in zip((rs, ln_cs, betas), oracle.gen_params(12345, 23, 128)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    m = csr_of(g)
    h = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    assert np.array_equal(h, oracle.calc_csr(rs, ln_cs, betas, m))


def test_deferred(L):
    """test.py:154-182: retrieve -> deferred generator -> assign -> same hashes."""
    g = golden("tiny.npz")
    m = csr_of(g)
    gen = L.minhash_cuda_init(23, 128, seed=7, devices=1)
    saved = L.minhash_cuda_retrieve_vars(gen)
    h1 = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    gen = L.minhash_cuda_init(23, 128, devices=1, deferred=True)
    L.minhash_cuda_assign_vars(gen, *saved)
    again = L.minhash_cuda_retrieve_vars(gen)
    h2 = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    assert all(np.array_equal(a, b) for a, b in zip(saved, again))
    assert np.array_equal(h1, h2)
    assert np.array_equal(h2, oracle.calc_csr(*saved, m))


def test_float(L):
    """test.py:184-210: weights < 1 give negative t (compared as int32)."""
    g = golden("float.npz")
    gen = L.minhash_cuda_init(20, 128, devices=1, seed=7)
    rs, ln_cs, betas = L.minhash_cuda_retrieve_vars(gen)
    assert np.array_equal(rs, g["rs"]) and np.array_equal(ln_cs, g["ln_cs"])
    assert np.array_equal(betas, g["betas"])
    h = L.minhash_cuda_calc(gen, csr_of(g)).astype(np.int32)
    L.minhash_cuda_fini(gen)
    assert np.array_equal(h, g["hashes"].astype(np.int32))


def test_random_vars(L):
    """test.py:72-90, plus bit-exact agreement with the golden Philox sample."""
    gen = L.minhash_cuda_init(1000, 128, devices=1, seed=1)
    rs, ln_cs, betas = L.minhash_cuda_retrieve_vars(gen)
    L.minhash_cuda_fini(gen)
    g = golden("philox_params.npz")
    for name, arr in (("rs", rs), ("ln_cs", ln_cs), ("betas", betas)):
        assert np.array_equal(arr.ravel()[g["index"]], g[name]), name
    a, loc, scale = gamma.fit(rs.ravel())
    assert 1.97 < a < 2.03 and -0.01 < loc < 0.01 and 0.98 < scale < 1.02
    a, loc, scale = gamma.fit(np.exp(ln_cs.ravel()))
    assert 1.97 < a < 2.03 and -0.01 < loc < 0.01 and 0.98 < scale < 1.02
    bmin, bmax = uniform.fit(betas.ravel())
    assert 0 <= bmin < 0.001 and 0.999 <= bmax <= 1


def test_integration(L):
    """test.py:92-110 with a sound tolerance: the mean over many row pairs of
    the hash-agreement rate tracks the true weighted Jaccard similarity."""
    data = oracle.c1_data(1)
    gen = L.minhash_cuda_init(130, 128, seed=1, verbosity=1)
    h = L.minhash_cuda_calc(gen, sp.csr_matrix(data, dtype=np.float32))
    L.minhash_cuda_fini(gen)
    pairs = [(i, i + 1) for i in range(0, 2000, 2)]
    est = np.array([oracle.hash_jaccard(h[a], h[b]) for a, b in pairs])
    true = np.array([oracle.weighted_jaccard(data[a], data[b]) for a, b in pairs])
    assert abs((est - true).mean()) < 0.01
    assert np.sqrt(np.mean((est - true) ** 2)) < 0.05


# ------------------------------------------------------------ beyond test.py

def test_edge_rows(L, capi):
    """Empty rows, zero/negative/NaN/inf/subnormal weights, duplicate columns,
    out-of-range columns: equal to the golden oracle output (C ABI)."""
    g = golden("edges.npz")
    dim = int(g["dim"])
    samples = g["rs"].shape[0]
    gen = L.minhash_cuda_init(dim, samples)
    L.minhash_cuda_assign_vars(gen, g["rs"], g["ln_cs"], g["betas"])
    data = np.ascontiguousarray(g["data"], np.float32)
    cols = np.ascontiguousarray(g["indices"], np.uint32)
    rows = np.ascontiguousarray(g["indptr"], np.uint32)
    out = np.zeros((len(rows) - 1, samples, 2), np.uint32)
    rc = capi.mhcuda_calc(ctypes.c_void_p(gen), data.ctypes.data, cols.ctypes.data,
                          rows.ctypes.data, len(rows) - 1, out.ctypes.data)
    L.minhash_cuda_fini(gen)
    assert rc == 0
    assert np.array_equal(out, g["hashes"])


@pytest.mark.parametrize("samples", [2, 64, 100, 128, 256, 200])
@pytest.mark.parametrize("spl", [1, 2, 4])
def test_samples_and_lane_widths(L, capi, samples, spl):
    """Every sample count (including ones that are not multiples of the
    wavefront tile) and every samples-per-lane variant give the oracle's
    hashes."""
    say("samples", samples, "spl", spl)
    data = oracle.c1_data(5, shape=(700, 90))
    data[3] = 0  # an empty row
    m = sp.csr_matrix(data, dtype=np.float32)
    gen = L.minhash_cuda_init(90, samples, seed=samples + spl)
    assert capi.mhcuda_set_tuning(ctypes.c_void_p(gen), spl, 0, 0) == 0
    params = L.minhash_cuda_retrieve_vars(gen)
    h = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    assert np.array_equal(h, oracle.calc_csr(*params, m))


def test_odd_samples_rejected(L):
    """minhashcuda.cc:602-618: no even sample_delta -> invalid arguments."""
    gen = L.minhash_cuda_init(10, 63)
    with pytest.raises(ValueError):
        L.minhash_cuda_calc(gen, sp.csr_matrix(np.eye(10, dtype=np.float32)))
    L.minhash_cuda_fini(gen)


def test_python_input_checks(L):
    """python.cc:305-347: non-CSR, float64 data, bad slices."""
    gen = L.minhash_cuda_init(10, 16)
    dense = np.eye(10, dtype=np.float32)
    with pytest.raises(TypeError):
        L.minhash_cuda_calc(gen, dense)
    with pytest.raises(TypeError):
        L.minhash_cuda_calc(gen, sp.csr_array(dense))
    with pytest.raises(ValueError):
        L.minhash_cuda_calc(gen, sp.csr_matrix(dense.astype(np.float64)))
    with pytest.raises(ValueError):
        L.minhash_cuda_calc(gen, sp.csr_matrix(dense), row_start=10)
    with pytest.raises(ValueError):
        L.minhash_cuda_calc(gen, sp.csr_matrix(dense), row_start=5, row_finish=5)
    with pytest.raises(ValueError):
        L.minhash_cuda_assign_vars(gen, np.zeros(5, np.float32), np.zeros(160, np.float32),
                                   np.zeros(160, np.float32))
    # int64 indices / indptr are force-cast (python.cc:317,323): same hashes
    m = sp.csr_matrix(dense)
    want = L.minhash_cuda_calc(gen, m)
    m64 = sp.csr_matrix(dense)
    m64.indices = m64.indices.astype(np.int64)
    m64.indptr = m64.indptr.astype(np.int64)
    assert np.array_equal(L.minhash_cuda_calc(gen, m64), want)
    # row_finish past the end is clamped (python.cc:341)
    assert np.array_equal(L.minhash_cuda_calc(gen, m, row_start=3, row_finish=1000), want[3:])
    L.minhash_cuda_fini(gen)


def test_python_batch_equals_single(L):
    mats = [sp.csr_matrix(oracle.c1_data(s, shape=(n, 64)), dtype=np.float32)
            for s, n in ((1, 50), (2, 1), (3, 333))]
    gen = L.minhash_cuda_init(64, 128, seed=4)
    params = L.minhash_cuda_retrieve_vars(gen)
    hs = L.minhash_cuda_calc_batch(gen, mats)
    L.minhash_cuda_fini(gen)
    for m, h in zip(mats, hs):
        assert np.array_equal(h, oracle.calc_csr(*params, m))


def poisson_csr(rows, dim, mean, seed):
    """rows x dim CSR, nnz/row ~ Poisson(mean), sorted distinct uniform
    columns, lognormal(0, 1) weights."""
    rng = np.random.default_rng(seed)
    nnz = np.minimum(rng.poisson(mean, rows), dim)
    indptr = np.zeros(rows + 1, np.int64)
    np.cumsum(nnz, out=indptr[1:])
    cols = np.empty(indptr[-1], np.uint32)
    # stratified draw: the i-th of n columns falls in [i*dim/n, (i+1)*dim/n)
    pos = np.arange(indptr[-1]) - np.repeat(indptr[:-1], nnz)
    n = np.repeat(nnz, nnz).astype(np.int64)
    lo = pos * dim // np.maximum(n, 1)
    hi = (pos + 1) * dim // np.maximum(n, 1)
    cols[:] = lo + (rng.random(indptr[-1]) * (hi - lo)).astype(np.int64)
    w = rng.lognormal(0.0, 1.0, indptr[-1]).astype(np.float32)
    return sp.csr_matrix((w, cols, indptr), shape=(rows, dim))


def test_plain_config_full_parity(L):
    """100,000 x 2^20, nnz/row ~ Poisson(100), lognormal(0,1) weights,
    128 samples: every hash equals the oracle's (OpenMP oracle)."""
    say("building 100000 x 2^20 input")
    m = poisson_csr(100000, 1 << 20, 100, seed=2)
    gen = L.minhash_cuda_init(1 << 20, 128, seed=2)
    params = L.minhash_cuda_retrieve_vars(gen)
    say("hashing on the GPU")
    h = L.minhash_cuda_calc(gen, m)
    L.minhash_cuda_fini(gen)
    say("oracle")
    ref, _ = oracle.calc_omp(*params, m.data, m.indices, m.indptr, 16)
    assert np.array_equal(h, ref)


def test_device_entry_point_with_torch_buffers(L, capi):
    """mhcuda_calc_device on torch-allocated device buffers and torch's
    current stream, and mhcuda_calc_batch_device over several matrices."""
    mats = [poisson_csr(n, 5000, 30, seed=n) for n in (1000, 7, 2500)]
    gen = L.minhash_cuda_init(5000, 128, seed=9)
    params = L.minhash_cuda_retrieve_vars(gen)
    dev = []
    descs = (capi_mod.Matrix * len(mats))()
    for i, m in enumerate(mats):
        w = torch.from_numpy(m.data.astype(np.float32)).cuda()
        c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
        r = torch.from_numpy(m.indptr.astype(np.int32)).cuda()
        o = torch.zeros((m.shape[0], 128, 2), dtype=torch.int32, device="cuda")
        dev.append((w, c, r, o))
        descs[i] = capi_mod.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), m.shape[0],
                                   o.data_ptr())
    stream = torch.cuda.current_stream().cuda_stream
    w, c, r, o = dev[0]
    assert capi.mhcuda_calc_device(ctypes.c_void_p(gen), w.data_ptr(), c.data_ptr(),
                                   r.data_ptr(), mats[0].shape[0], o.data_ptr(),
                                   ctypes.c_void_p(stream)) == 0
    torch.cuda.synchronize()
    got0 = o.cpu().numpy().view(np.uint32).copy()
    assert np.array_equal(got0, oracle.calc_csr(*params, mats[0]))
    for _, _, _, o in dev:
        o.zero_()
    for kernel in (capi_mod.KERNEL_PLAIN, capi_mod.KERNEL_AUTO):
        assert capi.mhcuda_calc_batch_device(ctypes.c_void_p(gen), len(mats), descs, kernel,
                                             ctypes.c_void_p(stream)) == 0
        torch.cuda.synchronize()
        for m, (_, _, _, o) in zip(mats, dev):
            assert np.array_equal(o.cpu().numpy().view(np.uint32), oracle.calc_csr(*params, m))
    L.minhash_cuda_fini(gen)
