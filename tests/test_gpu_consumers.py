"""GPU consumers of the hashes: weighted-Jaccard estimates of row pairs and
LSH band keys, bit-exact against the oracle's restatements."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hashed(libmhcuda, gpu_available):
    data = oracle.c1_data(1)
    data[7] = 0                       # an empty row
    data[9] = data[8]                 # a duplicate row
    gen = libmhcuda.minhash_cuda_init(130, 128, seed=1)
    h = libmhcuda.minhash_cuda_calc(gen, sp.csr_matrix(data, dtype=np.float32))
    yield gen, data, h
    libmhcuda.minhash_cuda_fini(gen)


def test_band_keys_match_oracle(capi, hashed):
    gen, data, h = hashed
    dh = torch.from_numpy(h.view(np.int32)).cuda()
    for bands in (1, 16, 32, 128):
        keys = torch.empty((h.shape[0], bands), dtype=torch.int64, device="cuda")
        assert capi.mhcuda_lsh_band_keys(ctypes.c_void_p(gen), dh.data_ptr(), h.shape[0], bands,
                                         keys.data_ptr(), None) == 0
        got = keys.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, oracle.band_keys(h, bands)), bands
        assert (got[7] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
        assert (got[8] == got[9]).all()
    assert capi.mhcuda_lsh_band_keys(ctypes.c_void_p(gen), dh.data_ptr(), 5, 3,
                                     keys.data_ptr(), None) == 1   # 3 does not divide 128


def test_jaccard_pairs_match_oracle(capi, hashed):
    gen, data, h = hashed
    dh = torch.from_numpy(h.view(np.int32)).cuda()
    rng = np.random.default_rng(0)
    pairs = rng.integers(0, h.shape[0], (5000, 2)).astype(np.uint32)
    pairs[:3] = [[8, 9], [7, 7], [0, 0]]
    dp = torch.from_numpy(pairs.view(np.int32)).cuda()
    out = torch.empty(len(pairs), dtype=torch.float32, device="cuda")
    assert capi.mhcuda_jaccard_pairs(ctypes.c_void_p(gen), dh.data_ptr(), dh.data_ptr(),
                                     dp.data_ptr(), len(pairs), out.data_ptr(), None) == 0
    got = out.cpu().numpy()
    assert np.array_equal(got, oracle.jaccard_pairs(h, h, pairs))
    assert got[0] == 1.0 and got[1] == 0.0 and got[2] == 1.0
    # the estimates track the true weighted Jaccard similarity
    true = np.array([oracle.weighted_jaccard(data[a], data[b]) for a, b in pairs[3:1003]])
    assert abs(float(np.mean(got[3:1003] - true))) < 0.01


def test_candidates_are_enriched_in_similar_pairs(capi, hashed):
    """Through 16 bands of 8 samples a duplicate row is always a candidate,
    and candidate pairs are far more similar than random pairs (collision
    probability 1 - (1 - J^8)^16 rises steeply with J)."""
    from minhashcuda_amd import lsh
    gen, data, h = hashed
    dh = torch.from_numpy(h.view(np.int32)).cuda()
    keys = torch.empty((h.shape[0], 16), dtype=torch.int64, device="cuda")
    assert capi.mhcuda_lsh_band_keys(ctypes.c_void_p(gen), dh.data_ptr(), h.shape[0], 16,
                                     keys.data_ptr(), None) == 0
    pairs = lsh.candidate_pairs_gpu(gen, keys)
    assert np.array_equal(pairs, lsh.candidate_pairs(keys))
    assert (8, 9) in {tuple(p) for p in pairs.tolist()}
    rng = np.random.default_rng(1)
    cand = pairs[rng.choice(len(pairs), min(len(pairs), 2000), replace=False)]
    rand = rng.integers(0, h.shape[0], (2000, 2))
    jc = np.mean([oracle.weighted_jaccard(data[a], data[b]) for a, b in cand])
    jr = np.mean([oracle.weighted_jaccard(data[a], data[b]) for a, b in rand if a != b])
    assert jc > jr + 0.15, (jc, jr)


@pytest.mark.parametrize("bands,max_bucket", [(32, 64), (16, 3), (128, 2)])
def test_gpu_candidates_equal_numpy_on_hashes(capi, hashed, bands, max_bucket):
    """GPU candidate generation (mhcuda_lsh_candidates) equals the numpy
    restatement lsh.candidate_pairs on the 6400 x 130 hashes, including the
    empty row's reserved keys and buckets above max_bucket."""
    from minhashcuda_amd import lsh
    gen, data, h = hashed
    dh = torch.from_numpy(h.view(np.int32)).cuda()
    keys = torch.empty((h.shape[0], bands), dtype=torch.int64, device="cuda")
    assert capi.mhcuda_lsh_band_keys(ctypes.c_void_p(gen), dh.data_ptr(), h.shape[0], bands,
                                     keys.data_ptr(), None) == 0
    got = lsh.candidate_pairs_gpu(gen, keys, max_bucket=max_bucket)
    want = lsh.candidate_pairs(keys, max_bucket=max_bucket)
    assert len(want) > 0 and np.array_equal(got, want)


def test_gpu_candidates_equal_numpy_on_a_million_rows(capi, hashed):
    """1.2M rows x 8 bands of keys with planted collisions (buckets of about
    3 rows), reserved keys and buckets larger than max_bucket: the GPU pairs
    equal the numpy ones; a too small output buffer reports the count."""
    from minhashcuda_amd import lsh
    gen = hashed[0]
    rows, bands = 1_200_000, 8
    rng = np.random.default_rng(5)
    k = rng.integers(0, rows // 3, (rows, bands)).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    k[rng.random((rows, bands)) < 0.01] = np.uint64(0xFFFFFFFFFFFFFFFF)
    k[:70, 3] = np.uint64(12345)          # one bucket of 70 rows (> 64: skipped)
    k[100:164, 5] = np.uint64(777)        # one bucket of exactly 64 rows (kept)
    keys = torch.from_numpy(k.view(np.int64)).cuda()
    got = lsh.candidate_pairs_gpu(gen, keys, max_bucket=64)
    want = lsh.candidate_pairs(k, max_bucket=64)
    assert len(want) > 1_000_000 and np.array_equal(got, want)
    count = ctypes.c_uint64(0)
    out = torch.empty(10, dtype=torch.int64, device="cuda")
    assert capi.mhcuda_lsh_candidates(ctypes.c_void_p(gen), keys.data_ptr(), rows, bands, 64,
                                      out.data_ptr(), 10, ctypes.byref(count), None) == 0
    assert count.value == len(want)
    codes = out.cpu().numpy().view(np.uint64)
    assert np.array_equal(codes, (want[:10, 0].astype(np.uint64) << np.uint64(32)) |
                          want[:10, 1].astype(np.uint64))
