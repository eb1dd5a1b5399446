"""Parameter files on the GPU: mhcuda_save_random_vars / mhcuda_load_random_vars
(Python: minhash_cuda_save_vars / minhash_cuda_load_vars).

The files the library writes must be byte-identical to the host writer's file
for the same tables (whose format the CPU tests pin against committed
fixtures), loading must reproduce the tables bit for bit and therefore the
hashes, and damaged or mismatched files must be refused.  The flow extends
the reference's deferred-init test (test.py:154-182) from arrays to a file.
"""
import ctypes
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from minhashcuda_amd import params_file as pf

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_params_file import recipe_tables  # noqa: E402


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def L(libmhcuda, gpu_available, monkeypatch):
    # small pinned chunks, so that even small tables stream in several pieces
    monkeypatch.setenv("MHCUDA_PARAMS_CHUNK_BYTES", "65536")
    return libmhcuda


def random_csr(rows, dim, seed, density=0.05):
    m = sp.random(rows, dim, density=density, format="csr", dtype=np.float32,
                  random_state=np.random.RandomState(seed))
    m.data = (m.data * 10 + 0.01).astype(np.float32)
    return m


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same_tables(a, b):
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


def test_save_load_tables(L, tmp_path):
    dim, samples = 1000, 128  # 128 columns per chunk: 8 chunks, the last partial
    say("save/load tables", dim, samples)
    path = str(tmp_path / "gen.wmhp")
    gen = L.minhash_cuda_init(dim, samples, seed=11, devices=1)
    try:
        L.minhash_cuda_save_vars(gen, path)
        tables = L.minhash_cuda_retrieve_vars(gen)
        m = random_csr(300, dim, 1)
        h1 = L.minhash_cuda_calc(gen, m)
    finally:
        L.minhash_cuda_fini(gen)
    head, *got = pf.read(path)  # verifies the checksums the library wrote
    assert (head["kind"], head["dim"], head["samples"]) == (1, dim, samples)
    assert same_tables(got, tables)
    ref = str(tmp_path / "host.wmhp")
    pf.write(ref, *tables)
    assert open(path, "rb").read() == open(ref, "rb").read()

    gen2 = L.minhash_cuda_init(dim, samples, deferred=True, devices=1)
    try:
        L.minhash_cuda_load_vars(gen2, path)
        assert same_tables(L.minhash_cuda_retrieve_vars(gen2), tables)
        h2 = L.minhash_cuda_calc(gen2, m)
    finally:
        L.minhash_cuda_fini(gen2)
    assert np.array_equal(h1, h2)
    assert np.array_equal(h2, oracle.calc_csr(*tables, m))


def test_seed_only(L, tmp_path):
    dim, samples = 777, 64
    say("seed-only file", dim, samples)
    path = str(tmp_path / "seed.wmhp")
    gen = L.minhash_cuda_init(dim, samples, seed=123, devices=1)
    try:
        L.minhash_cuda_save_vars(gen, path, seed_only=True)
        tables = L.minhash_cuda_retrieve_vars(gen)
    finally:
        L.minhash_cuda_fini(gen)
    assert os.path.getsize(path) == 64
    head = pf.read_header(path)
    assert (head["kind"], head["seed"], head["recipe"]) == (2, 123, 1)
    gen2 = L.minhash_cuda_init(dim, samples, deferred=True, devices=1)
    try:
        L.minhash_cuda_load_vars(gen2, path)
        assert same_tables(L.minhash_cuda_retrieve_vars(gen2), tables)
        # a generator loaded from a seed saves the same seed-only file again
        again = str(tmp_path / "again.wmhp")
        L.minhash_cuda_save_vars(gen2, again, seed_only=True)
        assert open(again, "rb").read() == open(path, "rb").read()
    finally:
        L.minhash_cuda_fini(gen2)


def test_seeded_fixture_regenerates_seeded_tables(L):
    """The committed seed-only file (dim 1000, samples 64, seed 7) yields the
    tables of mhcuda_init(1000, 64, seed=7)."""
    gen = L.minhash_cuda_init(1000, 64, seed=7, devices=1)
    gen2 = L.minhash_cuda_init(1000, 64, deferred=True, devices=1)
    try:
        L.minhash_cuda_load_vars(gen2, os.path.join(GOLDEN, "params_v1_seeded.wmhp"))
        assert same_tables(L.minhash_cuda_retrieve_vars(gen2), L.minhash_cuda_retrieve_vars(gen))
    finally:
        L.minhash_cuda_fini(gen)
        L.minhash_cuda_fini(gen2)


def test_table_fixture_hashes(L):
    """The committed 4 x 5 table file loads bit-exactly and hashes like the
    oracle with the recipe's tables."""
    rs, ln_cs, betas = recipe_tables()
    m = sp.csr_matrix(np.array([[0, 1.5, 0, 2.0, 0.25], [3.0, 0, 0, 0, 0], [0, 0, 0, 0, 0],
                                [1, 1, 1, 1, 1]], np.float32))
    gen = L.minhash_cuda_init(5, 4, deferred=True, devices=1)
    try:
        L.minhash_cuda_load_vars(gen, os.path.join(GOLDEN, "params_v1.wmhp"))
        assert same_tables(L.minhash_cuda_retrieve_vars(gen), (rs, ln_cs, betas))
        h = L.minhash_cuda_calc(gen, m)
    finally:
        L.minhash_cuda_fini(gen)
    assert np.array_equal(h, oracle.calc_csr(rs, ln_cs, betas, m))


def test_assigned_tables_are_saved_whole(L, tmp_path):
    dim, samples = 300, 32
    rng = np.random.RandomState(4)
    tabs = (rng.gamma(2, 1, (samples, dim)).astype(np.float32),
            np.log(rng.gamma(2, 1, (samples, dim))).astype(np.float32),
            rng.uniform(0, 1, (samples, dim)).astype(np.float32))
    path = str(tmp_path / "a.wmhp")
    gen = L.minhash_cuda_init(dim, samples, seed=5, devices=1)
    try:
        L.minhash_cuda_assign_vars(gen, *tabs)
        L.minhash_cuda_save_vars(gen, path, seed_only=True)  # not seeded any more: tables
    finally:
        L.minhash_cuda_fini(gen)
    head, *got = pf.read(path)
    assert head["kind"] == 1 and same_tables(got, tabs)


def test_refusals(L, tmp_path):
    say("refusals")
    path = str(tmp_path / "t.wmhp")
    gen = L.minhash_cuda_init(200, 16, seed=3, devices=1)
    deferred = L.minhash_cuda_init(200, 16, deferred=True, devices=1)
    other = L.minhash_cuda_init(201, 16, deferred=True, devices=1)
    try:
        with pytest.raises(ValueError):  # deferred, never assigned: nothing to save
            L.minhash_cuda_save_vars(deferred, path)
        assert not os.path.exists(path)
        L.minhash_cuda_save_vars(gen, path)
        tables = L.minhash_cuda_retrieve_vars(gen)
        with pytest.raises(ValueError):  # another dim
            L.minhash_cuda_load_vars(other, path)
        with pytest.raises(OSError):
            L.minhash_cuda_load_vars(deferred, str(tmp_path / "missing.wmhp"))
        with pytest.raises(OSError):
            L.minhash_cuda_save_vars(gen, str(tmp_path / "no" / "such" / "dir.wmhp"))
        # a flipped bit in betas: refused after streaming, tables zeroed
        L.minhash_cuda_load_vars(deferred, path)
        assert same_tables(L.minhash_cuda_retrieve_vars(deferred), tables)
        raw = bytearray(open(path, "rb").read())
        raw[64 + 2 * 4 * 200 * 16 + 4 * 777] ^= 0x01
        bad = str(tmp_path / "bad.wmhp")
        open(bad, "wb").write(bytes(raw))
        with pytest.raises(ValueError):
            L.minhash_cuda_load_vars(deferred, bad)
        for t in L.minhash_cuda_retrieve_vars(deferred):
            assert not t.any()
        with pytest.raises(ValueError):  # tables gone: nothing to save
            L.minhash_cuda_save_vars(deferred, str(tmp_path / "after.wmhp"))
        # a damaged header is refused before any table is touched
        raw = bytearray(open(path, "rb").read())
        raw[16] ^= 0x01
        open(bad, "wb").write(bytes(raw))
        L.minhash_cuda_load_vars(deferred, path)
        with pytest.raises(ValueError):
            L.minhash_cuda_load_vars(deferred, bad)
        assert same_tables(L.minhash_cuda_retrieve_vars(deferred), tables)
    finally:
        for g in (gen, deferred, other):
            L.minhash_cuda_fini(g)


def test_c_abi(capi, gpu_available, tmp_path):
    say("C ABI parameter files")
    path = str(tmp_path / "c.wmhp").encode()
    st = ctypes.c_int(0)
    gen = capi.mhcuda_init(96, 8, 9, 0, 1, 0, ctypes.byref(st))
    assert gen
    gen2 = capi.mhcuda_init(96, 8, 0, 1, 1, 0, ctypes.byref(st))
    assert gen2
    try:
        assert capi.mhcuda_save_random_vars(gen, path, 0) == 0
        d, s = ctypes.c_uint32(0), ctypes.c_uint16(0)
        assert capi.mhcuda_params_file_info(path, ctypes.byref(d), ctypes.byref(s)) == 0
        assert (d.value, s.value) == (96, 8)
        assert capi.mhcuda_load_random_vars(gen2, path) == 0
        assert capi.mhcuda_save_random_vars(None, path, 0) == 1
        assert capi.mhcuda_load_random_vars(gen2, None) == 1
        assert capi.mhcuda_load_random_vars(gen2, str(tmp_path / "nope").encode()) == 5
        bufs = [[np.empty((8, 96), np.float32) for _ in range(3)] for _ in range(2)]
        for g, b in zip((gen, gen2), bufs):
            assert capi.mhcuda_retrieve_random_vars(g, *[x.ctypes.data for x in b]) == 0
        assert same_tables(bufs[0], bufs[1])
    finally:
        capi.mhcuda_fini(gen)
        capi.mhcuda_fini(gen2)
