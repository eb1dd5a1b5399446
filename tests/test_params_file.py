"""Parameter-file format on the CPU: the committed version-1 fixtures, the host
reader/writer, checksums and the rejection of damaged files.

The fixtures (tests/golden/params_v1*.wmhp, written by make_params_file.py)
are checked field by field at fixed byte offsets with explicit struct
formats, and their checksums against a brute-force pure-Python evaluation of
the definition, so that the numpy implementation and the C++ one (checked
against these files on the GPU) are held to the same bytes.
"""
import os
import struct
import sys

import numpy as np
import pytest

from minhashcuda_amd import params_file as pf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_params_file import recipe_tables  # noqa: E402

FIXTURE = os.path.join(GOLDEN, "params_v1.wmhp")
SEEDED = os.path.join(GOLDEN, "params_v1_seeded.wmhp")
M64 = (1 << 64) - 1


def brute_checksum(table):
    """The table checksum, one Python int at a time."""
    bits = np.ascontiguousarray(table, np.float32).reshape(-1).view(np.uint32).tolist()
    s = 0
    for p, b in enumerate(bits):
        s = (s + ((b ^ ((p * 0x9E3779B97F4A7C15) & M64)) * 0xBF58476D1CE4E5B9)) & M64
    return s


def test_mix64_known_answer():
    """splitmix64 seeded with 0 first outputs 0xe220a8397b1dcdaf: its state
    after one step is the golden gamma, so mix64(gamma) must give it."""
    assert pf.mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


def test_fixture_layout():
    raw = open(FIXTURE, "rb").read()
    rs, ln_cs, betas = recipe_tables()
    assert len(raw) == 64 + 3 * 4 * 5 * 4
    assert raw[0:8] == b"WMHPARAM"
    assert struct.unpack_from("<I", raw, 8)[0] == 1      # version
    assert struct.unpack_from("<I", raw, 12)[0] == 1     # kind: tables
    assert struct.unpack_from("<I", raw, 16)[0] == 5     # dim
    assert struct.unpack_from("<I", raw, 20)[0] == 4     # samples
    assert struct.unpack_from("<II", raw, 24) == (0, 0)  # seed, recipe
    sums = struct.unpack_from("<3Q", raw, 32)
    assert sums == tuple(brute_checksum(t) for t in (rs, ln_cs, betas))
    h = 0
    for (w,) in struct.iter_unpack("<Q", raw[:56]):
        h = pf.mix64(h ^ w)
    assert struct.unpack_from("<Q", raw, 56)[0] == h
    body = np.frombuffer(raw, np.float32, offset=64).reshape(3, 4, 5)
    for got, want in zip(body, (rs, ln_cs, betas)):
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_seeded_fixture_layout():
    raw = open(SEEDED, "rb").read()
    assert len(raw) == 64
    assert raw[:8] == b"WMHPARAM"
    assert struct.unpack_from("<6I", raw, 8) == (1, 2, 1000, 64, 7, 1)
    assert struct.unpack_from("<3Q", raw, 32) == (0, 0, 0)
    h = pf.read_header(SEEDED)
    assert (h["kind"], h["dim"], h["samples"], h["seed"], h["recipe"]) == (2, 1000, 64, 7, 1)
    with pytest.raises(ValueError):
        pf.read(SEEDED)


def test_writer_reproduces_fixture(tmp_path):
    out = tmp_path / "p.wmhp"
    pf.write(str(out), *recipe_tables())
    assert out.read_bytes() == open(FIXTURE, "rb").read()
    seeded = tmp_path / "s.wmhp"
    pf.write_seeded(str(seeded), 1000, 64, 7)
    assert seeded.read_bytes() == open(SEEDED, "rb").read()


@pytest.mark.parametrize("mmap", [True, False])
def test_read_fixture(mmap):
    h, rs, ln_cs, betas = pf.read(FIXTURE, mmap=mmap)
    assert (h["kind"], h["dim"], h["samples"]) == (1, 5, 4)
    for got, want in zip((rs, ln_cs, betas), recipe_tables()):
        assert got.shape == (4, 5) and got.dtype == np.float32
        assert np.array_equal(np.asarray(got).view(np.uint32), want.view(np.uint32))


def test_checksum_matches_brute_force_and_chunking():
    rng = np.random.default_rng(5)
    t = rng.standard_normal((3, 1001)).astype(np.float32)
    assert pf.table_checksum(t) == brute_checksum(t)
    old = pf._CHUNK
    try:
        pf._CHUNK = 7  # chunk boundaries inside rows
        assert pf.table_checksum(t) == brute_checksum(t)
    finally:
        pf._CHUNK = old


def test_checksum_sees_single_bit_flips_and_swaps():
    rng = np.random.default_rng(6)
    t = rng.standard_normal((4, 64)).astype(np.float32)
    base = pf.table_checksum(t)
    bits = t.view(np.uint32)
    for k, j, b in [(0, 0, 0), (3, 63, 31), (2, 17, 22), (1, 40, 7)]:
        u = bits.copy()
        u[k, j] ^= np.uint32(1 << b)
        assert pf.table_checksum(u.view(np.float32)) != base
    u = bits.copy()
    u[0, 1], u[0, 2] = bits[0, 2], bits[0, 1]
    assert pf.table_checksum(u.view(np.float32)) != base


def test_round_trip_random_tables(tmp_path):
    rng = np.random.default_rng(7)
    tabs = [rng.standard_normal((6, 333)).astype(np.float32) for _ in range(3)]
    path = str(tmp_path / "r.wmhp")
    pf.write(path, *tabs)
    h, *got = pf.read(path)
    assert (h["dim"], h["samples"]) == (333, 6)
    for a, b in zip(got, tabs):
        assert np.array_equal(np.asarray(a), b)


def _damaged(tmp_path, name, edit):
    raw = bytearray(open(FIXTURE, "rb").read())
    raw = edit(raw)
    p = tmp_path / name
    p.write_bytes(bytes(raw))
    return str(p)


@pytest.mark.parametrize("name,edit", [
    ("magic", lambda r: r.__setitem__(0, ord("X")) or r),
    ("version", lambda r: r.__setitem__(8, 2) or r),
    ("headersum", lambda r: r.__setitem__(20, 5) or r),  # samples 4 -> 5, checksum stale
    ("truncated", lambda r: r[:-4]),
    ("extended", lambda r: r + b"\0\0\0\0"),
    ("short", lambda r: r[:40]),
])
def test_rejects_malformed(tmp_path, name, edit):
    with pytest.raises(ValueError):
        pf.read_header(_damaged(tmp_path, name, edit))


def test_rejects_corrupt_table(tmp_path):
    def flip(r):
        r[64 + 4 * 20 + 5] ^= 0x10  # a bit of ln_cs
        return r
    path = _damaged(tmp_path, "table", flip)
    pf.read_header(path)  # the header is intact
    with pytest.raises(ValueError, match="table 1"):
        pf.read(path)
    pf.read(path, verify=False)


def test_writer_rejects_bad_shapes(tmp_path):
    a = np.zeros((4, 5), np.float32)
    with pytest.raises(ValueError):
        pf.write(str(tmp_path / "x"), a, a, np.zeros((5, 4), np.float32))
    with pytest.raises(ValueError):
        pf.write(str(tmp_path / "y"), a[0], a[0], a[0])


def test_c_header_reader(capi, tmp_path):
    """mhcuda_params_file_info (file parsing only, no device work) reads the
    committed fixtures and refuses damaged or missing files."""
    import ctypes
    d, s = ctypes.c_uint32(0), ctypes.c_uint16(0)
    assert capi.mhcuda_params_file_info(FIXTURE.encode(), ctypes.byref(d), ctypes.byref(s)) == 0
    assert (d.value, s.value) == (5, 4)
    assert capi.mhcuda_params_file_info(SEEDED.encode(), ctypes.byref(d), ctypes.byref(s)) == 0
    assert (d.value, s.value) == (1000, 64)
    for name, edit in [("magic", lambda r: r.__setitem__(0, ord("X")) or r),
                       ("headersum", lambda r: r.__setitem__(20, 5) or r),
                       ("truncated", lambda r: r[:-4])]:
        bad = _damaged(tmp_path, name, edit)
        assert capi.mhcuda_params_file_info(bad.encode(), None, None) == 1  # InvalidArguments
    assert capi.mhcuda_params_file_info(str(tmp_path / "missing").encode(), None, None) == 5
    assert capi.mhcuda_params_file_info(None, None, None) == 1
