"""Parameter files on the host: read and write the versioned, checksummed format
that ``libMHCUDA.minhash_cuda_save_vars`` / ``minhash_cuda_load_vars`` (C:
``mhcuda_save_random_vars`` / ``mhcuda_load_random_vars``) use, without a GPU.

Typical uses: keep tables made elsewhere (for example the arrays the reference
engine's ``minhash_cuda_retrieve_vars`` returns, test.py:154-182) as a file the
engine loads, or inspect a file's tables as memory-mapped ``(samples, dim)``
float32 arrays.

Format version 1, little-endian: a 64-byte header

====== ==== ===============================================================
offset size field
====== ==== ===============================================================
0      8    magic ``b"WMHPARAM"``
8      4    version (1)
12     4    kind: 1 = tables follow, 2 = seeded (the engine regenerates the
            tables from the seed with its on-device generator)
16     4    dim
20     4    samples
24     4    seed (kind 2, else 0)
28     4    recipe of the on-device generator (kind 2: 1, else 0)
32     24   checksums of rs, ln_cs, betas (kind 1, else 0), uint64 each
56     8    header checksum: h = 0, then h = mix64(h ^ w) over the seven
            uint64 words at offsets 0..48 (mix64: the splitmix64 finalizer)
====== ==== ===============================================================

followed, for kind 1, by rs, ln_cs and betas, each ``[samples][dim]`` float32.
A table's checksum is the sum mod 2^64 over positions p = k * dim + j of
``(bits(table[k, j]) ^ p * 0x9E3779B97F4A7C15) * 0xBF58476D1CE4E5B9``.
"""
import os
import struct

import numpy as np

MAGIC = b"WMHPARAM"
VERSION = 1
KIND_TABLES = 1
KIND_SEEDED = 2
RECIPE = 1
HEADER_BYTES = 64
_HEADER = struct.Struct("<8s6I3QQ")
_POS_MUL = np.uint64(0x9E3779B97F4A7C15)
_SUM_MUL = np.uint64(0xBF58476D1CE4E5B9)
_MASK = (1 << 64) - 1
_CHUNK = 1 << 22  # values per checksum step


def mix64(z):
    """splitmix64 finalizer on a Python int."""
    z &= _MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return z ^ (z >> 31)


def header_checksum(first56):
    """Header checksum of the first 56 header bytes."""
    h = 0
    for (w,) in struct.iter_unpack("<Q", first56):
        h = mix64(h ^ w)
    return h


def table_checksum(table):
    """Checksum of one table (any float32 array of dim * samples values,
    row-major [samples][dim])."""
    bits = np.ascontiguousarray(table, dtype=np.float32).reshape(-1).view(np.uint32)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for s in range(0, bits.size, _CHUNK):
            b = bits[s:s + _CHUNK].astype(np.uint64)
            pos = np.arange(s, s + b.size, dtype=np.uint64)
            total += ((b ^ (pos * _POS_MUL)) * _SUM_MUL).sum(dtype=np.uint64)
    return int(total)


def _pack(kind, dim, samples, seed, recipe, sums):
    head = _HEADER.pack(MAGIC, VERSION, kind, dim, samples, seed, recipe, *sums, 0)
    return head[:56] + struct.pack("<Q", header_checksum(head[:56]))


def write(path, rs, ln_cs, betas):
    """Writes three [samples][dim] float32 tables as a kind-1 file."""
    tables = [np.ascontiguousarray(t, dtype=np.float32) for t in (rs, ln_cs, betas)]
    shape = tables[0].shape
    if len(shape) != 2 or any(t.shape != shape for t in tables):
        raise ValueError("rs, ln_cs and betas must be [samples][dim] arrays of one shape")
    samples, dim = shape
    if not (0 < samples < 1 << 16 and 0 < dim < 1 << 32):
        raise ValueError("samples must be in [1, 65535] and dim in [1, 2^32)")
    sums = [table_checksum(t) for t in tables]
    tmp = "%s.tmp.%d" % (path, os.getpid())
    with open(tmp, "wb") as f:
        f.write(_pack(KIND_TABLES, dim, samples, 0, 0, sums))
        for t in tables:
            t.tofile(f)
    os.replace(tmp, path)


def write_seeded(path, dim, samples, seed):
    """Writes a kind-2 file: the engine regenerates the tables from `seed`."""
    with open(path, "wb") as f:
        f.write(_pack(KIND_SEEDED, dim, samples, seed, RECIPE, [0, 0, 0]))


def read_header(path):
    """The header as a dict; ValueError if the file is not a valid version-1
    parameter file (bad magic, version, kind, header checksum or size)."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) != HEADER_BYTES:
        raise ValueError("%s: too short for a parameter file" % path)
    magic, version, kind, dim, samples, seed, recipe, s0, s1, s2, hsum = _HEADER.unpack(raw)
    if magic != MAGIC or version != VERSION:
        raise ValueError("%s: not a version-%d parameter file" % (path, VERSION))
    if hsum != header_checksum(raw[:56]):
        raise ValueError("%s: header checksum mismatch" % path)
    if dim == 0 or samples == 0 or samples > 0xFFFF or kind not in (KIND_TABLES, KIND_SEEDED):
        raise ValueError("%s: invalid header fields" % path)
    expect = HEADER_BYTES + (12 * dim * samples if kind == KIND_TABLES else 0)
    if size != expect:
        raise ValueError("%s: %d bytes, the header implies %d" % (path, size, expect))
    return {"kind": kind, "dim": dim, "samples": samples, "seed": seed, "recipe": recipe,
            "checksums": (s0, s1, s2)}


def read(path, verify=True, mmap=True):
    """(header, rs, ln_cs, betas) of a kind-1 file; the tables are (samples,
    dim) float32 arrays, memory-mapped read-only unless mmap=False.  With
    verify, each table's checksum is checked (ValueError on mismatch)."""
    h = read_header(path)
    if h["kind"] != KIND_TABLES:
        raise ValueError("%s: a seeded file holds no tables (seed %d)" % (path, h["seed"]))
    shape = (h["samples"], h["dim"])
    n = h["samples"] * h["dim"]
    tables = []
    for i in range(3):
        off = HEADER_BYTES + 4 * n * i
        if mmap:
            t = np.memmap(path, dtype=np.float32, mode="r", offset=off, shape=shape)
        else:
            t = np.fromfile(path, dtype=np.float32, count=n, offset=off).reshape(shape)
        if verify and table_checksum(t) != h["checksums"][i]:
            raise ValueError("%s: checksum mismatch in table %d" % (path, i))
        tables.append(t)
    return (h, *tables)
