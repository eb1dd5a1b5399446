"""CSR input from disk without loading it: memory-mapped ``.npz`` matrices and
row blocks of bounded size.

``scipy.sparse.save_npz(path, m, compressed=False)`` stores ``data``,
``indices`` and ``indptr`` as uncompressed ``.npy`` members of a zip file;
their bytes can be memory-mapped in place.  ``load_npz_csr`` returns a
``csr_matrix`` over such maps (pages are read from disk when the engine's
uploads touch them), which ``libMHCUDA.minhash_cuda_calc`` hashes as it is:
matrices of 2^32 or more nonzeros are cut into row slices inside the module
(the reference reads the whole matrix into memory and force-casts indptr to
uint32, python.cc:295-333).  ``iter_row_blocks`` yields consecutive row
blocks of at most ``max_nnz`` nonzeros for callers that want to bound the host
memory a block pins or to interleave hashing with other work.

The matrices are assembled without scipy's constructor, which would scan
64-bit index arrays in full to pick an index dtype.
"""
import struct
import zipfile

import numpy as np
import scipy.sparse as sp

_LOCAL_HEADER = struct.Struct("<IHHHHHIIIHH")  # zip local file header (30 bytes)
_LOCAL_SIG = 0x04034B50


def _npy_member(path, zf, name):
    """A read-only memmap over the .npy member `name` of the open zip."""
    info = zf.getinfo(name)
    if info.compress_type != zipfile.ZIP_STORED:
        raise ValueError("%s: member %s is compressed; save with compressed=False" % (path, name))
    with open(path, "rb") as f:
        f.seek(info.header_offset)
        fields = _LOCAL_HEADER.unpack(f.read(_LOCAL_HEADER.size))
        if fields[0] != _LOCAL_SIG:
            raise ValueError("%s: bad zip local header for %s" % (path, name))
        f.seek(info.header_offset + _LOCAL_HEADER.size + fields[9] + fields[10])
        version = np.lib.format.read_magic(f)
        if version == (1, 0):
            shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
        elif version in ((2, 0), (3, 0)):
            shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
        else:
            raise ValueError("%s: unsupported .npy version %s in %s" % (path, version, name))
        offset = f.tell()
    if dtype.hasobject or fortran or len(shape) != 1:
        raise ValueError("%s: %s is not a flat numeric array" % (path, name))
    if shape[0] == 0:
        return np.zeros(0, dtype)
    return np.memmap(path, dtype=dtype, mode="r", offset=offset, shape=shape)


def _small_member(zf, name):
    with zf.open(name) as f:
        return np.lib.format.read_array(f, allow_pickle=False)


def csr_from_arrays(data, indices, indptr, shape):
    """A csr_matrix holding exactly these arrays (no copies, no scans)."""
    m = sp.csr_matrix(tuple(int(x) for x in shape), dtype=data.dtype)
    m.data, m.indices, m.indptr = data, indices, indptr
    return m


def load_npz_csr(path):
    """The CSR matrix of an uncompressed scipy .npz file, memory-mapped."""
    with zipfile.ZipFile(path) as zf:
        names = set(zf.namelist())
        for need in ("data.npy", "indices.npy", "indptr.npy", "shape.npy", "format.npy"):
            if need not in names:
                raise ValueError("%s: not a scipy.sparse .npz file (no %s)" % (path, need))
        fmt = _small_member(zf, "format.npy").item()
        fmt = fmt.decode() if isinstance(fmt, bytes) else str(fmt)
        if fmt != "csr":
            raise ValueError("%s: holds a %s matrix, not csr" % (path, fmt))
        shape = _small_member(zf, "shape.npy")
        data = _npy_member(path, zf, "data.npy")
        indices = _npy_member(path, zf, "indices.npy")
        indptr = _npy_member(path, zf, "indptr.npy")
    if len(indptr) != int(shape[0]) + 1:
        raise ValueError("%s: indptr has %d entries for %d rows" % (path, len(indptr), shape[0]))
    return csr_from_arrays(data, indices, indptr, shape)


def iter_row_blocks(matrix, max_nnz):
    """Yields (row_start, row_finish, block): consecutive row blocks of
    `matrix` (a csr_matrix, e.g. from load_npz_csr) with at most `max_nnz`
    nonzeros each (a single longer row forms a block of its own).  data and
    indices of a block are views of the matrix's arrays; its indptr is a
    rebased copy."""
    if max_nnz < 1:
        raise ValueError("max_nnz must be positive")
    indptr = np.asarray(matrix.indptr, dtype=np.int64)
    rows = len(indptr) - 1
    r = 0
    while r < rows:
        base = int(indptr[r])
        # last row end within base + max_nnz (at least one row)
        e = int(np.searchsorted(indptr, base + max_nnz, side="right")) - 1
        e = min(max(e, r + 1), rows)
        ip = indptr[r:e + 1] - base
        if ip[-1] < 2 ** 31:
            ip = ip.astype(np.int32)
        end = int(indptr[e])
        block = csr_from_arrays(matrix.data[base:end], matrix.indices[base:end], ip,
                                (e - r, matrix.shape[1]))
        yield r, e, block
        r = e
