"""Synthetic CSR inputs of the benchmark configurations (seeded, reproducible).

* ``plain_csr``  -- rows x dim, nnz/row ~ Poisson(mean), distinct sorted
  uniform columns, lognormal weights (the plain-kernel configuration:
  100,000 x 2^20, Poisson(100), lognormal(0, 1), 128 samples).
* ``batched_layout`` / ``batched_csr`` -- the uneven batch: `count` matrices,
  rows per matrix log-uniform, nnz/row Zipf(s) clipped to [1, clip],
  lognormal weights (64 matrices, dim 2^20, rows in [1k, 200k], Zipf(1.2)
  clipped to [1, 20000], lognormal(0, 1.5), 128 samples).

Columns of a row with n nonzeros are n distinct sorted values, close to a
uniform n-subset of [0, dim): n draws x uniform in [0, dim - n] are sorted
within the row and the i-th gets i added (which makes them distinct).  Rows of
equal length therefore share no column pattern: a generator that placed the
i-th column of every row in the same stratum [i*dim/n, (i+1)*dim/n) would make
equal-length rows gather the same parameter blocks at the same time, a cache
reuse that real inputs do not have.
"""
import numpy as np
import scipy.sparse as sp


def _uniform_columns(indptr, dim, u):
    """Columns for every position given row offsets and uniforms u in [0,1)
    (a row with more nonzeros than dim gets sorted columns with repeats)."""
    nnz = np.diff(indptr)
    total = int(indptr[-1])
    row = np.repeat(np.arange(len(nnz), dtype=np.int64), nnz)
    n = np.repeat(nnz, nnz).astype(np.int64)
    distinct = n <= dim
    x = (u * np.where(distinct, dim - n + 1, dim)).astype(np.int64)
    key = np.sort(row * (dim + 1) + x)  # sorted within each row, rows stay in place
    pos = np.arange(total, dtype=np.int64) - np.repeat(indptr[:-1], nnz)
    return (key - row * (dim + 1) + np.where(distinct, pos, 0)).astype(np.uint32)


def plain_csr(rows=100000, dim=1 << 20, mean=100.0, sigma=1.0, seed=1):
    rng = np.random.default_rng(seed)
    nnz = np.minimum(rng.poisson(mean, rows), dim)
    indptr = np.zeros(rows + 1, np.int64)
    np.cumsum(nnz, out=indptr[1:])
    cols = _uniform_columns(indptr, dim, rng.random(int(indptr[-1])))
    w = rng.lognormal(0.0, sigma, int(indptr[-1])).astype(np.float32)
    return sp.csr_matrix((w, cols, indptr.astype(np.uint32)), shape=(rows, dim))


def batched_layout(count=64, min_rows=1000, max_rows=200000, zipf_s=1.2, clip=20000,
                   seed=1):
    """Row counts and nnz/row of the uneven batch: a list of int64 arrays."""
    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(np.log(min_rows), np.log(max_rows), count)).astype(np.int64)
    return [np.clip(rng.zipf(zipf_s, n), 1, clip).astype(np.int64) for n in lengths]


def batched_csr(nnz_per_row, dim=1 << 20, sigma=1.5, seed=1):
    """One matrix of the uneven batch on the host (for parity-sized cases)."""
    rng = np.random.default_rng(seed)
    indptr = np.zeros(len(nnz_per_row) + 1, np.int64)
    np.cumsum(nnz_per_row, out=indptr[1:])
    assert indptr[-1] < 2 ** 32
    cols = _uniform_columns(indptr, dim, rng.random(int(indptr[-1])))
    w = rng.lognormal(0.0, sigma, int(indptr[-1])).astype(np.float32)
    return sp.csr_matrix((w, cols, indptr.astype(np.uint32)), shape=(len(nnz_per_row), dim))


def batched_csr_torch(nnz_per_row, dim=1 << 20, sigma=1.5, seed=1, device="cuda",
                      weights="lognormal"):
    """One CSR matrix generated directly in device memory with torch (the
    benchmark batches hold billions of nonzeros, too many to build on the
    host).  weights: "lognormal" (lognormal(0, sigma)) or "uniform100"
    (uniform in (0, 100]).  Returns (weights f32, cols i32 holding u32 bits,
    indptr i32 holding u32 bits)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    nnz_host = np.asarray(nnz_per_row, dtype=np.int64)
    nnz = torch.as_tensor(nnz_host, device=device)
    indptr = torch.zeros(len(nnz_host) + 1, dtype=torch.int64, device=device)
    torch.cumsum(nnz, 0, out=indptr[1:])
    ip_host = np.zeros(len(nnz_host) + 1, np.int64)
    np.cumsum(nnz_host, out=ip_host[1:])
    total = int(ip_host[-1])
    assert total < 2 ** 32
    cols = torch.empty(total, dtype=torch.int32, device=device)
    vals = torch.empty(total, dtype=torch.float32, device=device)
    step = 1 << 27
    r0 = 0
    while r0 < len(nnz_host):  # whole rows, about `step` nonzeros at a time
        r1 = int(np.searchsorted(ip_host, ip_host[r0] + step, side="right")) - 1
        r1 = min(max(r1, r0 + 1), len(nnz_host))
        s, e = int(ip_host[r0]), int(ip_host[r1])
        r0_, r1_ = r0, r1
        r0 = r1
        if e == s:
            continue
        row = torch.repeat_interleave(torch.arange(r1_ - r0_, device=device), nnz[r0_:r1_])
        n = nnz[r0_:r1_][row]
        distinct = n <= dim  # longer rows get sorted columns with repeats
        u = torch.rand(e - s, generator=g, device=device, dtype=torch.float64)
        x = (u * torch.where(distinct, dim - n + 1, dim).to(torch.float64)).to(torch.int64)
        key = torch.sort(row * (dim + 1) + x).values
        pos = torch.arange(s, e, dtype=torch.int64, device=device) - indptr[r0_:r1_][row]
        cols[s:e] = (key - row * (dim + 1) + torch.where(distinct, pos, 0)).to(torch.int32)
        del row, n, u, x, key, pos, distinct
        if weights == "uniform100":
            u2 = torch.rand(e - s, generator=g, device=device, dtype=torch.float32)
            vals[s:e] = (1.0 - u2) * 100.0
        else:
            z = torch.randn(e - s, generator=g, device=device, dtype=torch.float32)
            vals[s:e] = torch.exp(z * sigma)
    return vals, cols, indptr.to(torch.int64).to(torch.int32)


def algorithmic_bytes(nnz, rows, samples):
    """HBM bytes one hashing pass must move at minimum: three FP32 parameters
    per (nonzero, sample), the nonzero's weight and column, the output pairs
    and the row offsets."""
    return 12 * nnz * samples + 8 * nnz + 8 * rows * samples + 4 * (rows + 1)
