"""Synthetic CSR inputs of the benchmark configurations (seeded, reproducible).

* ``plain_csr``  -- rows x dim, nnz/row ~ Poisson(mean), distinct sorted
  uniform columns, lognormal weights (the plain-kernel configuration:
  100,000 x 2^20, Poisson(100), lognormal(0, 1), 128 samples).
* ``batched_layout`` / ``batched_csr`` -- the uneven batch: `count` matrices,
  rows per matrix log-uniform, nnz/row Zipf(s) clipped to [1, clip],
  lognormal weights (64 matrices, dim 2^20, rows in [1k, 200k], Zipf(1.2)
  clipped to [1, 20000], lognormal(0, 1.5), 128 samples).

Columns of a row with n nonzeros are drawn stratified -- the i-th one
uniformly in [i*dim/n, (i+1)*dim/n) -- which makes them distinct and sorted
in one vectorised pass; each column is still uniform over its stratum.
"""
import numpy as np
import scipy.sparse as sp


def _stratified_columns(indptr, dim, u):
    """Columns for every position given row offsets and uniforms u in [0,1)."""
    nnz = np.diff(indptr)
    total = int(indptr[-1])
    pos = np.arange(total, dtype=np.int64) - np.repeat(indptr[:-1], nnz)
    n = np.repeat(nnz, nnz).astype(np.int64)
    lo = pos * dim // np.maximum(n, 1)
    hi = (pos + 1) * dim // np.maximum(n, 1)
    return (lo + (u * (hi - lo)).astype(np.int64)).astype(np.uint32)


def plain_csr(rows=100000, dim=1 << 20, mean=100.0, sigma=1.0, seed=1):
    rng = np.random.default_rng(seed)
    nnz = np.minimum(rng.poisson(mean, rows), dim)
    indptr = np.zeros(rows + 1, np.int64)
    np.cumsum(nnz, out=indptr[1:])
    cols = _stratified_columns(indptr, dim, rng.random(int(indptr[-1])))
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
    cols = _stratified_columns(indptr, dim, rng.random(int(indptr[-1])))
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
    nnz = torch.as_tensor(nnz_per_row, dtype=torch.int64, device=device)
    indptr = torch.zeros(len(nnz_per_row) + 1, dtype=torch.int64, device=device)
    torch.cumsum(nnz, 0, out=indptr[1:])
    total = int(indptr[-1].item())
    assert total < 2 ** 32
    cols = torch.empty(total, dtype=torch.int32, device=device)
    vals = torch.empty(total, dtype=torch.float32, device=device)
    step = 1 << 27
    row_of = torch.repeat_interleave(torch.arange(len(nnz_per_row), device=device), nnz)
    for s in range(0, total, step):
        e = min(total, s + step)
        r = row_of[s:e]
        pos = torch.arange(s, e, dtype=torch.int64, device=device) - indptr[r]
        n = nnz[r]
        lo = pos * dim // n
        hi = (pos + 1) * dim // n
        u = torch.rand(e - s, generator=g, device=device, dtype=torch.float64)
        cols[s:e] = (lo + (u * (hi - lo).to(torch.float64)).to(torch.int64)).to(torch.int32)
        if weights == "uniform100":
            u2 = torch.rand(e - s, generator=g, device=device, dtype=torch.float32)
            vals[s:e] = (1.0 - u2) * 100.0
        else:
            z = torch.randn(e - s, generator=g, device=device, dtype=torch.float32)
            vals[s:e] = torch.exp(z * sigma)
    del row_of
    return vals, cols, indptr.to(torch.int64).to(torch.int32)


def algorithmic_bytes(nnz, rows, samples):
    """HBM bytes one hashing pass must move at minimum: three FP32 parameters
    per (nonzero, sample), the nonzero's weight and column, the output pairs
    and the row offsets."""
    return 12 * nnz * samples + 8 * nnz + 8 * rows * samples + 4 * (rows + 1)
