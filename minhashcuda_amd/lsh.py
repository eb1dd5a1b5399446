"""Candidate generation from LSH band keys (mhcuda_lsh_band_keys).

Rows whose keys agree in at least one band are candidate pairs: with r
samples per band, two rows of weighted Jaccard similarity J collide in a band
with probability J^r, and in at least one of b bands with 1 - (1 - J^r)^b.
"""
import numpy as np

RESERVED = np.uint64(0xFFFFFFFFFFFFFFFF)  # key of bands holding an empty row's samples


def candidate_pairs(keys, max_bucket=64):
    """Distinct row pairs (i < j) sharing a band key.

    keys: array-like [rows][bands] of uint64 keys (a torch tensor of int64 is
    accepted).  Buckets with more than `max_bucket` rows are skipped, as are
    reserved keys.  Returns an int64 array [n][2].
    """
    if hasattr(keys, "cpu"):
        keys = keys.cpu().numpy()
    keys = np.asarray(keys).view(np.uint64)
    rows, bands = keys.shape
    found = []
    for g in range(bands):
        k = keys[:, g]
        order = np.argsort(k, kind="stable")
        ks = k[order]
        valid = ks != RESERVED
        # run lengths, to drop oversized buckets
        starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
        lengths = np.diff(np.r_[starts, rows])
        size = np.repeat(lengths, lengths)
        keep = valid & (size <= max_bucket) & (size > 1)
        # every pair inside a sorted run is (i, i + d) with equal keys
        for d in range(1, min(max_bucket, rows)):
            same = (ks[d:] == ks[:-d]) & keep[d:] & keep[:-d]
            if not same.any():
                break
            i = order[:-d][same]
            j = order[d:][same]
            found.append(np.stack([np.minimum(i, j), np.maximum(i, j)], axis=1))
    if not found:
        return np.zeros((0, 2), np.int64)
    pairs = np.concatenate(found).astype(np.int64)
    code = np.unique(pairs[:, 0] * rows + pairs[:, 1])
    return np.stack([code // rows, code % rows], axis=1)


def candidate_pairs_gpu(gen, keys, max_bucket=64, lib=None):
    """candidate_pairs on the GPU (mhcuda_lsh_candidates): `keys` is a torch
    tensor on the device, [rows][bands] int64 holding the uint64 band keys
    of mhcuda_lsh_band_keys.  Returns the same int64 array [n][2] (i < j,
    sorted) as candidate_pairs."""
    import ctypes

    import torch

    from . import LIB_PATH
    from . import capi as capi_mod
    lib = lib or capi_mod.load(LIB_PATH)
    rows, bands = keys.shape
    keys = keys.contiguous()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    count = ctypes.c_uint64(0)
    capacity = max(1024, 4 * rows * bands)
    while True:
        out = torch.empty(capacity, dtype=torch.int64, device=keys.device)
        capi_mod.check(lib.mhcuda_lsh_candidates(ctypes.c_void_p(gen), keys.data_ptr(), rows,
                                                 bands, max_bucket, out.data_ptr(), capacity,
                                                 ctypes.byref(count), stream),
                       "mhcuda_lsh_candidates")
        if count.value <= capacity:
            break
        capacity = count.value
    codes = out[:count.value].cpu().numpy().view(np.uint64)
    return np.stack([(codes >> np.uint64(32)).astype(np.int64),
                     (codes & np.uint64(0xFFFFFFFF)).astype(np.int64)], axis=1)
