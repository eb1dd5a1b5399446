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
