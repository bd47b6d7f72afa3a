"""Independent equi-join in numpy (sort + searchsorted), used to pin the oracle's
join at sizes where the nested loop is too slow, and shared test helpers.

It knows nothing about hashing or partitioning: the result of an equi-join on
full key equality is fixed by the keys alone (the reference's match predicate,
src/join/join_provider.cu:422-437, only uses the hash as a pre-filter).
"""
import numpy as np


def key_ids(r_rows, s_rows):
    """Dense integer id per distinct key row over R and S together."""
    r_rows = np.asarray(r_rows, dtype=np.uint64)
    s_rows = np.asarray(s_rows, dtype=np.uint64)
    if r_rows.shape[1] == 1:
        return r_rows[:, 0], s_rows[:, 0]
    allk = np.concatenate([r_rows, s_rows])
    _, inv = np.unique(allk, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    return inv[:len(r_rows)].astype(np.uint64), inv[len(r_rows):].astype(np.uint64)


def np_join(r_rows, s_rows, r_rowids=None, s_rowids=None):
    """All (r_id, s_id) with equal key rows, sorted by (r_id, s_id)."""
    rk, sk = key_ids(r_rows, s_rows)
    order = np.argsort(rk, kind="stable")
    rks = rk[order]
    lo = np.searchsorted(rks, sk, side="left")
    hi = np.searchsorted(rks, sk, side="right")
    cnt = hi - lo
    total = int(cnt.sum())
    s_idx = np.repeat(np.arange(len(sk), dtype=np.int64), cnt)
    starts = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) if total else np.zeros(0, np.int64)
    r_pos = np.arange(total, dtype=np.int64) + starts
    r_idx = order[r_pos] if total else np.zeros(0, np.int64)
    r_ids = (np.asarray(r_rowids)[r_idx] if r_rowids is not None else r_idx).astype(np.uint32)
    s_ids = (np.asarray(s_rowids)[s_idx] if s_rowids is not None else s_idx).astype(np.uint32)
    key = (r_ids.astype(np.uint64) << np.uint64(32)) | s_ids.astype(np.uint64)
    key.sort()
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def sorted_pairs(r, s):
    key = (np.asarray(r, dtype=np.uint64) << np.uint64(32)) | np.asarray(s, dtype=np.uint64)
    key.sort()
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def fnv1a_bytes(data, bits):
    """Textbook byte-wise FNV-1a (Noll)."""
    prime, h = {32: (16777619, 2166136261), 64: (1099511628211, 14695981039346656037)}[bits]
    for b in data:
        h = ((h ^ b) * prime) % (1 << bits)
    return h
