"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module. The oracle restates the reference's hashing and hash-join path on the
CPU (see vj_oracle.c for the reference citations); it is the checker, never the
thing measured as the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

HASH_KINDS = [
    "fnv", "custom_add", "custom_add_shift", "custom_add_hw", "custom_xor", "custom_xor_shift",
    "custom_xor_hw", "custom_mult", "custom_mult_shift", "custom_mult_hw", "custom_mult_xor_shift",
    "custom_n_mult_xor_shift", "custom_mult_xor_hw",
]
# Kinds whose reference result rests on undefined shifts (PTX clamp emulated): parity unpinned.
UNPINNED_KINDS = {"custom_add_shift", "custom_xor_shift", "custom_mult_shift", "custom_mult_xor_shift",
                  "custom_n_mult_xor_shift"}


class JoinCfg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("hash_bits", ctypes.c_int), ("chunk_bits", ctypes.c_int),
                ("passes", ctypes.c_int), ("bits1", ctypes.c_int), ("bits2", ctypes.c_int),
                ("load", ctypes.c_float), ("nthreads", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.or_hash_kind_from_name.argtypes = [ctypes.c_char_p]
        L.or_hash_kind_from_name.restype = i32
        L.or_hash_rows.argtypes = [vp, u64, i32, i32, i32, i32, vp, i32]
        L.or_hash_rows.restype = i32
        L.or_radix_partition.argtypes = [vp, u64, i32, i32, vp, vp, vp]
        L.or_radix_partition.restype = None
        L.or_join.argtypes = [vp, u64, vp, vp, u64, vp, i32, ctypes.POINTER(JoinCfg), vp, vp, u64]
        L.or_join.restype = ctypes.c_int64
        L.or_join_brute.argtypes = [vp, u64, vp, vp, u64, vp, i32, vp, vp, u64]
        L.or_join_brute.restype = ctypes.c_int64
        L.or_sort_pairs.argtypes = [vp, vp, u64]
        L.or_sort_pairs.restype = None
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def kind_id(name):
    k = lib().or_hash_kind_from_name(name.encode())
    if k < 0:
        raise ValueError("unknown hash kind %r" % name)
    return k


def reference_rows(cols, signed=False):
    """Row-major uint64 buffer as the reference stores keys (db_table.column_values).

    `cols` is a list of 1-D integer arrays (one per key column). 32-bit columns are
    widened to 64 bits (zero-extended, or sign-extended when `signed`)."""
    n = len(cols[0])
    rows = np.empty((n, len(cols)), dtype=np.uint64)
    for c, col in enumerate(cols):
        col = np.asarray(col)
        if col.dtype.itemsize == 4 and signed:
            rows[:, c] = col.view(np.int32).astype(np.int64).view(np.uint64)
        else:
            rows[:, c] = col.astype(np.uint64)
    return np.ascontiguousarray(rows)


def hash_rows(rows, kind, hash_bits=64, chunk_bits=64, nthreads=0):
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    n, ncols = rows.shape
    out = np.empty(n, dtype=np.uint64)
    k = kind_id(kind) if isinstance(kind, str) else int(kind)
    rc = lib().or_hash_rows(_ptr(rows), n, ncols, k, hash_bits, chunk_bits, _ptr(out), nthreads)
    if rc != 0:
        raise ValueError("oracle rejected hash arguments")
    return out


def radix_partition(hashes, shift, bits):
    hashes = np.ascontiguousarray(hashes, dtype=np.uint64)
    bins = 1 << bits
    hist = np.empty(bins, dtype=np.uint64)
    off = np.empty(bins, dtype=np.uint64)
    perm = np.empty(len(hashes), dtype=np.uint32)
    lib().or_radix_partition(_ptr(hashes), len(hashes), shift, bits, _ptr(hist), _ptr(off), _ptr(perm))
    return hist, off, perm


def join(r_rows, s_rows, kind="fnv", hash_bits=64, chunk_bits=64, passes=1, bits1=8, bits2=0, load=0.25,
         r_rowids=None, s_rowids=None, nthreads=0, cap=None):
    """Partitioned hash join (reference semantics). Returns (r_ids, s_ids) in oracle order."""
    r_rows = np.ascontiguousarray(r_rows, dtype=np.uint64)
    s_rows = np.ascontiguousarray(s_rows, dtype=np.uint64)
    ncols = r_rows.shape[1]
    cfg = JoinCfg(kind_id(kind) if isinstance(kind, str) else int(kind), hash_bits, chunk_bits, passes, bits1,
                  bits2, load, nthreads)
    rr = None if r_rowids is None else np.ascontiguousarray(r_rowids, dtype=np.uint32)
    sr = None if s_rowids is None else np.ascontiguousarray(s_rowids, dtype=np.uint32)
    if cap is None:
        cap = max(len(s_rows), 1)
    while True:
        out_r = np.empty(cap, dtype=np.uint32)
        out_s = np.empty(cap, dtype=np.uint32)
        m = lib().or_join(_ptr(r_rows), len(r_rows), _ptr(rr), _ptr(s_rows), len(s_rows), _ptr(sr), ncols,
                          ctypes.byref(cfg), _ptr(out_r), _ptr(out_s), cap)
        if m < 0:
            raise ValueError("oracle rejected join arguments")
        if m <= cap:
            return out_r[:m], out_s[:m]
        cap = int(m)


def join_brute(r_rows, s_rows, r_rowids=None, s_rowids=None):
    r_rows = np.ascontiguousarray(r_rows, dtype=np.uint64)
    s_rows = np.ascontiguousarray(s_rows, dtype=np.uint64)
    rr = None if r_rowids is None else np.ascontiguousarray(r_rowids, dtype=np.uint32)
    sr = None if s_rowids is None else np.ascontiguousarray(s_rowids, dtype=np.uint32)
    cap = max(len(s_rows), 1)
    while True:
        out_r = np.empty(cap, dtype=np.uint32)
        out_s = np.empty(cap, dtype=np.uint32)
        m = lib().or_join_brute(_ptr(r_rows), len(r_rows), _ptr(rr), _ptr(s_rows), len(s_rows), _ptr(sr),
                                r_rows.shape[1], _ptr(out_r), _ptr(out_s), cap)
        if m <= cap:
            return out_r[:m], out_s[:m]
        cap = int(m)


def sort_pairs(r, s):
    """Pairs sorted by (r, s); returns new arrays."""
    key = (np.asarray(r, dtype=np.uint64) << np.uint64(32)) | np.asarray(s, dtype=np.uint64)
    key.sort()
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
