"""GPU test of vj_join_into_noalloc, the entry point that allocates nothing:
its pair sets equal vj_join_into's (and the numpy join's) on two calls in a
row, the second smaller than the first, the context's arenas stay untouched,
and a join whose scratch exceeds the fixed array is refused before anything
is launched."""
import numpy as np
import pytest

import vjoin
import workloads
from joinref import np_join, sorted_pairs

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _out(n):
    import torch
    return (torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda"),
            torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda"))


def _pairs(o, n):
    return sorted_pairs(o[0][:n].cpu().numpy().view(np.uint32), o[1][:n].cpu().numpy().view(np.uint32))


def _same(a, b, what):
    assert len(a[0]) == len(b[0]), (what, len(a[0]), len(b[0]))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), what


CASES = [
    # (log build, log probe, key columns, key width, opts): the second is smaller than the first
    (16, 19, 1, 8, dict(passes=2, bits1=5, bits2=4)),
    (12, 15, 2, 4, dict(kind="custom_xor_hw", passes=1, bits1=6)),
]


def _inputs(log_r, log_s, ncols, kw, seed):
    rng = np.random.default_rng(seed)
    dt = np.uint32 if kw == 4 else np.uint64
    nd = 1 << (log_r - 1)
    distinct = rng.integers(0, 1 << 30, size=(nd, ncols), dtype=np.uint64).astype(dt)
    r = distinct[rng.integers(0, nd, size=1 << log_r)]  # duplicates on the build side
    s = distinct[rng.integers(0, nd, size=1 << log_s)]
    s[::3, 0] ^= dt(1 << 31)  # a third of the probe rows miss
    return [r[:, c].copy() for c in range(ncols)], [s[:, c].copy() for c in range(ncols)]


def test_noalloc_equals_join_into_twice_in_a_row():
    c = vjoin.Context(0)
    ref = vjoin.Context(0)
    try:
        assert c.memory() == 0
        for i, (lr, ls, nc, kw, kw_opts) in enumerate(CASES):
            rcols, scols = _inputs(lr, ls, nc, kw, 100 + i)
            dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
            opts = vjoin.join_opts(**kw_opts)
            want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
            m = len(want[0])
            o = _out(m)
            rc, n = c.join_into_noalloc(vjoin.table(dr), vjoin.table(ds), opts, o[0], o[1])
            assert rc == 0, c.last_error()
            assert n == m
            sm = c.summary()
            assert 0 < sm["scratch_bytes"] <= SCRATCH and sm["matches"] == m and sm["tasks"] > 0
            o2 = _out(m)
            rc2, n2 = ref.join_into(vjoin.table(dr), vjoin.table(ds), opts, o2[0], o2[1])
            assert rc2 == 0 and n2 == m
            got = _pairs(o, n)
            _same(got, _pairs(o2, n2), ("vs join_into", i))
            _same(got, want, ("vs numpy", i))
            assert c.memory() == 0  # the context's arenas were never used
        # capacity: the count is reported, the first `capacity` pairs are written
        rc, n = c.join_into_noalloc(vjoin.table(dr), vjoin.table(ds), opts, o[0], o[1], capacity=10)
        assert vjoin.STATUS[rc] == "VJ_ERR_CAPACITY" and n == m
    finally:
        c.close()
        ref.close()


def test_noalloc_refuses_joins_beyond_its_scratch():
    import torch
    c = vjoin.Context(0)
    try:
        # 2^20 + 2^23 rows of 4-byte keys over two passes need about 144 MiB > 64 MiB
        r = torch.arange(1 << 20, dtype=torch.int32, device="cuda")
        s = torch.arange(1 << 23, dtype=torch.int32, device="cuda")
        o = _out(16)
        rc, n = c.join_into_noalloc(vjoin.table([r]), vjoin.table([s]), vjoin.join_opts(passes=2), o[0], o[1])
        assert vjoin.STATUS[rc] == "VJ_ERR_UNSUPPORTED" and n == 0
        assert "bytes of scratch" in c.last_error()
        assert c.memory() == 0
        # the same context still joins a small input afterwards
        rcols, scols = _inputs(10, 13, 1, 4, 7)
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        o = _out(len(want[0]))
        rc, n = c.join_into_noalloc(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(), o[0], o[1])
        assert rc == 0, c.last_error()
        _same(_pairs(o, n), want, "small after refusal")
    finally:
        c.close()
