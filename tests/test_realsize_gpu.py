"""GPU parity at the configurations' real sizes, through the C ABI.

  * the multi-column skewed join at full size (build 2^22 rows of 4 x u32 keys,
    probe 2^25 rows, 80 % Zipf(1.0), 20 % misses), FNV-1a-64 and custom_xor_hw,
    against the oracle's join on the same hash: identical sorted pair sets;
  * a batch of 32 uneven int32 joins whose builds span 2^10..2^20 rows, so the
    planner gives 0-, 1- and 2-pass joins in one pipeline, eager and replayed
    from a hipGraph, and with autotuned plans; a forced 2-pass batch whose large
    jobs split their histogram segments over several blocks; every join against
    an independent numpy join;
  * resident 2-pass build sides re-probed by 2^18..2^22-row probes, directly and
    through the serving engine mixed with fresh joins, against the numpy join;
  * a 2^26 x 2^27 FK join of int64 keys (past the standard plan's 14 bits) with
    the wide build+probe block and with the standard one, against the oracle;
  * the 2^24 x 2^27 FK join in the region layout against the oracle, and the
    multi-column Zipf join in it (whose regions overflow: exact-layout rerun).

Pairs are compared as sorted (build_rowid << 32 | probe_rowid) keys; the sort
runs on the device (torch.sort) only to compare, the pairs come from libvj.so
and from the oracle / numpy join.
"""
import numpy as np
import pytest

import oracle
import vjoin
import workloads
from joinref import np_join

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vjoin.Context(0)
    yield c
    c.close()


def _dev_keys(r, s, n):
    """Sorted r << 32 | s of the first n pairs of two device int32 tensors."""
    import torch
    rr = r[:n].to(torch.int64) & 0xFFFFFFFF
    ss = s[:n].to(torch.int64) & 0xFFFFFFFF
    return torch.sort((rr << 32) | ss).values


def _host_keys(r, s):
    import torch
    k = (np.asarray(r, dtype=np.uint64) << np.uint64(32)) | np.asarray(s, dtype=np.uint64)
    return torch.sort(torch.from_numpy(k.view(np.int64)).cuda()).values


def _assert_same(got_keys, want_keys, what=""):
    import torch
    assert got_keys.numel() == want_keys.numel(), (what, got_keys.numel(), want_keys.numel())
    assert bool(torch.equal(got_keys, want_keys)), what


def _out(n):
    import torch
    return (torch.empty(max(n, 1), dtype=torch.int32, device="cuda"),
            torch.empty(max(n, 1), dtype=torch.int32, device="cuda"))


# --- multi-column skewed join, full size ----------------------------------------


@pytest.fixture(scope="module")
def multicol_full():
    rcols, scols = workloads.multicol_skew()  # 2^22 x 2^25
    n_hit = int(np.count_nonzero(scols[0] < np.uint32(0x80000000)))
    return {"rcols": rcols, "scols": scols, "m": 4 * n_hit,
            "r_rows": workloads.reference_rows(rcols), "s_rows": workloads.reference_rows(scols),
            "dr": workloads.to_device(rcols), "ds": workloads.to_device(scols)}


@pytest.mark.parametrize("kind,plan", [("fnv", None), ("custom_xor_hw", None), ("fnv", (5, 6, 1, 2)),
                                       ("custom_xor_hw", (5, 6, 2, 2)), ("fnv", (5, 6, 1, 2, 2))])
def test_multicol_skew_full_size_matches_oracle(ctx, multicol_full, kind, plan):
    """The planner's plan, and the autotuner's kind of plan (fewer bits, wide
    tables, either pass tile). In the region layout the Zipf keys overflow
    their regions and the join runs again with the exact layout."""
    w = multicol_full
    m = w["m"]
    assert len(w["rcols"][0]) == 1 << 22 and len(w["scols"][0]) == 1 << 25
    out_r, out_s = _out(m)
    layout = plan[4] if plan is not None and len(plan) > 4 else 0
    opts = vjoin.join_opts(kind, 64, 64, passes=2) if plan is None else \
        vjoin.join_opts(kind, 64, 64, passes=2, bits1=plan[0], bits2=plan[1], tile=plan[2], table=plan[3],
                        layout=layout)
    rc, cnt = ctx.join_into(vjoin.table(w["dr"]), vjoin.table(w["ds"]), opts, out_r, out_s)
    assert rc == 0, ctx.last_error()
    assert cnt == m
    sm = ctx.summary()
    assert sm["passes"] == 2
    if layout == vjoin.LAYOUT_REGIONS:
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_EXACT, 1)
    wr, ws = oracle.join(w["r_rows"], w["s_rows"], kind, 64, 64, passes=2, bits1=sm["bits1"], bits2=sm["bits2"],
                         cap=m)
    assert len(wr) == m
    _assert_same(_dev_keys(out_r, out_s, cnt), _host_keys(wr, ws), kind)


# --- the large FK join in the region layout: 2^24 x 2^27 -----------------------------


def test_large_fk_region_layout_matches_oracle(ctx):
    """The large FK join (build 2^24 unique int64 keys, probe 2^27 drawn from
    them, FNV-1a-64, 2 passes of 7 bits) with the region layout: no fall back,
    every probe row matched once, the oracle's pairs."""
    import torch
    rcols, scols = workloads.large_fk(log_r=24, log_s=27, seed=24)
    ns = len(scols[0])
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    out_r, out_s = _out(ns)
    opts = vjoin.join_opts("fnv", 64, 64, passes=2, bits1=7, bits2=7, layout=vjoin.LAYOUT_REGIONS)
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), opts, out_r, out_s)
    assert rc == 0, ctx.last_error()
    sm = ctx.summary()
    assert cnt == ns and (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_REGIONS, 0)
    got = _dev_keys(out_r, out_s, cnt)
    wr, ws = oracle.join(rcols[0].view(np.uint64).reshape(-1, 1), scols[0].view(np.uint64).reshape(-1, 1), "fnv",
                         64, 64, passes=2, bits1=7, bits2=7, cap=ns)
    assert len(wr) == ns
    _assert_same(got, _host_keys(wr, ws), "regions")
    del out_r, out_s, dr, ds, got
    torch.cuda.empty_cache()


# --- a build side past the 14-bit two-pass plan: 2^26 x 2^27 FK join ------------------


def test_build_of_2_26_rows_fits_wide_tables_and_matches_oracle(ctx):
    """A 2^26-row build of 8-byte keys needs 16 radix bits to reach half a
    standard (2048-row) table per partition, past the 14 bits of two passes.
    The planner then takes the wide build+probe block (8192-row tables), so 14
    bits leave one build chunk per partition; the forced standard block still
    gives the same pairs by cutting partitions into several build chunks."""
    import torch
    rcols, scols = workloads.large_fk(log_r=26, log_s=27, seed=26)
    nr, ns = len(rcols[0]), len(scols[0])
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    out_r, out_s = _out(ns)
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts("fnv", 64, 64), out_r, out_s)
    assert rc == 0, ctx.last_error()
    sm = ctx.summary()
    assert cnt == ns and (sm["passes"], sm["bits1"] + sm["bits2"]) == (2, 14)
    assert (sm["probe_threads"], sm["table_rows"]) == (1024, 8192)
    assert sm["tasks"] == 1 << 14  # one build chunk and one probe chunk (of 16384 rows) per partition
    got = _dev_keys(out_r, out_s, cnt)
    wr, ws = oracle.join(rcols[0].view(np.uint64).reshape(-1, 1), scols[0].view(np.uint64).reshape(-1, 1), "fnv",
                         64, 64, passes=2, bits1=sm["bits1"], bits2=sm["bits2"], cap=ns)
    assert len(wr) == ns
    _assert_same(got, _host_keys(wr, ws), "wide")
    del wr, ws
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts("fnv", 64, 64, table=1), out_r, out_s)
    assert rc == 0, ctx.last_error()
    assert ctx.summary()["table_rows"] == 2048 and ctx.summary()["tasks"] > 1 << 14
    _assert_same(_dev_keys(out_r, out_s, cnt), got, "standard")
    del out_r, out_s, dr, ds
    torch.cuda.empty_cache()


# --- batches of uneven joins at real sizes --------------------------------------


@pytest.fixture(scope="module")
def batch32():
    joins = workloads.batched(n_joins=30, log_min=10, log_max=20, seed=0x5EED0013)
    rng = np.random.default_rng(0x5EED0014)
    for n in (600, 1000):  # at most one build chunk: the planner joins these without partitioning
        r = rng.permutation(n).astype(np.int32)
        s = rng.integers(0, 2 * n, size=n * 8).astype(np.int32)
        joins.append(([r], [s]))
    order = rng.permutation(len(joins))
    joins = [joins[i] for i in order]
    want = [_host_keys(*np_join(workloads.reference_rows(r), workloads.reference_rows(s))) for r, s in joins]
    dev = [(workloads.to_device(r), workloads.to_device(s)) for r, s in joins]
    return joins, want, dev


def _planner_passes(n, key_bytes=4):
    """Pass count the planner picks for n build rows of 4-byte keys (partitions of at
    most half an LDS build chunk of 2048 rows; up to 8 bits per pass)."""
    need = 0 if n <= 1024 else int(np.ceil(np.log2((n + 1023) // 1024)))
    return 0 if need == 0 else (1 if need <= 8 else 2)


def _run_batch(ctx, dev, opts, use_graph):
    builds = [vjoin.table(r) for r, _ in dev]
    probes = [vjoin.table(s) for _, s in dev]
    outs = [_out(int(s[0].numel())) for _, s in dev]
    for o in outs:
        o[0].fill_(-1)
        o[1].fill_(-1)
    caps = [int(s[0].numel()) for _, s in dev]
    rc, counts = ctx.join_batch(builds, probes, opts, [o[0] for o in outs], [o[1] for o in outs], caps,
                                use_graph=use_graph)
    assert rc == 0, ctx.last_error()
    return outs, counts


def test_batch_real_sizes_mixes_pass_counts_eager_and_graph(ctx, batch32):
    joins, want, dev = batch32
    passes = {_planner_passes(len(r[0])) for r, _ in joins}
    assert passes == {0, 1, 2}, passes  # one pipeline holds all three kinds of join
    opts = vjoin.join_opts()
    for use_graph, what in ((False, "eager"), (True, "graph capture"), (True, "graph replay"),
                            (True, "graph replay 2")):
        outs, counts = _run_batch(ctx, dev, opts, use_graph)
        for i, ((r, s), o) in enumerate(zip(joins, outs)):
            _assert_same(_dev_keys(o[0], o[1], counts[i]), want[i], (what, i, len(r[0])))


def test_batch_real_sizes_with_tuned_plans(batch32):
    joins, want, dev = batch32
    c = vjoin.Context(0)
    try:
        c.set_profiling(True)
        opts = vjoin.join_opts()
        tuned = {}
        for i, (r, s) in enumerate(joins):
            cls = max(len(r[0]) - 1, 1).bit_length()
            if cls not in tuned:
                tuned[cls] = c.tune(vjoin.table(dev[i][0]), vjoin.table(dev[i][1]), opts, reps=1)[0]
        outs, counts = _run_batch(c, dev, opts, True)
        for i, o in enumerate(outs):
            _assert_same(_dev_keys(o[0], o[1], counts[i]), want[i], ("tuned", i, tuned))
    finally:
        c.close()


@pytest.mark.parametrize("bits", [(8, 3), (4, 5)])
def test_batch_forced_two_passes(ctx, batch32, bits):
    # 8 + 3 bits: the large jobs have one or two segments per pass-1 digit, so
    # each segment's histogram is split over many blocks (rows of a segment)
    joins, want, dev = batch32
    outs, counts = _run_batch(ctx, dev, vjoin.join_opts(passes=2, bits1=bits[0], bits2=bits[1]), False)
    for i, o in enumerate(outs):
        _assert_same(_dev_keys(o[0], o[1], counts[i]), want[i], ("2 passes", bits, i))


# --- resident build sides and the serving engine at real sizes ------------------


def _probe_keys(keys, m, seed):
    """m probe keys: half drawn from the build keys, half outside them."""
    rng = np.random.default_rng(seed)
    hit = keys[rng.integers(0, len(keys), size=m // 2)]
    miss = (rng.integers(0, 1 << 29, size=m - m // 2) | (1 << 30)).astype(keys.dtype)
    s = np.concatenate([hit, miss])
    return [s[rng.permutation(m)]]


@pytest.mark.parametrize("log_r,passes,bits", [(19, 2, (0, 0)), (20, 2, (6, 5)), (18, 1, (0, 0))])
def test_resident_tables_reprobed_at_real_sizes(ctx, log_r, passes, bits):
    rng = np.random.default_rng(log_r)
    rcols = [rng.permutation(1 << log_r).astype(np.int32)]
    dr = workloads.to_device(rcols)
    opts = vjoin.join_opts(passes=passes, bits1=bits[0], bits2=bits[1])
    built = ctx.build(vjoin.table(dr), opts)
    try:
        for log_s in (18, 20, 22):
            scols = _probe_keys(rcols[0], 1 << log_s, log_r * 100 + log_s)
            ds = workloads.to_device(scols)
            o_r, o_s = _out(1 << log_s)
            rc, n = ctx.probe_into(built, vjoin.table(ds), opts, o_r, o_s)
            assert rc == 0, ctx.last_error()
            want = _host_keys(*np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols)))
            _assert_same(_dev_keys(o_r, o_s, n), want, (log_r, log_s))
    finally:
        built.free()


def test_server_reprobes_at_real_sizes_in_the_mix():
    """Build sides of 2^18..2^20 rows kept resident through the server (2-pass
    tables for the larger ones), then re-probed by 2^18..2^22-row probes
    submitted together with fresh joins, so re-probes and fresh joins share
    launches: every result equals the numpy join."""
    rng = np.random.default_rng(77)
    srv = vjoin.Server(0, budget=64 << 30, opts=vjoin.join_opts(), max_batch=16)
    try:
        tables = []
        for log_r in (18, 19, 20):
            rcols = [rng.permutation(1 << log_r).astype(np.int32)]
            dr = workloads.to_device(rcols)
            probe = workloads.to_device([rcols[0][:1024].copy()])
            o = _out(1024)
            r = srv.wait(srv.submit(vjoin.table(probe), o[0], o[1], build=vjoin.table(dr), keep_build=True))
            assert r["status"] == 0 and r["table"] > 0, r
            tables.append((rcols, dr, r["table"]))
        fresh = workloads.batched(n_joins=6, log_min=10, log_max=16, seed=78)
        jobs = []
        for k, log_s in enumerate((18, 20, 22, 19, 21, 22)):
            rcols, _, tid = tables[k % len(tables)]
            scols = _probe_keys(rcols[0], 1 << log_s, 1000 + k)
            ds = workloads.to_device(scols)
            o = _out(1 << log_s)
            jobs.append((rcols, scols, o, srv.submit(vjoin.table(ds), o[0], o[1], table=tid), (ds,)))
            fr, fs = fresh[k]
            dfr, dfs = workloads.to_device(fr), workloads.to_device(fs)
            o2 = _out(len(fs[0]))
            jobs.append((fr, fs, o2, srv.submit(vjoin.table(dfs), o2[0], o2[1], build=vjoin.table(dfr)), (dfr, dfs)))
        for rcols, scols, o, t, _ in jobs:
            r = srv.wait(t)
            assert r["status"] == 0, r
            want = _host_keys(*np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols)))
            _assert_same(_dev_keys(o[0], o[1], r["count"]), want, len(scols[0]))
        st = srv.stats()
        assert st["failed"] == 0 and st["resident_tables"] == 3
    finally:
        srv.close()
