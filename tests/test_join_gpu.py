"""GPU parity of the join pipeline (vj_join / vj_join_into / vj_partition /
vj_join_batch) against the oracle: the sorted (build_rowid, probe_rowid) pair
sets must be identical. Full-size runs are checked by size-independent
properties (every output pair has equal keys; FK joins match each probe row
exactly once)."""
import numpy as np
import pytest

import oracle
import vjoin
import workloads
from joinref import np_join, sorted_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vjoin.Context(0)
    yield c
    c.close()


def _pairs(r, s):
    return sorted_pairs(r.cpu().numpy().view(np.uint32), s.cpu().numpy().view(np.uint32))


def _assert_same(got, want, what=""):
    assert len(got[0]) == len(want[0]), (what, len(got[0]), len(want[0]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), what


def _run(ctx, rcols, scols, rowids=(None, None), **kw):
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    rr = None if rowids[0] is None else workloads.to_device([rowids[0]])[0]
    sr = None if rowids[1] is None else workloads.to_device([rowids[1]])[0]
    key_signed = kw.pop("key_signed", False)
    opts = vjoin.join_opts(**kw)
    return ctx.join(vjoin.table(dr, rr, key_signed), vjoin.table(ds, sr, key_signed), opts)


def test_plain_small_config_matches_oracle(ctx):
    # build 2^16 u32 permutation, probe 2^20 uniform [0, 2^17), FNV-1a-32, 1 pass x 8 bits
    rcols, scols = workloads.plain_small()
    got = _pairs(*_run(ctx, rcols, scols, kind="fnv", hash_bits=32, passes=1, bits1=8))
    want = sorted_pairs(*oracle.join(workloads.reference_rows(rcols), workloads.reference_rows(scols), "fnv", 32,
                                     64, passes=1, bits1=8))
    _assert_same(got, want)
    s = ctx.summary()
    assert s["passes"] == 1 and s["bits1"] == 8 and s["matches"] == len(want[0])


@pytest.mark.parametrize("passes,b1,b2", [(0, 0, 0), (1, 4, 0), (1, 8, 0), (2, 3, 5), (2, 7, 7), (-1, 0, 0)])
@pytest.mark.parametrize("kw", [4, 8])
def test_passes_and_bits(ctx, passes, b1, b2, kw):
    rng = np.random.default_rng((passes + 1) * 31 + b1 * 7 + b2 + kw)
    nr, ns = 20000, 150000
    dt = np.uint32 if kw == 4 else np.uint64
    r = rng.integers(0, 60000, size=nr, dtype=np.uint64).astype(dt)
    s = rng.integers(0, 60000, size=ns, dtype=np.uint64).astype(dt)
    got = _pairs(*_run(ctx, [r], [s], passes=passes, bits1=b1, bits2=b2))
    _assert_same(got, np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64)),
                 (passes, b1, b2, kw))


@pytest.mark.parametrize("kind,hash_bits", [("fnv", 32), ("custom_xor_hw", 64), ("custom_xor_hw", 32),
                                            ("custom_add", 64), ("custom_mult_xor_hw", 32), ("fnv", 64)])
def test_hash_kinds(ctx, kind, hash_bits):
    rng = np.random.default_rng(hash_bits)
    r = rng.integers(0, 2**40, size=30000, dtype=np.uint64)
    s = np.concatenate([r[rng.integers(0, len(r), 60000)], rng.integers(0, 2**40, 60000, dtype=np.uint64)])
    got = _pairs(*_run(ctx, [r], [s], kind=kind, hash_bits=hash_bits, passes=2, bits1=4, bits2=4))
    want = sorted_pairs(*oracle.join(r.reshape(-1, 1), s.reshape(-1, 1), kind, hash_bits, 64, passes=2, bits1=4,
                                     bits2=4))
    _assert_same(got, want, kind)


@pytest.mark.parametrize("kw,ncols", [(4, 2), (4, 3), (4, 4), (8, 2), (8, 3), (8, 4)])
def test_multicolumn_keys_with_duplicates(ctx, kw, ncols):
    rng = np.random.default_rng(kw * 10 + ncols)
    dt = np.uint32 if kw == 4 else np.uint64
    distinct = rng.integers(0, 7, size=(3000, ncols), dtype=np.uint64).astype(dt)
    r = np.repeat(distinct, 3, axis=0)[rng.permutation(9000)]
    s = distinct[rng.integers(0, len(distinct), 40000)]
    s[::5, 0] += 11  # some misses
    got = _pairs(*_run(ctx, [r[:, c].copy() for c in range(ncols)], [s[:, c].copy() for c in range(ncols)],
                       passes=2, bits1=3, bits2=4))
    want = np_join(r.astype(np.uint64), s.astype(np.uint64))
    _assert_same(got, want, (kw, ncols))


def test_explicit_rowids_and_one_based(ctx):
    rng = np.random.default_rng(3)
    r = rng.integers(0, 5000, size=8000, dtype=np.uint64).astype(np.uint32)
    s = rng.integers(0, 5000, size=30000, dtype=np.uint64).astype(np.uint32)
    rr = rng.permutation(10**6)[:8000].astype(np.uint32)
    sr = rng.permutation(10**6)[:30000].astype(np.uint32)
    got = _pairs(*_run(ctx, [r], [s], rowids=(rr, sr), passes=1, bits1=5, one_based=True))
    want = np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64), rr + 1, sr + 1)
    _assert_same(got, want)


def test_signed_keys(ctx):
    rng = np.random.default_rng(4)
    r = rng.integers(-3000, 3000, size=5000).astype(np.int32)
    s = rng.integers(-3000, 3000, size=20000).astype(np.int32)
    got = _pairs(*_run(ctx, [r], [s], key_signed=True, passes=1, bits1=6))
    want = sorted_pairs(*oracle.join(oracle.reference_rows([r.view(np.uint32)], signed=True),
                                     oracle.reference_rows([s.view(np.uint32)], signed=True), passes=1, bits1=6))
    _assert_same(got, want)


def test_heavy_duplicates_chunk_the_build_side(ctx):
    # one key repeated far beyond an LDS build chunk: R is cut into chunks, each probed with the whole S side
    r = np.full(10000, 42, dtype=np.uint32)
    r[::3] = 7
    s = np.array([42, 7, 1, 42] * 300, dtype=np.uint32)
    got = _pairs(*_run(ctx, [r], [s], passes=1, bits1=2))
    want = np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64))
    assert len(want[0]) > 3_000_000
    _assert_same(got, want)


def test_zipf_skewed_multicolumn_sample(ctx):
    rcols, scols = workloads.multicol_skew(log_distinct=14, log_s=19)
    got = _pairs(*_run(ctx, rcols, scols, kind="custom_xor_hw", passes=2, bits1=4, bits2=4))
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    _assert_same(got, want)


@pytest.mark.parametrize("kind", ["fnv", "custom_xor_hw"])
def test_zipf_multicolumn_two_passes_mid_size(ctx, kind):
    # build 2^17 rows of 4 x u32 keys, probe 2^20 (80 % Zipf): 2048-row tiles leave few segments per
    # pass-1 digit, so the planner splits each segment over several histogram blocks, and the Zipf
    # head makes a run of heavy pass-2 chunks that the XCD groups spread out
    rcols, scols = workloads.multicol_skew(log_distinct=15, log_s=20)
    got = _pairs(*_run(ctx, rcols, scols, kind=kind, passes=2))
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    _assert_same(got, want)


# (key width, key columns, hash kind, passes, bits1, bits2, build rows): cases where a partition is larger
# than a standard table (several build chunks) but fits a wide one, and small ones
SHAPE_CASES = [(4, 1, "fnv", 1, 3, 0, 40000), (8, 1, "fnv", 2, 3, 2, 150000), (4, 4, "custom_xor_hw", 2, 2, 2, 20000),
               (8, 2, "fnv", 1, 4, 0, 50000), (8, 1, "custom_add", 1, 2, 0, 30000), (8, 3, "fnv", 0, 0, 0, 3000),
               (4, 2, "fnv", -1, 0, 0, 70000)]


@pytest.mark.parametrize("tile,table", [(1, 1), (2, 1), (1, 2), (2, 2), (3, 1), (3, 2)])
@pytest.mark.parametrize("case", SHAPE_CASES)
def test_kernel_shapes(ctx, tile, table, case):
    """Every pass tile and build+probe block shape gives the same pair set as the
    oracle (duplicate keys, misses, partitions over a standard table's rows)."""
    kw, ncols, kind, passes, b1, b2, nr = case
    rng = np.random.default_rng(nr + ncols * 7 + kw)
    dt = np.uint32 if kw == 4 else np.uint64
    hi = 2**31 if kw == 4 else 2**60
    distinct = rng.integers(0, hi, size=(nr // 2, ncols), dtype=np.uint64).astype(dt)
    r = np.repeat(distinct, 2, axis=0)[rng.permutation(2 * (nr // 2))]  # every key twice
    s = distinct[rng.integers(0, len(distinct), 3 * nr)]
    s[::4, 0] ^= dt(1 << 20)  # some misses
    rc, sc = [r[:, c].copy() for c in range(ncols)], [s[:, c].copy() for c in range(ncols)]
    got = _pairs(*_run(ctx, rc, sc, kind=kind, passes=passes, bits1=b1, bits2=b2, tile=tile, table=table))
    sm = ctx.summary()
    kb = kw * ncols
    rows = {1: (6, 4, 2), 2: (2, 2, 2), 3: (8, 6, 4)}[tile]
    assert sm["tile_rows"] == 512 * (rows[0] if kb <= 8 else (rows[1] if kb <= 16 else rows[2]))
    assert sm["probe_threads"] == (1024 if table == 2 else 512)
    assert sm["table_rows"] == ({4: 8192, 8: 8192, 16: 4096, 24: 4096}[kb] if table == 2 else (2048 if kb <= 8 else 1024))
    want = sorted_pairs(*oracle.join(workloads.reference_rows(rc), workloads.reference_rows(sc), kind, 64, 64,
                                     passes=max(passes, 0) if passes >= 0 else 0, bits1=b1, bits2=b2)) \
        if passes >= 0 else np_join(workloads.reference_rows(rc), workloads.reference_rows(sc))
    _assert_same(got, want, (case, tile, table))


def test_wide_table_heavy_duplicates(ctx):
    # one key far beyond a wide table's rows: the build side is still cut into chunks
    r = np.full(40000, 42, dtype=np.uint32)
    r[::3] = 7
    s = np.array([42, 7, 1, 42] * 100, dtype=np.uint32)
    got = _pairs(*_run(ctx, [r], [s], passes=1, bits1=2, table=2))
    want = np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64))
    assert len(want[0]) > 3_000_000
    _assert_same(got, want)


def test_shapes_in_batches_and_resident_tables(ctx):
    import torch
    joins = workloads.batched(n_joins=12, log_min=8, log_max=16, seed=21)
    builds, probes, outs_r, outs_s, caps, keep = _batch_inputs(joins)
    want = [np_join(workloads.reference_rows(r), workloads.reference_rows(s)) for r, s in joins]
    for tile, table in ((2, 2), (2, 1), (1, 2)):
        rc, counts = ctx.join_batch(builds, probes, vjoin.join_opts(tile=tile, table=table), outs_r, outs_s, caps)
        assert rc == 0, ctx.last_error()
        for i in range(len(joins)):
            _assert_same(_pairs(outs_r[i][:counts[i]], outs_s[i][:counts[i]]), want[i], (tile, table, i))
    # a resident build side probed through a wide table
    rcols, scols = joins[-1]
    opts = vjoin.join_opts(table=2)
    built = ctx.build(builds[-1], opts)
    try:
        cap = len(scols[0])
        o_r = torch.empty(cap, dtype=torch.int32, device="cuda")
        o_s = torch.empty(cap, dtype=torch.int32, device="cuda")
        rc, n = ctx.probe_into(built, probes[-1], opts, o_r, o_s)
        assert rc == 0, ctx.last_error()
        assert ctx.summary()["table_rows"] == 2048 * 4
        _assert_same(_pairs(o_r[:n], o_s[:n]), want[-1], "resident")
    finally:
        built.free()


def test_bad_shapes_are_rejected(ctx):
    import torch
    d = torch.arange(100, dtype=torch.int32, device="cuda")
    for tile, table in ((4, 0), (0, 3), (-1, 0)):
        out = torch.empty(100, dtype=torch.int32, device="cuda")
        rc, _ = ctx.join_into(vjoin.table([d]), vjoin.table([d]), vjoin.join_opts(tile=tile, table=table), out, out)
        assert rc == vjoin.STATUS_CODE["VJ_ERR_INVALID_ARGUMENT"], (tile, table, rc)


@pytest.mark.parametrize("nr,ns", [(0, 100), (100, 0), (0, 0), (1, 1), (5, 70000)])
def test_empty_and_tiny(ctx, nr, ns):
    import torch
    rng = np.random.default_rng(nr + ns)
    r = rng.integers(0, 10, size=nr, dtype=np.uint64).astype(np.uint32)
    s = rng.integers(0, 10, size=ns, dtype=np.uint64).astype(np.uint32)
    dr = torch.from_numpy(r.view(np.int32)).cuda() if nr else torch.empty(0, dtype=torch.int32, device="cuda")
    ds = torch.from_numpy(s.view(np.int32)).cuda() if ns else torch.empty(0, dtype=torch.int32, device="cuda")
    for passes in (0, 1, 2):
        got = _pairs(*ctx.join(vjoin.table([dr]), vjoin.table([ds]), vjoin.join_opts(passes=passes, bits1=2,
                                                                                      bits2=2)))
        want = np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64))
        _assert_same(got, want, (nr, ns, passes))


def test_join_into_capacity(ctx):
    import torch
    rcols, scols = workloads.plain_small(log_r=12, log_s=16)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    opts = vjoin.join_opts(passes=1, bits1=6)
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    m = len(want[0])
    small_r = torch.empty(100, dtype=torch.int32, device="cuda")
    small_s = torch.empty(100, dtype=torch.int32, device="cuda")
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), opts, small_r, small_s)
    assert vjoin.STATUS[rc] == "VJ_ERR_CAPACITY" and cnt == m
    out_r = torch.empty(m, dtype=torch.int32, device="cuda")
    out_s = torch.empty(m, dtype=torch.int32, device="cuda")
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), opts, out_r, out_s)
    assert rc == 0 and cnt == m
    _assert_same(_pairs(out_r, out_s), want)


def test_join_into_replays_graph_for_a_repeated_shape():
    """The second join_into of one launch shape captures a hipGraph and later
    ones replay it; with other work and synchronisations in between, every
    replay equals the oracle, and a new shape runs eagerly again."""
    import torch
    rcols, scols = workloads.plain_small(log_r=13, log_s=16)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    r2, s2 = workloads.plain_small(log_r=11, log_s=14, seed=5)
    dr2, ds2 = workloads.to_device(r2), workloads.to_device(s2)
    want2 = np_join(workloads.reference_rows(r2), workloads.reference_rows(s2))
    c = vjoin.Context(0)
    try:
        c.set_profiling(False)  # replay is automatic without per-stage timing
        opts = vjoin.join_opts(passes=2, bits1=4, bits2=3)
        replays = []
        for rep in range(5):
            o_r = torch.full((len(scols[0]),), -1, dtype=torch.int32, device="cuda")
            o_s = torch.full((len(scols[0]),), -1, dtype=torch.int32, device="cuda")
            rc, n = c.join_into(vjoin.table(dr), vjoin.table(ds), opts, o_r, o_s)
            assert rc == 0, c.last_error()
            sm = c.summary()
            replays.append(sm["graph_replay"])
            assert sm["ms_total"] > 0
            _assert_same(_pairs(o_r[:n], o_s[:n]), want, ("rep", rep))
            x = torch.arange(1000, device="cuda").sum()  # other work between replays
            torch.cuda.synchronize()
            assert int(x) == 499500
        assert replays[0] == 0 and all(replays[2:]), replays
        o_r = torch.empty(len(s2[0]), dtype=torch.int32, device="cuda")
        o_s = torch.empty(len(s2[0]), dtype=torch.int32, device="cuda")
        rc, n = c.join_into(vjoin.table(dr2), vjoin.table(ds2), opts, o_r, o_s)
        assert rc == 0 and c.summary()["graph_replay"] == 0
        _assert_same(_pairs(o_r[:n], o_s[:n]), want2, "other shape")
    finally:
        c.close()


def test_vj_join_grows_output_beyond_hint(ctx):
    # 1:n join whose output is 4x the probe side: vj_join reruns build+probe with an exact buffer
    rng = np.random.default_rng(9)
    base = rng.permutation(4000).astype(np.uint32)
    r = np.repeat(base, 4)
    s = base[rng.integers(0, 4000, 50000)]
    got = _run(ctx, [r], [s], passes=1, bits1=4)
    assert ctx.summary()["retries"] == 1
    want = np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64))
    _assert_same(_pairs(*got), want)


@pytest.mark.parametrize("bits,shift", [(8, 0), (5, 3), (1, 17), (6, 28), (4, 40)])
def test_partition_matches_oracle(ctx, bits, shift):
    """Digits below bit 32 come from the hash's low half (32-bit arithmetic for
    FNV-1a-64); digits reaching past it from the full 64-bit hash."""
    import torch
    rng = np.random.default_rng(bits)
    k = rng.integers(0, 2**63, size=100000, dtype=np.uint64)
    dk = workloads.to_device([k])
    out_k = torch.empty_like(dk[0])
    out_id = torch.empty(len(k), dtype=torch.int32, device="cuda")
    hist = torch.empty(1 << bits, dtype=torch.int32, device="cuda")
    off = torch.empty(1 << bits, dtype=torch.int32, device="cuda")
    hopts = vjoin.hash_opts("fnv", 64, 64)
    rc = ctx.partition(vjoin.table(dk), hopts, shift, bits, vjoin.table([out_k], out_id), hist, off)
    assert rc == 0, ctx.last_error()
    torch.cuda.synchronize()
    h = oracle.hash_rows(k.reshape(-1, 1), "fnv", 64, 64)
    whist, woff, wperm = oracle.radix_partition(h, shift, bits)
    assert np.array_equal(hist.cpu().numpy().astype(np.uint64), whist)
    assert np.array_equal(off.cpu().numpy().astype(np.uint64), woff)
    ids = out_id.cpu().numpy().view(np.uint32)
    keys = out_k.cpu().numpy().view(np.uint64)
    assert np.array_equal(keys, k[ids])  # rows moved with their ids
    for b in range(1 << bits):  # same multiset of rows per digit
        seg = slice(int(woff[b]), int(woff[b] + whist[b]))
        assert np.array_equal(np.sort(ids[seg]), np.sort(wperm[seg]))


def test_batch_of_uneven_joins(ctx):
    import torch
    joins = workloads.batched(n_joins=24, log_min=8, log_max=15)
    builds, probes, outs_r, outs_s, caps, keep = [], [], [], [], [], []
    for rcols, scols in joins:
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        keep += [dr, ds]
        builds.append(vjoin.table(dr))
        probes.append(vjoin.table(ds))
        cap = len(scols[0])
        outs_r.append(torch.empty(cap, dtype=torch.int32, device="cuda"))
        outs_s.append(torch.empty(cap, dtype=torch.int32, device="cuda"))
        caps.append(cap)
    rc, counts = ctx.join_batch(builds, probes, vjoin.join_opts(), outs_r, outs_s, caps)
    assert rc == 0, ctx.last_error()
    for i, (rcols, scols) in enumerate(joins):
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        got = _pairs(outs_r[i][:counts[i]], outs_s[i][:counts[i]])
        _assert_same(got, want, i)


def _batch_inputs(joins):
    import torch
    builds, probes, outs_r, outs_s, caps, keep = [], [], [], [], [], []
    for rcols, scols in joins:
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        keep += [dr, ds]
        builds.append(vjoin.table(dr))
        probes.append(vjoin.table(ds))
        cap = len(scols[0])
        outs_r.append(torch.empty(max(cap, 1), dtype=torch.int32, device="cuda"))
        outs_s.append(torch.empty(max(cap, 1), dtype=torch.int32, device="cuda"))
        caps.append(cap)
    return builds, probes, outs_r, outs_s, caps, keep


def test_batch_graph_replays_match_eager(ctx):
    """The captured pipeline, replayed several times with other work and
    synchronisations in between, gives the eager result every time; a batch of
    another shape is captured anew, and the first shape still replays right."""
    import torch
    joins_a = workloads.batched(n_joins=16, log_min=8, log_max=14, seed=11)
    joins_b = workloads.batched(n_joins=9, log_min=9, log_max=13, seed=12)
    ia, ib = _batch_inputs(joins_a), _batch_inputs(joins_b)
    opts = vjoin.join_opts()
    want_a = [np_join(workloads.reference_rows(r), workloads.reference_rows(s)) for r, s in joins_a]
    want_b = [np_join(workloads.reference_rows(r), workloads.reference_rows(s)) for r, s in joins_b]

    def run(inp, want, use_graph, what):
        builds, probes, outs_r, outs_s, caps, _ = inp
        for o in outs_r + outs_s:
            o.fill_(-1)
        rc, counts = ctx.join_batch(builds, probes, opts, outs_r, outs_s, caps, use_graph=use_graph)
        assert rc == 0, ctx.last_error()
        for i in range(len(builds)):
            _assert_same(_pairs(outs_r[i][:counts[i]], outs_s[i][:counts[i]]), want[i], (what, i))

    run(ia, want_a, False, "eager a")
    rcols, scols = workloads.plain_small(log_r=10, log_s=14)
    for rep in range(3):
        run(ia, want_a, True, "graph a %d" % rep)
        r, s = _run(ctx, rcols, scols)  # eager work between replays
        torch.cuda.synchronize()
        assert len(r) > 0
    run(ib, want_b, True, "graph b")
    run(ib, want_b, True, "graph b again")
    run(ia, want_a, True, "graph a after b")
    run(ia, want_a, False, "eager a after graphs")


def test_tune_keeps_fastest_plan_and_results():
    """vj_tune picks a plan among the candidates around the planner's choice,
    later planner-driven joins of the size class use it, and results match."""
    rcols, scols = workloads.plain_small(log_r=14, log_s=17)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    c = vjoin.Context(0)
    try:
        c.set_profiling(True)
        (P, b1, b2, tile, table, layout), ms = c.tune(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(), reps=2)
        assert ms > 0 and P in (0, 1, 2)
        assert (P == 0 and b1 == 0) or (P >= 1 and 1 <= b1 <= 8) and (P < 2 or 1 <= b2 <= 8)
        assert tile in (vjoin.TILE_STANDARD, vjoin.TILE_SMALL, vjoin.TILE_LARGE)
        assert table in (vjoin.TABLE_STANDARD, vjoin.TABLE_WIDE)
        assert layout == vjoin.LAYOUT_EXACT or (P == 2 and layout == vjoin.LAYOUT_REGIONS)
        r, s_ = c.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts())
        sm = c.summary()
        assert (sm["passes"], sm["bits1"], sm["bits2"] if P == 2 else 0) == (P, b1, b2 if P == 2 else 0)
        assert sm["layout"] == layout and sm["layout_fallback"] == 0
        # the tuned kernel shapes are used too (4-byte keys: tiles of 3072 or 1024 rows, tables of 2048 or 8192)
        assert sm["tile_rows"] == {vjoin.TILE_STANDARD: 3072, vjoin.TILE_SMALL: 1024, vjoin.TILE_LARGE: 4096}[tile]
        assert (sm["probe_threads"], sm["table_rows"]) == ((512, 2048) if table == vjoin.TABLE_STANDARD else (1024, 8192))
        _assert_same(_pairs(r, s_), want, "tuned")
        c.tune_clear()
        r, s_ = c.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts())
        _assert_same(_pairs(r, s_), want, "untuned")
    finally:
        c.close()


def test_budget_admission(ctx):
    rcols, scols = workloads.plain_small(log_r=14, log_s=18)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    c2 = vjoin.Context(0)
    c2.set_budget(1 << 16)
    with pytest.raises(vjoin.VJError) as e:
        c2.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(passes=1, bits1=8))
    assert e.value.name == "VJ_ERR_BUDGET"
    c2.set_budget(0)
    r, s = c2.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(passes=1, bits1=8))
    assert len(r) > 0
    c2.close()


def test_budget_holds_at_the_limit():
    """Under a budget the arenas take exactly what admission counts: a budget of
    exactly the join's scratch admits it and the context then holds no more
    than the budget; one byte less refuses it. vj_join's regrown output is
    admitted too: a budget that covers the first output but not the regrown
    one refuses the join instead of exceeding it."""
    import torch
    rcols, scols = workloads.plain_small(log_r=14, log_s=18)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    opts = vjoin.join_opts(passes=2, bits1=4, bits2=3)
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    o_r = torch.empty(len(scols[0]), dtype=torch.int32, device="cuda")
    o_s = torch.empty(len(scols[0]), dtype=torch.int32, device="cuda")
    probe = vjoin.Context(0)
    probe.set_budget(1 << 40)  # a budget: exact arena sizes
    rc, n = probe.join_into(vjoin.table(dr), vjoin.table(ds), opts, o_r, o_s)
    assert rc == 0, probe.last_error()
    need = probe.memory()
    probe.close()
    c = vjoin.Context(0)
    try:
        c.set_budget(need)
        rc, n = c.join_into(vjoin.table(dr), vjoin.table(ds), opts, o_r, o_s)
        assert rc == 0, c.last_error()
        assert c.memory() <= need
        _assert_same(_pairs(o_r[:n], o_s[:n]), want, "at the limit")
        c2 = vjoin.Context(0)
        c2.set_budget(need - 1)
        rc, _ = c2.join_into(vjoin.table(dr), vjoin.table(ds), opts, o_r, o_s)
        assert vjoin.STATUS[rc] == "VJ_ERR_BUDGET"
        assert c2.memory() == 0
        c2.close()
    finally:
        c.close()
    # 1:4 join: the output is 4x the probe side, so vj_join regrows it
    rng = np.random.default_rng(9)
    base = rng.permutation(4000).astype(np.uint32)
    r = np.repeat(base, 4)
    s = base[rng.integers(0, 4000, 50000)]
    dr2, ds2 = workloads.to_device([r]), workloads.to_device([s])
    o1 = vjoin.join_opts(passes=1, bits1=4)
    probe = vjoin.Context(0)
    probe.set_budget(1 << 40)
    probe.join(vjoin.table(dr2), vjoin.table(ds2), o1)
    scratch = probe.memory()
    probe.close()
    c3 = vjoin.Context(0)
    try:
        c3.set_budget(scratch + len(s) * 8 * 2)  # room for the first output, not for the 4x one
        with pytest.raises(vjoin.VJError) as e:
            c3.join(vjoin.table(dr2), vjoin.table(ds2), o1)
        assert e.value.name == "VJ_ERR_BUDGET"
        c3.set_budget(scratch + len(s) * 8 * 4)  # room for the regrown output
        got = c3.join(vjoin.table(dr2), vjoin.table(ds2), o1)
        assert c3.summary()["retries"] == 1 and c3.memory() <= scratch
        _assert_same(_pairs(*got), np_join(r.reshape(-1, 1).astype(np.uint64), s.reshape(-1, 1).astype(np.uint64)))
    finally:
        c3.close()


def test_large_fk_join_full_size(ctx):
    """2^24 unique int64 build keys, 2^27 FK probe keys, 2 passes: every probe row
    matches exactly one build row with an equal key."""
    import torch
    rcols, scols = workloads.large_fk()
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    n = len(scols[0])
    out_r = torch.empty(n, dtype=torch.int32, device="cuda")
    out_s = torch.empty(n, dtype=torch.int32, device="cuda")
    rc, cnt = ctx.join_into(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(passes=2), out_r, out_s)
    assert rc == 0, ctx.last_error()
    assert cnt == n
    torch.cuda.synchronize()
    sid = out_s.long()
    assert torch.equal(torch.sort(sid).values, torch.arange(n, device="cuda"))
    assert torch.equal(dr[0][out_r.long()], ds[0][sid])
