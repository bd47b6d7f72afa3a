"""GPU parity of the region partition layout (VJ_LAYOUT_REGIONS): no histogram
read, fixed-capacity regions reserved with device atomics, and the fall back
to the exact layout when a region overflows. Every result is compared with the
numpy join (its result depends on the keys alone) through the C ABI; the
summary must name the layout the result came from."""
import numpy as np
import pytest

import vjoin
from joinref import np_join, sorted_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vjoin.Context(0)
    yield c
    c.close()


def _dev(cols, kw):
    import torch
    it = np.int32 if kw == 4 else np.int64
    return [torch.from_numpy(np.ascontiguousarray(c).view(it)).cuda() for c in cols]


def _pairs(r, s):
    return sorted_pairs(r.cpu().numpy().view(np.uint32), s.cpu().numpy().view(np.uint32))


def _check(got, want, what):
    assert len(got[0]) == len(want[0]), (what, len(got[0]), len(want[0]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), what


def _fk(rng, nr, ns, kw, nc, miss=0.2):
    """Unique build keys (nc columns of kw bytes), probe rows drawn from them, a share of misses."""
    dt = np.uint32 if kw == 4 else np.uint64
    hi = np.iinfo(dt).max
    r = rng.integers(0, hi, size=(nr, nc), dtype=np.uint64, endpoint=True).astype(dt)
    s = r[rng.integers(0, nr, size=ns)] if nr else \
        rng.integers(0, hi, size=(ns, nc), dtype=np.uint64, endpoint=True).astype(dt)
    m = rng.random(ns) < miss
    s[m, 0] = rng.integers(0, hi, size=int(m.sum()), dtype=np.uint64, endpoint=True).astype(dt)
    return r, s


def _join(ctx, r, s, kw, **kw_opts):
    dr = _dev([r[:, c] for c in range(r.shape[1])], kw)
    ds = _dev([s[:, c] for c in range(s.shape[1])], kw)
    a, b = ctx.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(**kw_opts))
    return _pairs(a, b), ctx.summary()


CASES = [  # key bytes, columns, build rows, probe rows, bits1, bits2, hash, table shape, falls back
    (8, 1, 1 << 20, 1 << 22, 7, 6, "fnv", 0, False),
    (4, 1, 70000, 300000, 4, 3, "fnv", 0, False),
    (4, 4, 1 << 18, 1 << 20, 5, 6, "custom_xor_hw", vjoin.TABLE_WIDE, False),
    (8, 2, 100000, 500000, 6, 6, "custom_add", 0, False),
    (4, 1, 3000, 5000, 3, 2, "fnv", 0, False),
    # the low 14 bits of custom_mult_xor_shift are the same for all these keys
    # (one partition holds every row): every region overflows
    (8, 1, 1 << 16, 1 << 20, 8, 6, "custom_mult_xor_shift", vjoin.TABLE_STANDARD, True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_region_layout_matches_numpy_join(ctx, case):
    kw, nc, nr, ns, b1, b2, kind, table, falls_back = CASES[case]
    print("case", CASES[case], flush=True)
    rng = np.random.default_rng(500 + case)
    r, s = _fk(rng, nr, ns, kw, nc)
    want = np_join(r.astype(np.uint64), s.astype(np.uint64))
    got, sm = _join(ctx, r, s, kw, kind=kind, hash_bits=64, chunk_bits=64, passes=2, bits1=b1, bits2=b2, table=table,
                    layout=vjoin.LAYOUT_REGIONS)
    if falls_back:
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_EXACT, 1), sm
    else:
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_REGIONS, 0), sm
        assert sm["bytes_histogram"] == 0  # no histogram read in this layout
    _check(got, want, CASES[case])


def test_region_overflow_falls_back_to_exact_layout(ctx):
    """Half the probe rows carry one key: its partition's region overflows, the
    join runs again with the exact layout and the pairs are still exact."""
    rng = np.random.default_rng(7)
    r, s = _fk(rng, 1 << 16, 1 << 20, 8, 1)
    s[rng.random(len(s)) < 0.5] = r[12345]
    want = np_join(r, s)
    got, sm = _join(ctx, r, s, 8, passes=2, bits1=4, bits2=4, layout=vjoin.LAYOUT_REGIONS)
    assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_EXACT, 1), sm
    _check(got, want, "overflow")
    # the same skew on the build side (a run of equal build keys past its region)
    r2 = r.copy()
    r2[: 1 << 14] = r2[0]
    want = np_join(r2, s[: 1 << 14])
    got, sm = _join(ctx, r2, s[: 1 << 14], 8, passes=2, bits1=4, bits2=4, layout=vjoin.LAYOUT_REGIONS)
    assert sm["layout_fallback"] == 1
    _check(got, want, "build-side overflow")


@pytest.mark.parametrize("nr,ns", [(0, 1000), (1000, 0), (0, 0), (1, 1)])
def test_region_layout_with_empty_and_tiny_sides(ctx, nr, ns):
    rng = np.random.default_rng(nr * 3 + ns)
    r, s = _fk(rng, nr, ns, 4, 1, miss=0.0)
    want = np_join(r.astype(np.uint64), s.astype(np.uint64))
    got, sm = _join(ctx, r, s, 4, passes=2, bits1=3, bits2=3, layout=vjoin.LAYOUT_REGIONS)
    assert sm["layout"] == vjoin.LAYOUT_REGIONS and sm["matches"] == len(want[0])
    _check(got, want, (nr, ns))


def test_region_layout_in_batches_and_graph_replays(ctx):
    """A batch of three 2-pass joins in the region layout, eager and replayed
    from a hipGraph (every replay checked), and repeated single joins (the
    automatic capture after the second)."""
    import torch
    rng = np.random.default_rng(11)
    joins = [_fk(rng, nr, ns, 8, 1) for nr, ns in ((1 << 15, 1 << 18), (5000, 90000), (1 << 17, 1 << 19))]
    wants = [np_join(r, s) for r, s in joins]
    builds = [vjoin.table(_dev([r[:, 0]], 8)) for r, _ in joins]
    probes = [vjoin.table(_dev([s[:, 0]], 8)) for _, s in joins]
    opts = vjoin.join_opts(passes=2, bits1=5, bits2=5, layout=vjoin.LAYOUT_REGIONS)
    caps = [len(w[0]) for w in wants]
    outs_r = [torch.empty(c, dtype=torch.int32, device="cuda") for c in caps]
    outs_s = [torch.empty(c, dtype=torch.int32, device="cuda") for c in caps]
    for graph in (False, True, True, True):
        for o in outs_r + outs_s:
            o.fill_(-1)
        rc, counts = ctx.join_batch(builds, probes, opts, outs_r, outs_s, caps, use_graph=graph)
        assert rc == 0, ctx.last_error()
        assert list(counts) == caps
        for i in range(3):
            _check(_pairs(outs_r[i], outs_s[i]), wants[i], ("batch", graph, i))
        torch.cuda.synchronize()
    out_r = torch.empty(caps[2], dtype=torch.int32, device="cuda")
    out_s = torch.empty(caps[2], dtype=torch.int32, device="cuda")
    replays = []
    for _ in range(4):
        out_r.fill_(-1)
        rc, cnt = ctx.join_into(builds[2], probes[2], opts, out_r, out_s)
        assert rc == 0 and cnt == caps[2], ctx.last_error()
        sm = ctx.summary()
        assert sm["layout"] == vjoin.LAYOUT_REGIONS
        replays.append(sm["graph_replay"])
        _check(_pairs(out_r, out_s), wants[2], "repeated join")
    assert replays[-1] == 1


def test_tuned_region_layout_and_remembered_overflow():
    """A size class tuned to the region layout uses it for planner-driven
    joins; after one overflow (skewed keys) the class keeps to the exact layout
    without another fall back, and vj_tune_clear forgets both."""
    rng = np.random.default_rng(21)
    r, s = _fk(rng, 1 << 17, 1 << 20, 8, 1)
    c = vjoin.Context(0)
    try:
        dr, ds = _dev([r[:, 0]], 8), _dev([s[:, 0]], 8)
        plan, ms = c.tune(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts(passes=2, layout=vjoin.LAYOUT_REGIONS),
                          reps=1)
        assert plan[0] == 2 and plan[5] == vjoin.LAYOUT_REGIONS and ms > 0
        got, sm = _join(c, r, s, 8)
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_REGIONS, 0)
        _check(got, np_join(r, s), "tuned")
        sk = s.copy()
        sk[rng.random(len(sk)) < 0.6] = r[99]
        want = np_join(r, sk)
        got, sm = _join(c, r, sk, 8)
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_EXACT, 1)
        _check(got, want, "first skewed")
        got, sm = _join(c, r, sk, 8)
        assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_EXACT, 0)
        _check(got, want, "second skewed")
        c.tune_clear()
        got, sm = _join(c, r, s, 8)
        assert sm["layout"] == vjoin.LAYOUT_EXACT
    finally:
        c.close()


def test_layout_out_of_range_is_rejected(ctx):
    import torch
    d = torch.arange(10, dtype=torch.int64, device="cuda")
    out = torch.empty(10, dtype=torch.int32, device="cuda")
    o = vjoin.join_opts(layout=3)
    rc, _ = ctx.join_into(vjoin.table([d]), vjoin.table([d]), o, out, out)
    assert rc == vjoin.STATUS_CODE["VJ_ERR_INVALID_ARGUMENT"]


@pytest.mark.parametrize("kw,nc,signed,one_based", [(4, 3, True, True), (8, 3, False, True), (8, 1, True, False)])
def test_region_layout_with_row_ids_one_based_and_signed_keys(ctx, kw, nc, signed, one_based):
    """Explicit row ids, 1-based output, signed keys and three key columns (keys up to 24 bytes, wide table)."""
    import torch
    rng = np.random.default_rng(kw * 10 + nc)
    r, s = _fk(rng, 70000, 400000, kw, nc)
    rr = rng.permutation(1 << 24)[:len(r)].astype(np.uint32)
    sr = rng.permutation(1 << 24)[:len(s)].astype(np.uint32)
    ob = 1 if one_based else 0
    want = np_join(r.astype(np.uint64), s.astype(np.uint64), rr + ob, sr + ob)
    dr = _dev([r[:, c] for c in range(nc)], kw)
    ds = _dev([s[:, c] for c in range(nc)], kw)
    drr = torch.from_numpy(rr.view(np.int32)).cuda()
    dsr = torch.from_numpy(sr.view(np.int32)).cuda()
    opts = vjoin.join_opts(passes=2, bits1=5, bits2=4, one_based=one_based, table=vjoin.TABLE_WIDE,
                           layout=vjoin.LAYOUT_REGIONS)
    a, b = ctx.join(vjoin.table(dr, drr, signed), vjoin.table(ds, dsr, signed), opts)
    sm = ctx.summary()
    assert (sm["layout"], sm["layout_fallback"]) == (vjoin.LAYOUT_REGIONS, 0), sm
    _check(_pairs(a, b), want, (kw, nc, signed, one_based))


def test_paths_that_keep_the_exact_layout(ctx):
    """The region layout asked for on the no-allocation entry point and on a
    resident build side: both run the exact layout and give the same pairs."""
    import torch
    rng = np.random.default_rng(33)
    r, s = _fk(rng, 1 << 15, 1 << 18, 8, 1)
    want = np_join(r, s)
    dr, ds = _dev([r[:, 0]], 8), _dev([s[:, 0]], 8)
    opts = vjoin.join_opts(passes=2, bits1=4, bits2=4, layout=vjoin.LAYOUT_REGIONS)
    m = len(want[0])
    o_r = torch.empty(m, dtype=torch.int32, device="cuda")
    o_s = torch.empty(m, dtype=torch.int32, device="cuda")
    rc, cnt = ctx.join_into_noalloc(vjoin.table(dr), vjoin.table(ds), opts, o_r, o_s)
    assert rc == 0 and cnt == m, ctx.last_error()
    assert ctx.summary()["layout"] == vjoin.LAYOUT_EXACT
    _check(_pairs(o_r, o_s), want, "no-allocation entry point")
    built = ctx.build(vjoin.table(dr), opts)
    try:
        o_r.fill_(-1)
        rc, cnt = ctx.probe_into(built, vjoin.table(ds), opts, o_r, o_s)
        assert rc == 0 and cnt == m, ctx.last_error()
        assert ctx.summary()["layout"] == vjoin.LAYOUT_EXACT
        _check(_pairs(o_r, o_s), want, "resident build side")
    finally:
        built.free()


N_RANDOM = 48


@pytest.mark.parametrize("seed", range(N_RANDOM))
def test_random_region_configurations(ctx, seed):
    """Seeded random two-pass plans in the region layout: key width and columns,
    hash kind, radix bits, pass tile, table shape, duplicate density (dense
    duplicates overflow and fall back), row ids, 1-based output, signed keys
    and the entry point; every case against the numpy join."""
    import torch
    rng = np.random.default_rng(7000 + seed)
    kw = int(rng.choice([4, 8]))
    nc = int(rng.integers(1, 5))
    kind = str(rng.choice(["fnv", "custom_xor_hw", "custom_add", "custom_mult"]))
    b1 = int(rng.integers(3, 9))
    b2 = int(rng.integers(1, min(8, 14 - b1) + 1))
    nr, ns = int(2 ** rng.uniform(0, 17)), int(2 ** rng.uniform(0, 19))
    dens = float(rng.choice([0.05, 0.5, 1.0, 20.0]))
    tile, table = int(rng.integers(0, 4)), int(rng.integers(0, 3))
    rowids, one_based, signed = bool(rng.random() < 0.3), bool(rng.random() < 0.3), bool(rng.random() < 0.2)
    entry = str(rng.choice(["join", "into", "batch"]))
    cfg = dict(seed=seed, kw=kw, nc=nc, kind=kind, b1=b1, b2=b2, nr=nr, ns=ns, dens=dens, tile=tile, table=table,
               rowids=rowids, one_based=one_based, signed=signed, entry=entry)
    print("case", cfg, flush=True)
    dt = np.uint32 if kw == 4 else np.uint64
    hi = np.iinfo(dt).max
    nd = max(1, int(nr * dens))
    distinct = rng.integers(0, hi, size=(nd, nc), dtype=np.uint64, endpoint=True).astype(dt)
    r = distinct[rng.integers(0, nd, size=nr)]
    s = distinct[rng.integers(0, nd, size=ns)]
    miss = rng.random(ns) < 0.3
    s[miss, 0] = rng.integers(0, hi, size=int(miss.sum()), dtype=np.uint64, endpoint=True).astype(dt)
    rr = rng.permutation(1 << 24)[:nr].astype(np.uint32) if rowids else None
    sr = rng.permutation(1 << 24)[:ns].astype(np.uint32) if rowids else None
    ob = 1 if one_based else 0
    want = np_join(r.astype(np.uint64), s.astype(np.uint64), None if rr is None else rr + ob,
                   None if sr is None else sr + ob)
    if rr is None and ob:
        want = (want[0] + np.uint32(1), want[1] + np.uint32(1))
    m = len(want[0])
    dr = _dev([r[:, c] for c in range(nc)], kw)
    ds = _dev([s[:, c] for c in range(nc)], kw)
    drr = None if rr is None else torch.from_numpy(rr.view(np.int32)).cuda()
    dsr = None if sr is None else torch.from_numpy(sr.view(np.int32)).cuda()
    bt, pt = vjoin.table(dr, drr, signed), vjoin.table(ds, dsr, signed)
    opts = vjoin.join_opts(kind, 64, 64, passes=2, bits1=b1, bits2=b2, one_based=one_based, tile=tile, table=table,
                           layout=vjoin.LAYOUT_REGIONS)
    if entry == "join":
        got = _pairs(*ctx.join(bt, pt, opts))
        assert ctx.summary()["layout"] in (vjoin.LAYOUT_REGIONS, vjoin.LAYOUT_EXACT)
    elif entry == "into":
        o_r = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
        o_s = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
        rc, cnt = ctx.join_into(bt, pt, opts, o_r, o_s, capacity=m)
        assert rc == 0 and cnt == m, (rc, cnt, m, ctx.last_error())
        got = _pairs(o_r[:cnt], o_s[:cnt])
    else:  # the second of two copies in one batch
        outs = [(torch.empty(max(m, 1), dtype=torch.int32, device="cuda"),
                 torch.empty(max(m, 1), dtype=torch.int32, device="cuda")) for _ in range(2)]
        rc, counts = ctx.join_batch([bt, bt], [pt, pt], opts, [o[0] for o in outs], [o[1] for o in outs], [m, m])
        assert rc == 0 and list(counts) == [m, m], (rc, counts, ctx.last_error())
        got = _pairs(outs[1][0][:m], outs[1][1][:m])
    _check(got, want, cfg)
