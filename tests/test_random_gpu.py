"""Seeded random join configurations through the C ABI, each against the
numpy join (whose result depends on the keys alone): key width and column
count, hash kind, HASH_BITS and chunk width, pass count and radix bits, pass
tile (standard, small, large) and build+probe block, partition layout (two-pass plans in the region
layout overflow and fall back where duplicates are dense), duplicate density,
sizes from a few rows to 2^18, explicit row ids, 1-based output, signed keys,
and the entry points (vj_join, vj_join_into with an exact or a larger capacity,
vj_join_batch).
Every case prints its parameters, so a failure names it."""
import numpy as np
import pytest

import vjoin
from joinref import np_join, sorted_pairs

pytestmark = pytest.mark.gpu

N_CASES = 96


@pytest.fixture(scope="module")
def ctx():
    c = vjoin.Context(0)
    yield c
    c.close()


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    kw = int(rng.choice([4, 8]))
    nc = int(rng.integers(1, 5))
    kind = str(rng.choice(vjoin.HASH_KINDS))
    hash_bits = int(rng.choice([32, 64]))
    chunk_bits = int(rng.choice([32, 64, 128] if nc % 2 == 0 else [32, 64]))
    passes = int(rng.choice([-1, 0, 1, 2]))
    b1 = b2 = 0
    if passes == 1:
        b1 = int(rng.integers(1, 9))
    elif passes == 2:
        b1 = int(rng.integers(1, 9))
        b2 = int(rng.integers(1, min(8, 14 - b1) + 1))
    nr = int(2 ** rng.uniform(2, 16))
    ns = int(2 ** rng.uniform(2, 18))
    distinct_per_row = float(rng.choice([0.02, 0.3, 1.0, 50.0]))  # duplicate density of the build side
    return dict(seed=seed, kw=kw, nc=nc, kind=kind, hash_bits=hash_bits, chunk_bits=chunk_bits, passes=passes,
                b1=b1, b2=b2, nr=nr, ns=ns, dens=distinct_per_row, tile=int(rng.integers(0, 3)),
                table=int(rng.integers(0, 3)), signed=bool(rng.random() < 0.2), rowids=bool(rng.random() < 0.3),
                one_based=bool(rng.random() < 0.3), entry=str(rng.choice(["join", "into", "into_big", "batch"])),
                layout=int(rng.integers(0, 3)), large_tile=bool(rng.random() < 0.25))


def _data(p, rng):
    dt = np.uint32 if p["kw"] == 4 else np.uint64
    nd = max(1, int(p["nr"] * p["dens"]))
    hi = np.iinfo(dt).max
    distinct = rng.integers(0, hi, size=(nd, p["nc"]), dtype=np.uint64, endpoint=True).astype(dt)
    r = distinct[rng.integers(0, nd, size=p["nr"])]
    s = distinct[rng.integers(0, nd, size=p["ns"])]
    miss = rng.random(p["ns"]) < 0.3
    s[miss, 0] = rng.integers(0, hi, size=int(miss.sum()), dtype=np.uint64, endpoint=True).astype(dt)
    return r, s


def _dev(cols, kw):
    import torch
    t = torch.int32 if kw == 4 else torch.int64
    return [torch.from_numpy(np.ascontiguousarray(c).view(np.int32 if kw == 4 else np.int64)).to("cuda").view(t)
            for c in cols]


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_configuration_matches_numpy_join(ctx, seed):
    import torch
    p = _case(seed)
    print("case", p, flush=True)
    rng = np.random.default_rng(seed)
    r, s = _data(p, rng)
    rr = rng.permutation(1 << 24)[:p["nr"]].astype(np.uint32) if p["rowids"] else None
    sr = rng.permutation(1 << 24)[:p["ns"]].astype(np.uint32) if p["rowids"] else None
    ob = 1 if p["one_based"] else 0
    want = np_join(r.astype(np.uint64), s.astype(np.uint64), None if rr is None else rr + ob,
                   None if sr is None else sr + ob)
    if rr is None and ob:
        want = (want[0] + np.uint32(1), want[1] + np.uint32(1))
    dr = _dev([r[:, c] for c in range(p["nc"])], p["kw"])
    ds = _dev([s[:, c] for c in range(p["nc"])], p["kw"])
    drr = None if rr is None else torch.from_numpy(rr.view(np.int32)).cuda()
    dsr = None if sr is None else torch.from_numpy(sr.view(np.int32)).cuda()
    bt = vjoin.table(dr, drr, p["signed"])
    pt = vjoin.table(ds, dsr, p["signed"])
    tile = vjoin.TILE_LARGE if p["large_tile"] else p["tile"]
    opts = vjoin.join_opts(p["kind"], p["hash_bits"], p["chunk_bits"], passes=p["passes"], bits1=p["b1"],
                           bits2=p["b2"], one_based=p["one_based"], tile=tile, table=p["table"],
                           layout=p["layout"])
    m = len(want[0])
    if p["entry"] == "join":
        a, b = ctx.join(bt, pt, opts)
        got = sorted_pairs(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))
    elif p["entry"] in ("into", "into_big"):
        cap = m if p["entry"] == "into" else m + 1000
        o_r = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
        o_s = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
        rc, cnt = ctx.join_into(bt, pt, opts, o_r, o_s, capacity=cap)
        assert rc == 0 and cnt == m, (rc, cnt, m, ctx.last_error())
        got = sorted_pairs(o_r[:cnt].cpu().numpy().view(np.uint32), o_s[:cnt].cpu().numpy().view(np.uint32))
    else:  # the join in the middle of a batch of three copies (the others with 1-pair outputs)
        caps = [1, m, 1]
        o = [(torch.empty(max(c, 1), dtype=torch.int32, device="cuda"),
              torch.empty(max(c, 1), dtype=torch.int32, device="cuda")) for c in caps]
        rc, counts = ctx.join_batch([bt, bt, bt], [pt, pt, pt], opts, [x[0] for x in o], [x[1] for x in o], caps)
        assert rc == (0 if m <= 1 else vjoin.STATUS_CODE["VJ_ERR_CAPACITY"]), (rc, ctx.last_error())
        assert list(counts) == [m, m, m]
        got = sorted_pairs(o[1][0][:m].cpu().numpy().view(np.uint32), o[1][1][:m].cpu().numpy().view(np.uint32))
    assert len(got[0]) == m, (p, len(got[0]), m)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), p
