"""GPU tests of resident build sides (vj_build / vj_probe_into) and of the
serving engine (vj_server_*): every result equals the same join run alone
and the oracle's pair set; budget refusal and table lifetime are errors, not
crashes."""
import numpy as np
import pytest

import vjoin
import workloads
from joinref import np_join, sorted_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vjoin.Context(0)
    yield c
    c.close()


def _pairs(r, s, n):
    return sorted_pairs(r[:n].cpu().numpy().view(np.uint32), s[:n].cpu().numpy().view(np.uint32))


def _same(got, want, what=""):
    assert len(got[0]) == len(want[0]), (what, len(got[0]), len(want[0]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), what


def _out(n):
    import torch
    return (torch.empty(max(n, 1), dtype=torch.int32, device="cuda"),
            torch.empty(max(n, 1), dtype=torch.int32, device="cuda"))


def _probe_keys(rcols, m, seed):
    """Probe keys: half drawn from the build keys, half outside them."""
    rng = np.random.default_rng(seed)
    k = rcols[0]
    hit = k[rng.integers(0, len(k), size=m // 2)]
    miss = rng.integers(0, 1 << 30, size=m - m // 2).astype(k.dtype) | np.array(1 << 30, dtype=k.dtype)
    s = np.concatenate([hit, miss])
    return [s[rng.permutation(m)]]


@pytest.mark.parametrize("log_r,passes", [(8, -1), (12, -1), (15, -1), (15, 1)])
def test_resident_build_probe_equals_join(ctx, log_r, passes):
    rcols, _ = workloads.plain_small(log_r=log_r, log_s=10)
    dr = workloads.to_device(rcols)
    opts = vjoin.join_opts(passes=passes)
    built = ctx.build(vjoin.table(dr), opts)
    assert built.nbytes() > 0
    for seed in range(3):  # several probes of one resident table
        scols = _probe_keys(rcols, 5000 + 777 * seed, seed)
        ds = workloads.to_device(scols)
        o_r, o_s = _out(len(scols[0]))
        rc, n = ctx.probe_into(built, vjoin.table(ds), opts, o_r, o_s)
        assert rc == 0, ctx.last_error()
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        _same(_pairs(o_r, o_s, n), want, (log_r, seed))
        r2, s2 = ctx.join(vjoin.table(dr), vjoin.table(ds), opts)
        _same(_pairs(r2, s2, len(r2)), want, ("join", log_r, seed))
    built.free()


def test_resident_rejects_other_hash(ctx):
    rcols, scols = workloads.plain_small(log_r=10, log_s=12)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    built = ctx.build(vjoin.table(dr), vjoin.join_opts("fnv", 64))
    o_r, o_s = _out(len(scols[0]))
    with pytest.raises(vjoin.VJError) as e:
        ctx.probe_into(built, vjoin.table(ds), vjoin.join_opts("custom_xor_hw", 64), o_r, o_s)
    assert e.value.name == "VJ_ERR_INVALID_ARGUMENT"
    built.free()


def _requests(n_fresh, seed):
    joins = workloads.batched(n_joins=n_fresh, log_min=8, log_max=14, fmin=2.0, fmax=6.0, seed=seed)
    out = []
    for rcols, scols in joins:
        out.append((rcols, scols, workloads.to_device(rcols), workloads.to_device(scols)))
    return out


def test_server_results_equal_alone_and_in_the_mix(ctx):
    """Fresh joins (some keeping their build side) submitted together, then
    re-probes of the kept tables submitted while other fresh joins run: each
    result equals the same join run alone and the oracle."""
    reqs = _requests(12, seed=21)
    srv = vjoin.Server(0, budget=0, max_batch=8)
    tickets, outs = [], []
    for i, (rcols, scols, dr, ds) in enumerate(reqs):
        o = _out(len(scols[0]))
        outs.append(o)
        tickets.append(srv.submit(vjoin.table(ds), o[0], o[1], build=vjoin.table(dr), keep_build=(i % 3 == 0)))
    resp = [srv.wait(t) for t in tickets]
    kept = {}
    for i, r in enumerate(resp):
        assert r["status"] == 0, r
        rcols, scols, dr, ds = reqs[i]
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        _same(_pairs(outs[i][0], outs[i][1], r["count"]), want, ("fresh", i))
        r2, s2 = ctx.join(vjoin.table(dr), vjoin.table(ds), vjoin.join_opts())  # alone
        _same(_pairs(r2, s2, len(r2)), want, ("alone", i))
        assert r["t_submit_ns"] <= r["t_launch_ns"] <= r["t_done_ns"]
        if i % 3 == 0:
            assert r["table"] > 0
            kept[i] = r["table"]
        else:
            assert r["table"] == 0
    # re-probes of resident tables interleaved with new fresh joins
    more = _requests(6, seed=22)
    mix = []
    for k, (i, tid) in enumerate(sorted(kept.items())):
        rcols = reqs[i][0]
        scols = _probe_keys(rcols, 3000 + 500 * k, 100 + k)
        ds = workloads.to_device(scols)
        o = _out(len(scols[0]))
        mix.append(("re", rcols, scols, ds, o, srv.submit(vjoin.table(ds), o[0], o[1], table=tid)))
        rc2, sc2, dr2, ds2 = more[k % len(more)]
        o2 = _out(len(sc2[0]))
        mix.append(("fresh", rc2, sc2, ds2, o2, srv.submit(vjoin.table(ds2), o2[0], o2[1], build=vjoin.table(dr2))))
    for kind, rcols, scols, ds, o, t in mix:
        r = srv.wait(t)
        assert r["status"] == 0, (kind, r)
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        _same(_pairs(o[0], o[1], r["count"]), want, kind)
    st = srv.stats()
    assert st["completed"] == len(reqs) + len(mix) and st["failed"] == 0
    assert st["resident_tables"] == len(kept) and st["resident_bytes"] > 0
    for tid in kept.values():
        srv.drop_table(tid)
    assert srv.stats()["resident_tables"] == 0
    o = _out(10)
    with pytest.raises(vjoin.VJError) as e:
        srv.submit(vjoin.table(reqs[0][3]), o[0], o[1], table=list(kept.values())[0])
    assert e.value.name == "VJ_ERR_NOT_FOUND"
    srv.close()


def test_server_budget_refuses_only_what_cannot_fit():
    small = _requests(4, seed=31)
    big_r, big_s = workloads.plain_small(log_r=19, log_s=23)
    dbr, dbs = workloads.to_device(big_r), workloads.to_device(big_s)
    srv = vjoin.Server(0, budget=64 << 20, max_batch=4)  # 64 MiB: the big join needs more
    o_big = _out(len(big_s[0]))
    t_big = srv.submit(vjoin.table(dbs), o_big[0], o_big[1], build=vjoin.table(dbr))
    ts, outs = [], []
    for rcols, scols, dr, ds in small:
        o = _out(len(scols[0]))
        outs.append(o)
        ts.append(srv.submit(vjoin.table(ds), o[0], o[1], build=vjoin.table(dr)))
    assert srv.wait(t_big)["status_name"] == "VJ_ERR_BUDGET"
    for (rcols, scols, dr, ds), o, t in zip(small, outs, ts):
        r = srv.wait(t)
        assert r["status"] == 0, r
        want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
        _same(_pairs(o[0], o[1], r["count"]), want)
    st = srv.stats()
    assert st["peak_bytes"] <= 64 << 20, st
    srv.close()


def test_server_capacity_reports_needed_count():
    rcols, scols, dr, ds = _requests(1, seed=41)[0]
    want = np_join(workloads.reference_rows(rcols), workloads.reference_rows(scols))
    srv = vjoin.Server(0)
    o = _out(4)
    r = srv.wait(srv.submit(vjoin.table(ds), o[0], o[1], build=vjoin.table(dr), capacity=4))
    assert r["status_name"] == "VJ_ERR_CAPACITY" and r["count"] == len(want[0])
    srv.close()
