#!/usr/bin/env python3
"""Benchmark of the single-GPU hash join (one MI355X).

A step is one complete join of the large FK workload: build 2^24 unique int64
keys, probe 2^27 int64 keys drawn from the build keys (100 % match), FNV-1a-64,
2 radix passes (bits autotuned), all (build_rowid, probe_rowid) pairs written
to HBM. Inputs are resident in HBM before the timed region; the output buffers
are allocated before it too (allocation is timed separately).

Prints one JSON line. `value` is probe tuples per second of the whole join.
`roofline` is the dominant kernel's algorithmic bytes per launch over its
average duration, from HIP events the library records on the join's stream
around each stage. Further legs, in the same line:
  multicol_skew  the multi-column Zipf join (2^22 x 2^25, 4 x uint32 keys)
  hash           the standalone hash kernel over HASH_BITS x chunk widths
  bigbuild       a 2^26 x 2^27 FK join, wide vs. standard build+probe block
  batched        256 uneven joins, one batched pipeline vs. sequential launches
  serving        a Poisson request stream, served vs. one join per launch
  cpu_baseline   the same algorithm (the OpenMP oracle) on the host cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--legs fk,multicol,hash,bigbuild,batched,serving,cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("vectorized-join_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
KEY_BYTES = 8
ROWID_BYTES = 4
PAIR_BYTES = 8


def algorithmic_bytes(nr, ns, m, passes):
    """B = (2P + 1) * (|R| w + |S| w) + M * 2 * 4 with w = 8 B key + 4 B row id."""
    w = KEY_BYTES + ROWID_BYTES
    return (2 * passes + 1) * (nr + ns) * w + m * PAIR_BYTES


def stage_bytes(nr, ns, m, regions=False):
    """Bytes each stage must move (its algorithmic traffic; the region layout has no histogram read)."""
    n = nr + ns
    return {
        "histogram": 0 if regions else n * KEY_BYTES,                 # read keys
        "pass1": n * (KEY_BYTES + KEY_BYTES + ROWID_BYTES),            # read keys, write key + row id
        "pass2": n * 2 * (KEY_BYTES + ROWID_BYTES),                   # read + write key + row id
        "build_probe": n * (KEY_BYTES + ROWID_BYTES) + m * PAIR_BYTES,  # read tuples, write pairs
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(r_keys, s_keys, passes, bits1, bits2, runs=5, one_thread_log_shrink=3):
    """The same algorithm (oracle/vj_oracle.c: word-wise FNV-1a-64, LSB radix
    digits with the GPU run's passes and bits, per-partition hash tables) on the
    host cores: OpenMP on all of this process's cores with OMP_PROC_BIND=close,
    the median of `runs` timed joins after one warm-up join, on the full large
    FK join (the same seeded keys as the GPU steps; generation not timed); plus
    a 1-thread figure on a 1/2^shrink sample of it (median of 3 after a warm-up)."""
    import numpy as np
    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the oracle's OpenMP runtime starts
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    r = r_keys.view(np.uint64).reshape(-1, 1)
    s = s_keys.view(np.uint64).reshape(-1, 1)

    def timed(rr, ss, nthreads, n_runs):
        ts = []
        for k in range(n_runs + 1):  # the first run is the warm-up
            t0 = time.perf_counter()
            out_r, _ = oracle.join(rr, ss, "fnv", 64, 64, passes=passes, bits1=bits1, bits2=bits2, load=0.25,
                                   nthreads=nthreads, cap=len(ss))
            dt = time.perf_counter() - t0
            assert len(out_r) == len(ss)  # FK join: every probe row matches once
            if k:
                ts.append(dt)
        return float(np.median(ts))

    t_all = timed(r, s, threads, runs)
    shrink = 1 << one_thread_log_shrink
    r1 = r[: len(r) // shrink]
    s1 = s[: len(s) // shrink]
    s1 = r1[np.random.default_rng(1).integers(0, len(r1), size=len(s1))]  # FK keys of the smaller build side
    t_one = timed(r1, s1, 1, 3)
    return {"value": len(s) / t_all, "unit": "probe tuples/s", "cores": threads, "kind": "port",
            "sample": "full large FK join (build %d, probe %d int64 keys), FNV-1a-64, %d passes, bits %d+%d, "
                      "OpenMP oracle oracle/vj_oracle.c, median of %d runs after a warm-up: %.3f s" % (
                          len(r), len(s), passes, bits1, bits2, runs, t_all),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "one_thread": {"value": len(s1) / t_one, "unit": "probe tuples/s", "cores": 1,
                           "sample": "build %d, probe %d (1/%d of the join), median of 3 after a warm-up: %.3f s" % (
                               len(r1), len(s1), shrink, t_one)}}


# Stage -> kernel (the HIP events around a stage bracket exactly this kernel, except
# the histogram stage, which also holds the one-block k_prepare).
STAGE_KERNEL = {"histogram": "k_hist", "pass1": "k_pass", "pass2": "k_pass", "build_probe": "k_probe"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "large_fk_traffic.txt")
CALIB_BYTES = 1 << 30  # plain copy used to scale the FETCH_SIZE / WRITE_SIZE counters


def _pmc_rows(path, counter):
    """(dispatch id, kernel name, counter value in bytes) from a rocprofv3 --pmc
    result: its sqlite database (counters_collection view) or counter_collection.csv."""
    rows = []
    if path.endswith(".db"):
        import sqlite3
        db = sqlite3.connect(path)
        for d, name, c, v in db.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection"):
            if c == counter:
                rows.append((int(d), name, float(v) * 1024.0))
        db.close()
    else:
        import csv
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"]) * 1024.0))
    rows.sort()
    return rows


def _stage_of_dispatches(rows):
    """Maps each join-kernel dispatch to its stage; the two k_pass dispatches of a
    join come in order (pass 1, then pass 2)."""
    out, npass = [], 0
    for _, name, val in rows:
        if "k_hist" in name:
            out.append(("histogram", val))
        elif "k_pass" in name:
            out.append(("pass1" if npass % 2 == 0 else "pass2", val))
            npass += 1
        elif "k_probe" in name and "reset" not in name:
            out.append(("build_probe", val))
    return out


def traffic_from_pmc(fetch_csv, write_csv, workload, leg_args="--legs fk --steps 1 --warmup 1 --no-tune"):
    """HBM bytes per launch of each stage's kernel from two rocprofv3 --pmc runs,
    each counter scaled by what it reads for a plain 1 GiB device copy."""
    import numpy as np
    lines = ["# HBM traffic per launch, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch),",
             "# each scaled by counter(1 GiB device copy) / 1 GiB. Workload: " + workload,
             "# made by: rocprofv3 --pmc FETCH_SIZE -d F -o run -- python3 bench.py %s --calib-copy" % leg_args,
             "#          rocprofv3 --pmc WRITE_SIZE -d W -o run -- python3 bench.py %s --calib-copy" % leg_args,
             "#          python3 bench.py --traffic-from F/run_results.db W/run_results.db"]
    factors, per = {}, {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = _pmc_rows(path, counter)
        # the calibration copies are the largest copy dispatches of the run
        copies = [v for _, n, v in rows if "copy" in n.lower()]
        big = max(copies)
        calib = float(np.median([v for v in copies if v >= 0.98 * big]))
        factors[counter] = CALIB_BYTES / calib
        for st, v in _stage_of_dispatches(rows):
            per.setdefault((st, counter), []).append(v)
        lines.append("calib %s raw_bytes_for_1GiB_copy=%.0f scale=%.4f" % (counter, calib, factors[counter]))
    for st in ("histogram", "pass1", "pass2", "build_probe"):
        if (st, "FETCH_SIZE") not in per:
            continue
        fb = float(np.median(per[(st, "FETCH_SIZE")])) * factors["FETCH_SIZE"]
        wb = float(np.median(per[(st, "WRITE_SIZE")])) * factors["WRITE_SIZE"]
        lines.append("stage=%s kernel=%s fetch_bytes=%.0f write_bytes=%.0f traffic_bytes=%.0f dispatches=%d" %
                     (st, STAGE_KERNEL[st], fb, wb, fb + wb, len(per[(st, "FETCH_SIZE")])))
    return "\n".join(lines) + "\n"


def counters_from_pmc(db_paths, rows, command, workload):
    """Per kernel (the two k_pass launches of a join split into pass 1 and pass
    2), the mean per dispatch of every counter in the given rocprofv3 --pmc
    databases, and per tuple: instruction counts x 64 / tuples per dispatch
    (instructions per lane and row), and LDS bank-conflict cycles per LDS
    instruction cycle."""
    import collections
    import re
    import sqlite3
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in db_paths:
        db = sqlite3.connect(path)
        got = db.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection").fetchall()
        db.close()
        names = {}
        for d, n, _, _ in got:
            names[d] = n
        label, npass = {}, collections.Counter()
        for d in sorted(names):
            m = re.search(r"vj::(?:\(anonymous namespace\)::)?(k_[a-z0-9_]+(<[^>]*>)?)", names[d])
            short = m.group(1) if m else names[d][:40]
            if short.startswith("k_pass"):
                npass[short] += 1
                short += " pass%d" % (2 - npass[short] % 2)
            label[d] = short
        for d, _, c, v in got:
            if label[d].startswith("k_"):
                per[label[d]][c].append(float(v))
    lines = ["# SQ counters per dispatch (mean) and per tuple. Workload: " + workload,
             "# tuples per dispatch: %d (build + probe rows)" % rows,
             "# made by: " + command,
             "# per-row columns: instructions x 64 lanes / tuples; bank = SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS"]
    for k in sorted(per):
        c = {n: sum(v) / len(v) for n, v in per[k].items()}
        parts = ["%s=%.4g" % (n, c[n]) for n in sorted(c)]
        for n in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR"):
            if n in c:
                parts.append("%s_per_row=%.2f" % (n[9:], c[n] * 64.0 / rows))
        if c.get("SQ_ACTIVE_INST_LDS"):
            parts.append("bank=%.2f" % (c.get("SQ_LDS_BANK_CONFLICT", 0.0) / c["SQ_ACTIVE_INST_LDS"]))
        lines.append("%-28s %s" % (k, " ".join(parts)))
    return "\n".join(lines) + "\n"


def trace_summary(db_path, command, workload):
    """Markdown summary of a rocprofv3 --kernel-trace database: per kernel, the
    dispatch count, average / min / max duration and its resources."""
    import sqlite3
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, count(*), avg(duration), min(duration), max(duration), sum(duration), "
                      "max(vgpr_count), max(sgpr_count), max(lds_size), max(scratch_size), max(grid_x), "
                      "max(workgroup_x) from kernels group by name order by sum(duration) desc").fetchall()
    db.close()
    lines = ["# Kernel trace: " + workload, "", "Made by:", "", "```", command, "python3 bench.py --trace-summary "
             "<that run's run_results.db> OUT.md --trace-command '...' --trace-workload '...'", "```", "",
             "| kernel | calls | avg us | min us | max us | total us | VGPRs | SGPRs | LDS B | scratch B/lane | "
             "grid | block |", "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, n, avg, mn, mx, tot, vg, sg, lds, scr, gx, wx in rows:
        short = name.replace("(anonymous namespace)::", "")
        if len(short) > 70:
            short = short[:67] + "..."
        lines.append("| `%s` | %d | %.1f | %.1f | %.1f | %.1f | %d | %d | %d | %d | %d | %d |" % (
            short, n, avg / 1e3, mn / 1e3, mx / 1e3, tot / 1e3, vg, sg, lds, scr, gx, wx))
    return "\n".join(lines) + "\n"


def read_traffic(stage, workload):
    """Traffic per launch for `stage` from the committed profile summary, if it
    was measured on this workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            text = f.read()
    except OSError:
        return None
    if ("Workload: " + workload) not in text:
        return None
    for line in text.splitlines():
        if line.startswith("stage=%s " % stage):
            kv = dict(x.split("=", 1) for x in line.split())
            return float(kv["traffic_bytes"])
    return None


def _pair_checksums(outs_r, outs_s, counts):
    """Order-independent checksum of each join's pairs: sum of r * 2^32 + s (mod 2^64)."""
    import torch
    sums = []
    for r, s_, n in zip(outs_r, outs_s, counts):
        rr = r[:n].to(torch.int64) & 0xFFFFFFFF
        ss = s_[:n].to(torch.int64) & 0xFFFFFFFF
        sums.append(int(((rr << 32) | ss).sum().item()) if n else 0)
    return sums


def batched_leg(ctx, n_joins, reps, seed=None):
    """256 independent int32 joins of uneven sizes (build 2^U(10,20), probe 4-16x,
    ~50 % match): one batched pipeline replayed from a hipGraph against the same
    joins as sequential plain vj_join_into calls. Both produce every join's pairs;
    counts and pair checksums must agree join by join."""
    import numpy as np
    import torch

    import vjoin
    import workloads
    joins = workloads.batched(n_joins=n_joins, seed=seed)
    builds, probes, outs_r, outs_s, caps, keep = [], [], [], [], [], []
    rows = 0
    for rcols, scols in joins:
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        keep += [dr, ds]
        builds.append(vjoin.table(dr))
        probes.append(vjoin.table(ds))
        cap = len(scols[0])
        rows += len(rcols[0]) + cap
        outs_r.append(torch.empty(cap, dtype=torch.int32, device="cuda"))
        outs_s.append(torch.empty(cap, dtype=torch.int32, device="cuda"))
        caps.append(cap)
    opts = vjoin.join_opts("fnv", 64, 64)
    # autotune the radix plan of every build-size class present (one sample join
    # each); the batched and the plain runs below both use the tuned plans
    t0 = time.perf_counter()
    tuned = {}
    for i, (rcols, _) in enumerate(joins):
        cls = max(len(rcols[0]) - 1, 1).bit_length()
        if cls not in tuned:
            tuned[cls] = ctx.tune(builds[i], probes[i], opts, reps=2)[0]
    t_tune = time.perf_counter() - t0
    del joins
    torch.cuda.synchronize()

    def batch(graph):
        rc, counts = ctx.join_batch(builds, probes, opts, outs_r, outs_s, caps, use_graph=graph)
        if rc != 0:
            raise RuntimeError("batch failed: %s %s" % (vjoin.STATUS.get(rc), ctx.last_error()))
        return counts

    def plain():
        counts = []
        for i in range(len(builds)):
            rc, cnt = ctx.join_into(builds[i], probes[i], opts, outs_r[i], outs_s[i])
            if rc != 0:
                raise RuntimeError("join failed: %s %s" % (vjoin.STATUS.get(rc), ctx.last_error()))
            counts.append(cnt)
        return counts

    # warm-up (allocates the arenas, captures the graph), then each mode timed
    # as a block of back-to-back runs after one untimed run of that mode (the
    # first batch after 256 plain joins runs ~0.25 ms slower in either mode)
    c_plain = plain()
    sum_plain = _pair_checksums(outs_r, outs_s, c_plain)
    c_batch = batch(True)
    sum_batch = _pair_checksums(outs_r, outs_s, c_batch)

    def timed_block(f):
        f()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return ts

    t_plain = timed_block(plain)
    t_batch = timed_block(lambda: batch(True))
    t_eager = timed_block(lambda: batch(False))
    c_last = batch(True)
    same = c_plain == c_batch == c_last and sum_plain == sum_batch == _pair_checksums(outs_r, outs_s, c_last)
    tp, tb, te = float(np.median(t_plain)), float(np.median(t_batch)), float(np.median(t_eager))
    return {"workload": "%d int32 joins, build 2^U(10,20), probe U(4,16)x build, ~50%% match, FNV-1a-64" % n_joins,
            "input_rows": rows, "matches": int(sum(c_plain)), "reps": reps,
            "joins_per_s_batched_graph": n_joins / tb, "joins_per_s_batched_eager": n_joins / te,
            "joins_per_s_plain": n_joins / tp, "speedup_vs_plain": tp / tb,
            "ms_batched_graph": tb * 1e3, "ms_batched_eager": te * 1e3, "ms_plain": tp * 1e3,
            "autotuned_plans": {"2^%d" % k: list(v) for k, v in sorted(tuned.items())}, "tuning_s": t_tune,
            "same_counts_and_checksums": bool(same)}


def serving_leg(n_requests, rates, budget_gb=64, n_pool=64, n_tables=8, n_reprobes=16, seed=7):
    """Poisson arrivals of join requests: 70 % fresh joins with sizes as the
    batched workload, 30 % re-probes of one of `n_tables` resident build sides
    with a fresh probe batch of 2^U(18,24) rows. Served by vj_server (HBM
    budget, continuous batching over two lanes) and, at the same arrival
    times, by a plain loop that runs one join per launch in arrival order.
    Latency = completion - arrival. Every request's pairs must agree (count and
    checksum) between the two."""
    import threading

    import numpy as np
    import torch

    import vjoin
    import workloads
    rng = np.random.default_rng(seed)
    pool = []
    for rcols, scols in workloads.batched(n_joins=n_pool, seed=seed):
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        # the tensors stay referenced: vjoin.table only holds their pointers
        pool.append((vjoin.table(dr), vjoin.table(ds), len(scols[0]), (dr, ds)))
    tables = []
    reprobes = []
    for t in range(n_tables):
        rcols, _ = workloads.batched(n_joins=1, log_min=16, log_max=20, seed=seed * 1000 + t)[0]
        tables.append((rcols, workloads.to_device(rcols)))
    for k in range(n_reprobes):
        t = k % n_tables
        keys = tables[t][0][0]
        m = int(2.0 ** rng.uniform(18, 24))
        s = np.where(rng.random(m) < 0.5, keys[rng.integers(0, len(keys), size=m)],
                     rng.integers(len(keys) * 4, 1 << 30, size=m).astype(keys.dtype))
        ds = workloads.to_device([s.astype(keys.dtype)])
        reprobes.append((t, vjoin.table(ds), m, ds))
    opts = vjoin.join_opts("fnv", 64, 64)
    results = {"workload": "%d requests: 70%% fresh int32 joins (build 2^U(10,20), probe U(4,16)x), 30%% re-probes "
                           "of %d resident tables with 2^U(18,24)-row probes; Poisson arrivals" % (n_requests, n_tables),
               "budget_GB": budget_gb, "rates": []}

    # resident tables: the server keeps them from keep_build requests; the plain loop builds them with vj_build
    srv = vjoin.Server(0, budget=int(budget_gb * (1 << 30)), opts=opts, max_batch=64)
    ctx = vjoin.Context(0)
    tid, built = [], []
    for rcols, dr in tables:
        probe = vjoin.table(workloads.to_device([rcols[0][:1024]]))
        o = (torch.empty(1024, dtype=torch.int32, device="cuda"), torch.empty(1024, dtype=torch.int32, device="cuda"))
        r = srv.wait(srv.submit(probe, o[0], o[1], build=vjoin.table(dr), keep_build=True))
        assert r["status"] == 0 and r["table"] > 0, r
        tid.append(r["table"])
        built.append(ctx.build(vjoin.table(dr), opts))

    for lam in rates:
        arr = np.cumsum(rng.exponential(1.0 / lam, size=n_requests))
        kinds = []
        for i in range(n_requests):
            if rng.random() < 0.3:
                kinds.append(("re", int(rng.integers(0, n_reprobes))))
            else:
                kinds.append(("fresh", int(rng.integers(0, n_pool))))
        outs = []
        for kind, j in kinds:
            n = pool[j][2] if kind == "fresh" else reprobes[j][2]
            outs.append((torch.empty(n, dtype=torch.int32, device="cuda"),
                         torch.empty(n, dtype=torch.int32, device="cuda")))
        torch.cuda.synchronize()

        # served: a submitter thread releases requests at their arrival times
        tickets = [None] * n_requests
        t0_ns = [0]

        def submitter():
            t0 = time.perf_counter()
            t0_ns[0] = time.monotonic_ns()
            for i, (kind, j) in enumerate(kinds):
                while time.perf_counter() - t0 < arr[i]:
                    pass
                o = outs[i]
                if kind == "fresh":
                    tickets[i] = srv.submit(pool[j][1], o[0], o[1], build=pool[j][0])
                else:
                    t, probe, _, _ = reprobes[j]
                    tickets[i] = srv.submit(probe, o[0], o[1], table=tid[t])

        st0 = srv.stats()
        th = threading.Thread(target=submitter)
        th.start()
        resp = []
        for i in range(n_requests):
            while tickets[i] is None:
                time.sleep(0)
            resp.append(srv.wait(tickets[i]))
        th.join()
        st1 = srv.stats()
        assert all(r["status"] == 0 for r in resp), [r for r in resp if r["status"]][:3]
        lat = np.array([(r["t_done_ns"] - r["t_submit_ns"]) * 1e-6 for r in resp])
        span = (max(r["t_done_ns"] for r in resp) - min(r["t_submit_ns"] for r in resp)) * 1e-9
        served_sum = _pair_checksums([o[0] for o in outs], [o[1] for o in outs], [r["count"] for r in resp])
        served_cnt = [r["count"] for r in resp]
        for o in outs:
            o[0].fill_(-1)
            o[1].fill_(-1)
        torch.cuda.synchronize()

        # plain: one join per launch, in arrival order, never before its arrival
        plain_lat, plain_cnt = [], []
        t0 = time.perf_counter()
        for i, (kind, j) in enumerate(kinds):
            while time.perf_counter() - t0 < arr[i]:
                pass
            o = outs[i]
            if kind == "fresh":
                rc, cnt = ctx.join_into(pool[j][0], pool[j][1], opts, o[0], o[1])
            else:
                t, probe, _, _ = reprobes[j]
                rc, cnt = ctx.probe_into(built[t], probe, opts, o[0], o[1])
            assert rc == 0
            plain_lat.append((time.perf_counter() - t0 - arr[i]) * 1e3)
            plain_cnt.append(cnt)
        plain_span = time.perf_counter() - t0 - arr[0]
        plain_sum = _pair_checksums([o[0] for o in outs], [o[1] for o in outs], plain_cnt)
        plain_lat = np.array(plain_lat)
        same = served_cnt == plain_cnt and served_sum == plain_sum
        results["rates"].append({
            "arrival_rate_per_s": lam,
            "served": {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                       "joins_per_s": n_requests / span, "batches": st1["batches"] - st0["batches"],
                       "overlapped_batches": st1["overlapped_batches"] - st0["overlapped_batches"]},
            "plain": {"p50_ms": float(np.percentile(plain_lat, 50)), "p99_ms": float(np.percentile(plain_lat, 99)),
                      "joins_per_s": n_requests / plain_span},
            "same_counts_and_checksums": bool(same)})
        del outs
        torch.cuda.empty_cache()
    st = srv.stats()
    results["server_stats"] = {k: st[k] for k in ("peak_bytes", "resident_tables", "resident_bytes", "max_batch",
                                                  "budget_waits", "budget_deferred")}
    srv.close()
    for b in built:
        b.free()
    ctx.close()
    return results


def serving_budget_leg(budget_gb=64, n_requests=48, seed=11):
    """A burst of large requests that together need far more than the HBM
    budget: 60 % fresh FK joins (build 2^22..2^24 int64 keys, probe 8x, every
    probe row matching once), 40 % re-probes of two resident 2^24-row tables
    with 2^26-row FK probes, all submitted at once. The server admits what fits
    the budget (scratch of the launch + the other lane's arenas + resident
    tables) and holds the rest back until memory frees up; the plain loop runs
    the same requests one join per launch. Every request's pair count and
    checksum must agree between the two."""
    import numpy as np
    import torch

    import vjoin
    import workloads
    rng = np.random.default_rng(seed)
    pool = []
    for k, (lr, ls) in enumerate(((22, 25), (23, 26), (24, 27), (24, 26))):
        rcols, scols = workloads.large_fk(log_r=lr, log_s=ls, seed=seed * 100 + k)
        dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
        pool.append((vjoin.table(dr), vjoin.table(ds), 1 << ls, (dr, ds)))
    res_src = []
    for k in range(2):
        rcols, _ = workloads.large_fk(log_r=24, log_s=10, seed=seed * 100 + 50 + k)
        keys = rcols[0]
        s = keys[rng.integers(0, len(keys), size=1 << 26)]
        dr, ds = workloads.to_device(rcols), workloads.to_device([s])
        res_src.append((vjoin.table(dr), vjoin.table(ds), 1 << 26, (dr, ds)))
    opts = vjoin.join_opts("fnv", 64, 64)
    budget = int(budget_gb * (1 << 30))
    srv = vjoin.Server(0, budget=budget, opts=opts, max_batch=64)
    ctx = vjoin.Context(0)
    tid, built = [], []
    for b, _, _, keep in res_src:
        o = (torch.empty(1024, dtype=torch.int32, device="cuda"), torch.empty(1024, dtype=torch.int32, device="cuda"))
        probe = vjoin.table([keep[0][0][:1024]])
        r = srv.wait(srv.submit(probe, o[0], o[1], build=b, keep_build=True))
        assert r["status"] == 0 and r["table"] > 0, r
        tid.append(r["table"])
        built.append(ctx.build(b, opts))
    kinds = [("fresh", int(rng.integers(0, len(pool)))) if rng.random() < 0.6 else ("re", int(rng.integers(0, 2)))
             for _ in range(n_requests)]
    outs = []
    for kind, j in kinds:
        n = pool[j][2] if kind == "fresh" else res_src[j][2]
        outs.append((torch.empty(n, dtype=torch.int32, device="cuda"), torch.empty(n, dtype=torch.int32, device="cuda")))
    torch.cuda.synchronize()
    st0 = srv.stats()
    t0 = time.monotonic_ns()
    tickets = []
    for (kind, j), o in zip(kinds, outs):
        if kind == "fresh":
            tickets.append(srv.submit(pool[j][1], o[0], o[1], build=pool[j][0]))
        else:
            tickets.append(srv.submit(res_src[j][1], o[0], o[1], table=tid[j]))
    resp = [srv.wait(t) for t in tickets]
    st1 = srv.stats()
    assert all(r["status"] == 0 for r in resp), [r for r in resp if r["status"]][:3]
    lat = np.array([(r["t_done_ns"] - t0) * 1e-6 for r in resp])
    span = (max(r["t_done_ns"] for r in resp) - t0) * 1e-9
    served_cnt = [r["count"] for r in resp]
    served_sum = _pair_checksums([o[0] for o in outs], [o[1] for o in outs], served_cnt)
    for o in outs:
        o[0].fill_(-1)
        o[1].fill_(-1)
    torch.cuda.synchronize()
    plain_cnt, plain_lat = [], []
    t0p = time.perf_counter()
    for (kind, j), o in zip(kinds, outs):
        if kind == "fresh":
            rc, cnt = ctx.join_into(pool[j][0], pool[j][1], opts, o[0], o[1])
        else:
            rc, cnt = ctx.probe_into(built[j], res_src[j][1], opts, o[0], o[1])
        assert rc == 0
        plain_cnt.append(cnt)
        plain_lat.append((time.perf_counter() - t0p) * 1e3)
    plain_span = time.perf_counter() - t0p
    plain_sum = _pair_checksums([o[0] for o in outs], [o[1] for o in outs], plain_cnt)
    st = srv.stats()
    out = {"workload": "burst of %d requests: 60%% fresh FK joins (build 2^22..2^24 int64 keys, probe 8x), 40%% "
                       "re-probes of 2 resident 2^24-row tables with 2^26-row probes" % n_requests,
           "budget_GB": budget_gb,
           "served": {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                      "joins_per_s": n_requests / span, "batches": st1["batches"] - st0["batches"],
                      "budget_waits": st1["budget_waits"] - st0["budget_waits"],
                      "budget_deferred": st1["budget_deferred"] - st0["budget_deferred"]},
           "plain": {"p50_ms": float(np.percentile(plain_lat, 50)), "p99_ms": float(np.percentile(plain_lat, 99)),
                     "joins_per_s": n_requests / plain_span},
           "peak_bytes": st["peak_bytes"], "peak_within_budget": st["peak_bytes"] <= budget,
           "same_counts_and_checksums": served_cnt == plain_cnt and served_sum == plain_sum}
    srv.close()
    for b in built:
        b.free()
    ctx.close()
    del outs, pool, res_src
    torch.cuda.empty_cache()
    return out


def multicol_data():
    """Device tables of the multi-column skewed join and its number of hitting probe rows."""
    import numpy as np

    import workloads
    rcols, scols = workloads.multicol_skew()
    n_hit = int(np.count_nonzero(scols[0] < np.uint32(0x80000000)))
    return workloads.to_device(rcols), workloads.to_device(scols), n_hit


def multicol_leg(ctx, steps, data, tune=True, fixed_plan=None):
    """Multi-column skewed join: 4 x uint32 key columns, build 2^22 rows (2^20
    distinct keys, each 4 times), probe 2^25 rows (80 % Zipf(1.0) over the
    distinct keys, 20 % outside them), with FNV-1a-64 and custom_xor_hw-64;
    two radix passes, bits and kernel shapes autotuned per hash (vj_tune on
    these tables) unless tune is off (then the planner's plan).
    Checked by size-independent properties: the match count is 4 x the hitting
    probe rows, every pair has equal keys in all 4 columns, every probe row
    appears 0 or 4 times. (The tests compare the same join with the oracle.)"""
    import numpy as np
    import torch

    import vjoin
    dr, ds, n_hit = data
    nr, ns = int(dr[0].numel()), int(ds[0].numel())
    m_want = 4 * n_hit
    out_r = torch.empty(m_want, dtype=torch.int32, device="cuda")
    out_s = torch.empty(m_want, dtype=torch.int32, device="cuda")
    res = {"workload": "build 2^22 rows of 4 x uint32 keys (2^20 distinct, x4), probe 2^25 rows (80% Zipf(1.0), 20% "
                       "misses), 2 radix passes, bits and kernel shapes autotuned" if tune else
                       "build 2^22 rows of 4 x uint32 keys (2^20 distinct, x4), probe 2^25 rows (80% Zipf(1.0), 20% "
                       "misses), 2 radix passes, planner bits", "build_rows": nr, "probe_rows": ns, "matches": m_want}
    for kind in ("fnv", "custom_xor_hw"):
        b, p = vjoin.table(dr), vjoin.table(ds)
        opts = vjoin.join_opts(kind, 64, 64, passes=2)
        plan = None
        if fixed_plan:
            opts = vjoin.join_opts(kind, 64, 64, passes=fixed_plan[0], bits1=fixed_plan[1], bits2=fixed_plan[2],
                                   tile=fixed_plan[3], table=fixed_plan[4],
                                   layout=fixed_plan[5] if len(fixed_plan) > 5 else 0)
        elif tune:
            plan, _ = ctx.tune(b, p, opts, reps=3)
            opts = vjoin.join_opts(kind, 64, 64, passes=plan[0], bits1=plan[1], bits2=plan[2], tile=plan[3],
                                   table=plan[4], layout=plan[5])
        ms, stages = [], {k: [] for k in ("ms_histogram", "ms_pass1", "ms_pass2", "ms_build_probe")}
        for i in range(steps + 1):
            rc, cnt = ctx.join_into(b, p, opts, out_r, out_s)
            if rc != 0:
                raise RuntimeError("join failed: %s %s" % (vjoin.STATUS.get(rc), ctx.last_error()))
            if i:
                sm = ctx.summary()
                ms.append(sm["ms_total"])
                for k in stages:
                    stages[k].append(sm[k])
        sid, rid = out_s.long(), out_r.long()
        eq = all(bool(torch.equal(dr[c][rid], ds[c][sid])) for c in range(4))
        per_row = torch.bincount(sid, minlength=ns)
        ok = cnt == m_want and eq and int((per_row == 4).sum()) == n_hit and int((per_row != 0).sum()) == n_hit
        del sid, rid, per_row
        t = float(np.median(ms))
        B = 5 * (nr + ns) * 20 + m_want * 8
        res[kind] = {"ms": t, "stage_ms": {k[3:]: float(np.median(v)) for k, v in stages.items()},
                     "plan": {"passes": sm["passes"], "bits": [sm["bits1"], sm["bits2"]], "tile_rows": sm["tile_rows"],
                              "probe_threads": sm["probe_threads"], "table_rows": sm["table_rows"],
                              "autotuned": plan is not None},
                     "probe_tuples_per_s": ns / (t * 1e-3), "algorithmic_bytes": B,
                     "achieved_GBps": B / (t * 1e-3) / 1e9, "frac": B / (t * 1e-3) / HBM_PEAK, "parity": bool(ok)}
    return res


def bigbuild_leg(ctx, steps=5, log_r=26, log_s=27):
    """A build side past the standard plan: 2^26 unique int64 keys (FK probe of
    2^27 rows), FNV-1a-64. With 2048-row tables it would need 16 radix bits,
    past two passes' 14; the planner's wide build+probe block (1024 threads,
    8192-row tables) keeps one build chunk per partition at 14 bits. Timed
    against the same join forced onto the standard block (several build chunks
    per partition, so every probe chunk is read once per build chunk), and the
    wide block in the region layout (no histogram read). Checked as the FK
    join: every probe row once, with an equal key."""
    import numpy as np
    import torch

    import vjoin
    import workloads
    rcols, scols = workloads.large_fk(log_r=log_r, log_s=log_s, seed=26)
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    del rcols, scols
    nr, ns = int(dr[0].numel()), int(ds[0].numel())
    out_r = torch.empty(ns, dtype=torch.int32, device="cuda")
    out_s = torch.empty(ns, dtype=torch.int32, device="cuda")
    res = {"workload": "build 2^%d unique int64 keys, FK probe 2^%d rows, FNV-1a-64, 2 radix passes" % (log_r, log_s),
           "build_rows": nr, "probe_rows": ns}
    for name, table, layout in (("wide_table", 0, 0), ("wide_table_regions", 0, vjoin.LAYOUT_REGIONS),
                                ("standard_table", 1, 0)):
        opts = vjoin.join_opts("fnv", 64, 64, table=table, layout=layout)
        b, p = vjoin.table(dr), vjoin.table(ds)
        ms, stages = [], {k: [] for k in ("ms_histogram", "ms_pass1", "ms_pass2", "ms_build_probe")}
        for i in range(steps + 1):
            rc, cnt = ctx.join_into(b, p, opts, out_r, out_s)
            if rc != 0:
                raise RuntimeError("join failed: %s %s" % (vjoin.STATUS.get(rc), ctx.last_error()))
            if i:
                sm = ctx.summary()
                ms.append(sm["ms_total"])
                for k in stages:
                    stages[k].append(sm[k])
        sid = out_s.long()
        ok = cnt == ns and bool(torch.equal(torch.sort(sid).values, torch.arange(ns, device="cuda"))) and \
            bool(torch.equal(dr[0][out_r.long()], ds[0][sid]))
        del sid
        t = float(np.median(ms))
        B = algorithmic_bytes(nr, ns, ns, sm["passes"])
        res[name] = {"ms": t, "stage_ms": {k[3:]: float(np.median(v)) for k, v in stages.items()},
                     "passes": sm["passes"], "bits": [sm["bits1"], sm["bits2"]], "probe_threads": sm["probe_threads"],
                     "table_rows": sm["table_rows"], "tasks": sm["tasks"], "probe_tuples_per_s": ns / (t * 1e-3),
                     "layout": "regions" if sm["layout"] == vjoin.LAYOUT_REGIONS else "exact",
                     "tile_rows": sm["tile_rows"],
                     "algorithmic_bytes": B, "frac": B / (t * 1e-3) / HBM_PEAK, "parity": bool(ok)}
    return res


def hash_leg(shapes, reps=10):
    """vj_hash (the standalone hash kernel, k_hash) over the probe-side keys of the
    large FK join and of the multi-column join, for FNV-1a and custom_xor_hw,
    HASH_BITS 32 / 64 and chunk widths 32 / 64 / 128 bits (where the row is a
    whole number of chunks), as the reference's hash benchmark sweeps
    (src/playground/hash_benchmark.cu:59-169). Bytes per launch:
    B_hash = n x (key bytes read) + n x (hash bytes written); time from HIP
    events on the launch stream, median of `reps` after a warm-up."""
    import numpy as np
    import torch

    import vjoin
    rows = []
    for name, cols in shapes:
        n = int(cols[0].numel())
        key_bytes = sum(c.element_size() for c in cols)
        words = 2 * len(cols)  # reference words: every key column widened to 64 bits
        for kind in ("fnv", "custom_xor_hw"):
            for hb in (32, 64):
                out = torch.empty(n, dtype=torch.int32 if hb == 32 else torch.int64, device="cuda")
                for cb in (32, 64, 128):
                    if (words * 32) % cb:
                        continue
                    vjoin.hash_table(cols, kind, hb, cb, out=out)  # warm-up
                    ts = []
                    for _ in range(reps):
                        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        a.record()
                        vjoin.hash_table(cols, kind, hb, cb, out=out)
                        b.record()
                        b.synchronize()
                        ts.append(a.elapsed_time(b))
                    ms = float(np.median(ts))
                    B = n * key_bytes + n * hb // 8
                    rows.append({"keys": name, "hash": kind, "hash_bits": hb, "chunk_bits": cb, "rows": n,
                                 "bytes": B, "ms": ms, "GBps": B / (ms * 1e-3) / 1e9,
                                 "frac": B / (ms * 1e-3) / HBM_PEAK, "tuples_per_s": n / (ms * 1e-3)})
                del out
    return {"kernel": "k_hash", "unit": "GB/s", "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-r", type=int, default=24)
    ap.add_argument("--log-s", type=int, default=27)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--legs", default="fk,multicol,hash,bigbuild,batched,serving,cpu",
                    help="comma list of legs to run: fk (the headline line), multicol, hash, bigbuild, batched, "
                         "serving, cpu")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hash", action="store_true", help="skip the hash kernel leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batched-joins", type=int, default=256)
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--no-serving", action="store_true", help="skip the served-vs-plain request stream")
    ap.add_argument("--no-tune", action="store_true", help="use the planner's radix bits instead of vj_tune")
    ap.add_argument("--mc-plan", default="",
                    help="fixed multicol plan 'passes,bits1,bits2,tile,table[,layout]' (profiling runs)")
    ap.add_argument("--layout", type=int, default=0,
                    help="partition layout of the FK join with --no-tune (0 planner, 1 exact, 2 regions)")
    ap.add_argument("--no-multicol", action="store_true", help="skip the multi-column skewed join leg")
    ap.add_argument("--serving-requests", type=int, default=2048,
                    help="requests per arrival rate (p99 is the 20th-worst of 2048)")
    ap.add_argument("--serving-rates", type=float, nargs="+", default=[2000.0, 8000.0, 32000.0])
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device to device three times (counter calibration)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU: write the traffic summary from two rocprofv3 --pmc results (run_results.db or "
                         "counter_collection.csv)")
    ap.add_argument("--counters-from", nargs="+", metavar="RUN_RESULTS_DB",
                    help="summarize rocprofv3 --pmc SQ counter runs (per kernel, per tuple) and exit")
    ap.add_argument("--counters-rows", type=int, default=0, help="tuples per dispatch (for --counters-from)")
    ap.add_argument("--counters-out", default="", help="output file (for --counters-from)")
    ap.add_argument("--traffic-workload", default="", help="workload named in the --traffic-from summary")
    ap.add_argument("--traffic-out", default="", help="output file of --traffic-from (default: the FK summary)")
    ap.add_argument("--traffic-leg", default="--legs fk --steps 1 --warmup 1 --no-tune",
                    help="the bench arguments the --traffic-from runs used")
    ap.add_argument("--trace-summary", nargs=2, metavar=("RUN_RESULTS_DB", "OUT_MD"),
                    help="no GPU: write a markdown per-kernel summary of a rocprofv3 --kernel-trace database")
    ap.add_argument("--trace-command", default="", help="the command that made the trace (for --trace-summary)")
    ap.add_argument("--trace-workload", default="", help="what the trace covers (for --trace-summary)")
    args = ap.parse_args()
    if args.trace_summary:
        text = trace_summary(args.trace_summary[0], args.trace_command, args.trace_workload)
        with open(args.trace_summary[1], "w") as f:
            f.write(text)
        print(text, end="")
        return
    workload_tag = "large_fk build 2^%d probe 2^%d passes %d" % (args.log_r, args.log_s, args.passes)
    if args.traffic_from:
        text = traffic_from_pmc(args.traffic_from[0], args.traffic_from[1], args.traffic_workload or workload_tag,
                                args.traffic_leg)
        with open(args.traffic_out or TRAFFIC_FILE, "w") as f:
            f.write(text)
        print(text, end="")
        return
    if args.counters_from:
        cmd = ("rocprofv3 --pmc <counters> -d D -o run -- python3 bench.py %s, once per counter set; "
               "python3 bench.py --counters-from D/run_results.db ... --counters-rows %d" % (args.traffic_leg,
                                                                                         args.counters_rows))
        text = counters_from_pmc(args.counters_from, args.counters_rows, cmd, args.traffic_workload or workload_tag)
        if args.counters_out:
            with open(args.counters_out, "w") as f:
                f.write(text)
        print(text, end="")
        return
    if args.gpus != 1:
        print("bench.py: only one GPU is measured; across GPUs the join would shard by a top radix digit "
              "of the hash, but no multi-GPU path exists. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import vjoin
    import workloads

    legs = set(x for x in args.legs.split(",") if x)
    for off, leg in ((args.no_cpu, "cpu"), (args.no_batched, "batched"), (args.no_serving, "serving"),
                     (args.no_multicol, "multicol"), (args.no_hash, "hash")):
        if off:
            legs.discard(leg)
    torch.cuda.set_device(0)
    ctx = vjoin.Context(0)
    ctx.set_profiling(True)
    out = {}
    fk = None
    if "fk" in legs:
        fk = fk_leg(ctx, args, workload_tag, keep_host="cpu" in legs)
        out = fk["line"]
    mc = multicol_data() if ("multicol" in legs or "hash" in legs) else None
    if "multicol" in legs:
        print("bench: multi-column skewed join leg", file=sys.stderr, flush=True)
        fixed = [int(x) for x in args.mc_plan.split(",")] if args.mc_plan else None
        out["multicol_skew"] = multicol_leg(ctx, 5, mc, tune=not args.no_tune, fixed_plan=fixed)
    if "hash" in legs:
        print("bench: hash kernel leg", file=sys.stderr, flush=True)
        shapes = []
        if fk:
            shapes.append(("large_fk probe keys: 2^%d x int64" % args.log_s, [fk["ds"][0]]))
        shapes.append(("multicol probe keys: 2^25 x 4 uint32", mc[1]))
        out["hash"] = hash_leg(shapes)
    if fk:
        fk.pop("ds", None), fk.pop("dr", None), fk.pop("out", None)
    mc = None
    torch.cuda.empty_cache()
    if "bigbuild" in legs:
        print("bench: 2^26 x 2^27 build leg (wide and standard build+probe blocks)", file=sys.stderr, flush=True)
        out["bigbuild"] = bigbuild_leg(ctx)
        torch.cuda.empty_cache()
    if args.calib_copy:  # a plain 1 GiB device copy, three times: the scale of FETCH_SIZE / WRITE_SIZE
        x = torch.empty(CALIB_BYTES // 8, dtype=torch.int64, device="cuda").fill_(1)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    if "batched" in legs:
        print("bench: batched leg (%d joins)" % args.batched_joins, file=sys.stderr, flush=True)
        out["batched"] = batched_leg(ctx, args.batched_joins, args.batched_reps)
    if "serving" in legs:
        torch.cuda.empty_cache()
        print("bench: serving leg (%d requests at %s /s)" % (args.serving_requests, args.serving_rates),
              file=sys.stderr, flush=True)
        out["serving"] = serving_leg(args.serving_requests, args.serving_rates)
        torch.cuda.empty_cache()
        print("bench: serving leg under the 64 GB budget", file=sys.stderr, flush=True)
        out["serving_budget"] = serving_budget_leg()
    if "cpu" in legs and fk:
        print("bench: cpu baseline (OpenMP oracle)", file=sys.stderr, flush=True)
        summ = fk["summary"]
        out["cpu_baseline"] = cpu_baseline(fk["r_host"], fk["s_host"], summ["passes"], summ["bits1"], summ["bits2"])
    ctx.close()
    print(json.dumps(out))


def fk_leg(ctx, args, workload_tag, keep_host):
    """The headline: the large FK join, `args.steps` timed steps after `args.warmup`."""
    import numpy as np
    import torch

    import vjoin
    import workloads
    print("bench: generating large_fk build 2^%d probe 2^%d" % (args.log_r, args.log_s), file=sys.stderr, flush=True)
    rcols, scols = workloads.large_fk(log_r=args.log_r, log_s=args.log_s)
    nr, ns = len(rcols[0]), len(scols[0])
    dr, ds = workloads.to_device(rcols), workloads.to_device(scols)
    r_host, s_host = (rcols[0], scols[0]) if keep_host else (None, None)
    del rcols, scols
    build, probe = vjoin.table(dr), vjoin.table(ds)
    if args.no_tune:  # planner bits (profiling runs: no extra joins in the trace)
        opts = vjoin.join_opts("fnv", 64, 64, passes=args.passes, layout=args.layout)
    else:  # radix bits autotuned for the configured pass count (vj_tune, on these tables)
        (tp, tb1, tb2, tt, ttab, tlay), tune_ms = ctx.tune(build, probe,
                                                           vjoin.join_opts("fnv", 64, 64, passes=args.passes), reps=3)
        opts = vjoin.join_opts("fnv", 64, 64, passes=tp, bits1=tb1, bits2=tb2, tile=tt, table=ttab, layout=tlay)
        print("bench: autotuned plan passes=%d bits=%d+%d tile=%d table=%d layout=%d (%.3f ms)"
              % (tp, tb1, tb2, tt, ttab, tlay, tune_ms), file=sys.stderr, flush=True)
    t_alloc = time.perf_counter()
    out_r = torch.empty(ns, dtype=torch.int32, device="cuda")
    out_s = torch.empty(ns, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t_alloc = time.perf_counter() - t_alloc

    def step():
        rc, cnt = ctx.join_into(build, probe, opts, out_r, out_s)
        if rc != 0:
            raise RuntimeError("join failed: %s %s" % (vjoin.STATUS.get(rc), ctx.last_error()))
        return cnt

    t0 = time.perf_counter()
    for _ in range(args.warmup):
        m = step()
    t_warm = time.perf_counter() - t0
    print("bench: warmup done (%.2f s), first-join scratch %.2f GB" % (t_warm, ctx.memory() / 1e9), file=sys.stderr,
          flush=True)
    # correctness of the last warm-up output (size-independent property of an FK join:
    # with unique build keys, every probe row once with an equal key is the whole join)
    sid = out_s.long()
    ok = bool(torch.equal(torch.sort(sid).values, torch.arange(ns, device="cuda"))) and \
        bool(torch.equal(dr[0][out_r.long()], ds[0][sid]))
    del sid
    stages = {"histogram": [], "pass1": [], "pass2": [], "build_probe": [], "total": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m = step()
        s = ctx.summary()
        stages["total"].append(s["ms_total"])
        stages["histogram"].append(s["ms_histogram"])
        stages["pass1"].append(s["ms_pass1"])
        stages["pass2"].append(s["ms_pass2"])
        stages["build_probe"].append(s["ms_build_probe"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    summ = ctx.summary()
    ms_step = elapsed / args.steps * 1e3
    med = {k: float(np.median(v)) for k, v in stages.items()}
    regions = summ["layout"] == vjoin.LAYOUT_REGIONS
    layout_name = "regions" if regions else "exact"
    sb = stage_bytes(nr, ns, m, regions)
    dominant = max(("histogram", "pass1", "pass2", "build_probe"), key=lambda k: med[k])
    achieved = sb[dominant] / (float(np.mean(stages[dominant])) * 1e-3)
    B = algorithmic_bytes(nr, ns, m, summ["passes"])
    join_gbs = B / (med["total"] * 1e-3)
    line = {
        "metric": "probe tuples/s + HBM GB/s (frac of peak); batched vs plain joins/s, p99 latency",
        "value": ns / (ms_step * 1e-3),
        "unit": "probe tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded: workloads.large_fk)",
        "config": {"workload": "large_fk: build 2^%d unique int64 keys (mix64 of a permutation), probe 2^%d FK "
                               "keys (100%% match), FNV-1a-64, %d radix passes, bits and layout %s" % (
                                  args.log_r, args.log_s, summ["passes"],
                                  "of the planner" if args.no_tune else "autotuned"),
                   "build_rows": nr, "probe_rows": ns, "matches": m, "passes": summ["passes"],
                   "bits": [summ["bits1"], summ["bits2"]], "layout": layout_name, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "%s (%s)" % (STAGE_KERNEL[dominant], dominant),
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": read_traffic(dominant, "%s layout %s" % (workload_tag, layout_name)),
                     "algorithmic_bytes_per_launch": sb[dominant]},
        "join_roofline": {"algorithmic_bytes": B, "ms": med["total"], "achieved_GBps": join_gbs / 1e9,
                          "frac": join_gbs / HBM_PEAK},
        "stage_ms_median": med,
        "stage_GBps": {k: sb[k] / (med[k] * 1e-3) / 1e9 for k in sb if med[k] > 0},
        "library_summary": {k: summ[k] for k in ("tasks", "partitions", "bytes_total", "gbps_total",
                                                 "probe_tuples_per_s", "scratch_bytes")},
        "output_alloc_ms": t_alloc * 1e3,
        "parity": {"fk_every_probe_row_once_equal_keys": ok},
    }
    return {"line": line, "summary": summ, "dr": dr, "ds": ds, "out": (out_r, out_s), "r_host": r_host,
            "s_host": s_host}


if __name__ == "__main__":
    main()
