#!/usr/bin/env python3
"""Sweep driver with the reference's profile schema and result columns.

Reads a TOML profile in the schema of the reference's join benchmark
(src/benchmark/configuration.hpp:307-570: launch, data, join, hash and probe
fields of one [profile] table), expands the same cartesian product of
configurations in the same order (get_benchmark_configs, get_join_configs,
get_hash_benchmark_configs, get_probe_benchmark_configs,
configuration.hpp:180-300), runs every configuration `runs` times on one GPU
through the reference-layout join (vj_join_rows in libvj.so) and appends one
line per configuration to OUTPUT/run.csv with the reference's columns
(JoinBenchmarkResults, src/join_benchmark.cu:15-33), averaged over the runs
like join_benchmark.cu:278-291. With profile = true, OUTPUT/partition.csv
gets the partition summary lines (PartitionSummary, join_provider.cu:56-83).

Data: R has elements / rs_scale rows and S has elements rows, every key
column of a row holding the same value, primary keys 1..n, as the reference's
generate_tables (src/benchmark/data_generator.cu:148-230): S uniform over
[offset, offset + max_value) with offset = ceil((skew + 0.1) * max_value), or
Zipf(skew) ranks 1..min(max_value, 10000); R uniform with skew 0. The
reference draws with time-seeded cuRAND; here a seeded numpy generator
(seed per configuration and run) makes sweeps repeatable.

Knobs the library sizes itself (hash / probe thread counts, elements per
thread, vector_bytes_size, probe_modes, max_streams, join_modes) are carried
into the CSV unchanged. gpus > 1 is recorded as an unsuccessful run (this
build joins on one GPU). verify = true checks every run: each pair's keys are
equal and the pair count equals a numpy count of the equi-join.

Usage: python sweep.py PROFILE.toml PROFILE_NAME [--output DIR] [--dry-run]
"""
import argparse
import csv
import math
import os
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

# (field, type) per group, as configuration.hpp loads them
LAUNCH = [("gpus", int), ("gpu_modes", int), ("max_streams", int), ("runs", int), ("elements", int),
          ("max_values", int), ("skews", float), ("distribution", str)]
FLAGS = ["verify", "profile"]
JOIN = [("rs_scale", float), ("rs_join_columns", int), ("buckets", int), ("vector_bytes_sizes", int),
        ("join_modes", int)]
HASH = [("hash_elements_sizes", int), ("hash_algorithms", str), ("hash_threads", int), ("hashes_p_threads", int)]
PROBE = [("probe_build_table_loads", float), ("probe_build_n_per_threads", int), ("probe_build_threads", int),
         ("probe_extract_n_per_threads", int), ("probe_extract_threads", int), ("probe_modes", int)]

BENCH_HEADER = ["runs", "gpus", "gpu_mode", "elements", "max_value", "max_column_value", "skew", "distribution",
                "max_streams_p_gpu", "hash_bytes"]
PROBE_HEADER = ["probe_mode", "probe_build_table_load", "probe_build_threads", "probe_build_n_per_thread",
                "probe_extract_threads", "probe_extract_n_per_thread"]
JOIN_HEADER = ["rs_columns", "rs_scale", "buckets", "vector_bytes_size", "join_mode"]
RESULT_HEADER = ["hash_tuple_p_second", "partition_tuple_p_second", "probe_tuple_p_second", "merge_tuple_p_second",
                 "join_tuple_p_second", "join_gb_p_second", "successful"]
RUN_HEADER = BENCH_HEADER + PROBE_HEADER + JOIN_HEADER + RESULT_HEADER
PARTITION_HEADER = ["device_index", "max_streams", "hash", "stream", "depth", "buckets", "elements", "vector_bytes",
                    "kernel", "runtime", "gb_p_s", "tuples_p_s"]


class ProfileError(ValueError):
    pass


def load_profile(path, name):
    """The [name] table of a TOML profile with every field the join benchmark
    reads; a missing field is an error naming it (configuration.hpp:307-570)."""
    try:
        import tomllib as toml_lib
    except ImportError:
        import tomli as toml_lib
    with open(path, "rb") as f:
        doc = toml_lib.load(f)
    if name not in doc:
        raise ProfileError("profile %s not found in %s" % (name, path))
    p = doc[name]
    setup = {}
    for field, typ in LAUNCH + JOIN + HASH + PROBE:
        if field not in p:
            raise ProfileError("%s.%s not found" % (name, field))
        setup[field] = [typ(v) for v in p[field]]
    for field in FLAGS:
        if field not in p:
            raise ProfileError("%s.%s not found" % (name, field))
        setup[field] = bool(p[field])
    if "output" not in p:
        raise ProfileError("%s.output not found" % name)
    setup["output"] = str(p["output"])
    return setup


def expand(setup):
    """Every configuration, in the reference's loop order (benchmark x join x hash x probe)."""
    bench = [dict(runs=runs, max_streams=ms, gpus=g, gpu_mode=gm, elements=e, max_value=mv, skew=sk,
                  distribution=d)
             for runs in setup["runs"] for ms in setup["max_streams"] for g in setup["gpus"]
             for gm in setup["gpu_modes"] for e in setup["elements"] for mv in setup["max_values"]
             for sk in setup["skews"] for d in setup["distribution"]]
    join = [dict(rs_scale=sc, rs_columns=c, buckets=b, vector_bytes_size=vb, join_mode=jm)
            for sc in setup["rs_scale"] for c in setup["rs_join_columns"] for b in setup["buckets"]
            for vb in setup["vector_bytes_sizes"] for jm in setup["join_modes"]]
    hashc = [dict(hash_threads=t, hashes_p_thread=h, hash_elements_size=es, algorithm=a)
             for t in setup["hash_threads"] for h in setup["hashes_p_threads"]
             for es in setup["hash_elements_sizes"] for a in setup["hash_algorithms"]]
    probe = [dict(probe_mode=pm, build_n_per_thread=bn, build_threads=bt, build_table_load=bl,
                  extract_n_per_thread=en, extract_threads=et)
             for pm in setup["probe_modes"] for bn in setup["probe_build_n_per_threads"]
             for bt in setup["probe_build_threads"] for bl in setup["probe_build_table_loads"]
             for en in setup["probe_extract_n_per_threads"] for et in setup["probe_extract_threads"]]
    return [(b, j, h, p) for b in bench for j in join for h in hashc for p in probe]


def generate(n, cols, max_value, skew, distribution, rng):
    """Row-major uint64 keys (every column of a row equal) and primary keys 1..n."""
    import numpy as np
    if distribution == "zipf":
        max_rank = min(10000, max_value)
        ranks = np.arange(1, max_rank + 1, dtype=np.float64)
        cdf = np.cumsum(1.0 / ranks ** skew)
        cdf /= cdf[-1]
        u = np.clip(rng.random(n), 1e-5, 1.0 - 1e-5)
        v = (np.searchsorted(cdf, u, side="left") + 1).astype(np.uint64)
        effective_max = max_rank
    else:
        offset = int(math.ceil((skew + 0.1) * max_value))
        v = (rng.random(n) * max_value).astype(np.uint64) + np.uint64(offset)
        effective_max = max_value
    keys = np.repeat(v.reshape(-1, 1), cols, axis=1)
    pks = np.arange(1, n + 1, dtype=np.uint64)
    return np.ascontiguousarray(keys), pks, effective_max


def numpy_match_count(r_keys, s_keys):
    """Pairs of an equi-join on the first key column (all columns of a row are equal here)."""
    import numpy as np
    rs = np.sort(r_keys[:, 0])
    return int((np.searchsorted(rs, s_keys[:, 0], side="right") - np.searchsorted(rs, s_keys[:, 0], side="left")).sum())


def run_config(ctx, bench, join, hashc, seed, verify):
    """One configuration, bench['runs'] times: averaged rates (join_benchmark.cu:278-291)."""
    import numpy as np
    import torch

    import vjoin
    rates = dict.fromkeys(RESULT_HEADER[:-1], 0.0)
    ok = bench["gpus"] == 1
    partition_rows = []
    max_column_value = bench["max_value"]
    for run in range(bench["runs"]):
        if not ok:
            break
        rng = np.random.default_rng([seed, run])
        ns = int(bench["elements"])
        nr = int(ns / join["rs_scale"])
        s_keys, s_pk, max_column_value = generate(ns, join["rs_columns"], bench["max_value"], bench["skew"],
                                                  bench["distribution"], rng)
        r_keys, r_pk, _ = generate(nr, join["rs_columns"], max_column_value, 0.0, "uniform", rng)
        dev = [torch.from_numpy(a.view(np.int64)).cuda() for a in (r_keys, r_pk, s_keys, s_pk)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            # an unknown algorithm name fails the configuration (the reference's
            # hash_func silently hashed nothing, hash.cu:515-541)
            opts = vjoin.join_opts(hashc["algorithm"], 64, 64)
            if join["buckets"] > 1:
                bits = max(1, min(8, int(math.ceil(math.log2(join["buckets"])))))
                opts.bits1 = opts.bits2 = bits
            vals, pks = ctx.join_rows(dev[0], dev[1], dev[2], dev[3], opts)
        except vjoin.VJError as e:
            print("Join failed %s" % e, file=sys.stderr)
            ok = False
            break
        wall = time.perf_counter() - t0
        sm = ctx.summary()
        n = nr + ns

        def rate(ms):
            return n / (ms * 1e-3) if ms > 0 else 0.0
        rates["hash_tuple_p_second"] += rate(sm["ms_histogram"])
        rates["partition_tuple_p_second"] += rate(sm["ms_histogram"] + sm["ms_pass1"] + sm["ms_pass2"])
        rates["probe_tuple_p_second"] += rate(sm["ms_build_probe"])
        rates["join_tuple_p_second"] += n / wall
        rates["join_gb_p_second"] += n * join["rs_columns"] * 8 / 1e9 / wall
        if verify:
            m = int(vals.shape[0])
            rv = torch.from_numpy(r_keys.view(np.int64)).cuda()
            sv = torch.from_numpy(s_keys.view(np.int64)).cuda()
            same = m == numpy_match_count(r_keys, s_keys)
            if m:
                ri, si = vals[:, 0] - 1, vals[:, 1] - 1  # primary keys are row + 1
                same = same and bool(torch.equal(rv[ri], sv[si]))
                same = same and bool(torch.equal(pks, torch.arange(1, m + 1, device="cuda")))
            if not same:
                print("Join verification failed", file=sys.stderr)
                ok = False
        for kernel, ms, gbps in (("histogram", sm["ms_histogram"], sm["gbps_histogram"]),
                                 ("prefix", 0.0, 0.0),
                                 ("swap", sm["ms_pass1"] + sm["ms_pass2"],
                                  (sm["bytes_pass1"] + sm["bytes_pass2"]) / ((sm["ms_pass1"] + sm["ms_pass2"]) * 1e6)
                                  if sm["ms_pass1"] + sm["ms_pass2"] > 0 else 0.0)):
            partition_rows.append([0, bench["max_streams"], hashc["algorithm"], 0, sm["passes"], join["buckets"], n,
                                   join["vector_bytes_size"], kernel, ms * 1e-3, gbps, rate(ms) if ms else 0.0])
        del dev, vals, pks
    runs = max(1, bench["runs"])
    for k in rates:
        rates[k] /= runs
    return rates, ok, max_column_value, partition_rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("profile_file")
    ap.add_argument("profile_name")
    ap.add_argument("--output", help="results directory (default: the profile's output); must not exist")
    ap.add_argument("--dry-run", action="store_true", help="list the configurations and exit")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    setup = load_profile(args.profile_file, args.profile_name)
    configs = expand(setup)
    if args.dry_run:
        for i, (b, j, h, p) in enumerate(configs):
            print(i, b, j, h, p)
        print("%d configurations" % len(configs))
        return 0
    out_dir = args.output or setup["output"]
    if os.path.exists(out_dir):  # as load_benchmark_results_location (configuration.hpp:552-568)
        print("%s already exists" % out_dir, file=sys.stderr)
        return 1
    os.makedirs(out_dir)
    import torch

    import vjoin
    torch.cuda.set_device(args.device)
    ctx = vjoin.Context(args.device)
    ctx.set_profiling(True)
    run_f = open(os.path.join(out_dir, "run.csv"), "w", newline="")
    run_w = csv.writer(run_f)
    run_w.writerow(RUN_HEADER)
    part_w = None
    if setup["profile"]:
        part_f = open(os.path.join(out_dir, "partition.csv"), "w", newline="")
        part_w = csv.writer(part_f)
        part_w.writerow(PARTITION_HEADER)
    try:
        for i, (b, j, h, p) in enumerate(configs):
            rates, ok, max_column_value, part_rows = run_config(ctx, b, j, h, seed=i, verify=setup["verify"])
            run_w.writerow([b["runs"], b["gpus"], b["gpu_mode"], b["elements"], b["max_value"], max_column_value,
                            b["skew"], b["distribution"], b["max_streams"], 8,
                            p["probe_mode"], p["build_table_load"], p["build_threads"], p["build_n_per_thread"],
                            p["extract_threads"], p["extract_n_per_thread"],
                            j["rs_columns"], j["rs_scale"], j["buckets"], j["vector_bytes_size"], j["join_mode"]] +
                           [rates[k] for k in RESULT_HEADER[:-1]] + [1 if ok else 0])
            run_f.flush()
            if part_w:
                for row in part_rows:
                    part_w.writerow(row)
            print("%d/%d" % (i + 1, len(configs)), flush=True)
    finally:
        run_f.close()
        if part_w:
            part_f.close()
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
