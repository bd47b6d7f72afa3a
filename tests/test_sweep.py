"""The sweep driver (vectorized-join_amd/sweep.py): the reference's profile
schema, configuration product and result columns; on a GPU, a small sweep
whose every run is verified."""
import csv
import os
import subprocess
import sys

import pytest

import sweep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILE = os.path.join(ROOT, "tests", "golden", "sweep_profile.toml")
SWEEP = os.path.join(ROOT, "vectorized-join_amd", "sweep.py")

# run.csv columns of the reference: BenchmarkConfig, ProbeBenchmarkConfig and
# JoinBenchmarkConfig headers (src/benchmark/configuration.hpp), then the join
# rates (src/join_benchmark.cu:15-33)
REFERENCE_RUN_COLUMNS = (
    "runs,gpus,gpu_mode,elements,max_value,max_column_value,skew,distribution,max_streams_p_gpu,hash_bytes,"
    "probe_mode,probe_build_table_load,probe_build_threads,probe_build_n_per_thread,probe_extract_threads,"
    "probe_extract_n_per_thread,rs_columns,rs_scale,buckets,vector_bytes_size,join_mode,hash_tuple_p_second,"
    "partition_tuple_p_second,probe_tuple_p_second,merge_tuple_p_second,join_tuple_p_second,join_gb_p_second,"
    "successful").split(",")


def test_run_columns_are_the_reference_columns():
    assert sweep.RUN_HEADER == REFERENCE_RUN_COLUMNS
    assert sweep.PARTITION_HEADER == ["device_index", "max_streams", "hash", "stream", "depth", "buckets",
                                      "elements", "vector_bytes", "kernel", "runtime", "gb_p_s", "tuples_p_s"]


def test_configuration_product_and_order():
    setup = sweep.load_profile(PROFILE, "tiny")
    configs = sweep.expand(setup)
    # 2 elements x 2 distributions (benchmark) x 2 column counts (join) x 2 algorithms (hash) x 1 probe
    assert len(configs) == 16
    # the benchmark loop is outermost and distribution its innermost field; the hash algorithm varies
    # fastest of all but the probe fields
    assert [c[0]["elements"] for c in configs[:8]] == [20000] * 8
    assert [c[0]["distribution"] for c in configs[:8]] == ["uniform"] * 4 + ["zipf"] * 4
    assert [c[2]["algorithm"] for c in configs[:4]] == ["fnv", "custom_xor_hw", "fnv", "custom_xor_hw"]
    assert [c[1]["rs_columns"] for c in configs[:4]] == [1, 1, 3, 3]


def test_missing_field_is_named():
    with pytest.raises(sweep.ProfileError, match="missing_field.gpus not found"):
        sweep.load_profile(PROFILE, "missing_field")
    with pytest.raises(sweep.ProfileError, match="not found"):
        sweep.load_profile(PROFILE, "no_such_profile")


def test_generated_data_follows_the_reference_distributions():
    import numpy as np
    rng = np.random.default_rng(0)
    keys, pks, mx = sweep.generate(50000, 3, 1000, 0.5, "uniform", rng)
    assert keys.shape == (50000, 3) and (keys == keys[:, :1]).all()
    assert pks[0] == 1 and pks[-1] == 50000
    assert keys.min() >= 600 and keys.max() < 1600 and mx == 1000  # offset ceil((0.5 + 0.1) * 1000)
    z, _, zmax = sweep.generate(50000, 1, 10**6, 1.0, "zipf", rng)
    assert zmax == 10000 and z.min() >= 1 and z.max() <= 10000
    assert (z == 1).mean() > 0.05  # the head of a Zipf(1) over 10^4 ranks holds about 10 % of the draws


def test_dry_run_lists_configurations():
    p = subprocess.run([sys.executable, SWEEP, PROFILE, "tiny", "--dry-run"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip().splitlines()[-1] == "16 configurations"


@pytest.mark.gpu
def test_sweep_runs_and_verifies_every_configuration(tmp_path):
    out = str(tmp_path / "tiny")
    assert sweep.main([PROFILE, "tiny", "--output", out]) == 0
    rows = list(csv.DictReader(open(os.path.join(out, "run.csv"))))
    assert len(rows) == 16 and list(rows[0].keys()) == REFERENCE_RUN_COLUMNS
    for r in rows:
        assert r["successful"] == "1", r
        assert float(r["join_tuple_p_second"]) > 0 and float(r["probe_tuple_p_second"]) > 0
    zipf = [r for r in rows if r["distribution"] == "zipf"]
    assert all(r["max_column_value"] == "10000" for r in zipf)
    part = list(csv.reader(open(os.path.join(out, "partition.csv"))))
    assert part[0] == sweep.PARTITION_HEADER and len(part) == 1 + 16 * 2 * 3
    # an unknown algorithm name and two GPUs are unsuccessful configurations, not crashes
    out2 = str(tmp_path / "unknown")
    assert sweep.main([PROFILE, "unknown_hash", "--output", out2]) == 0
    rows = list(csv.DictReader(open(os.path.join(out2, "run.csv"))))
    assert [(r["gpus"], r["successful"]) for r in rows] == [("1", "0"), ("1", "1"), ("2", "0"), ("2", "0")]
    # an existing results directory is refused, as the reference does
    assert sweep.main([PROFILE, "tiny", "--output", out]) == 1
