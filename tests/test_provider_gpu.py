"""The reference-layout join (vj_join_rows) and the JoinProvider facade
(include/vj/join_provider.hpp) against the oracle.

A caller written like the reference's join benchmark
(src/join_benchmark.cu:209-275), compiled with hipcc against the facade
(vectorized-join_amd/examples/reference_caller), joins row-major uint64
tables with arbitrary 64-bit primary keys on the device and on the host. Its
rs tables must hold exactly the oracle's pairs mapped to primary keys
((r pk, s pk) per row, primary keys 1..M), the reference's comparison rule
(verify, src/join_benchmark.cu:97-148: per r pk, the sorted list of s pks).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import vjoin
import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "vectorized-join_amd", "examples", "reference_caller")


def _tables(nr, ns, cols, distinct, seed):
    """Row-major uint64 keys with duplicates on both sides, some probe misses, and
    primary keys that are not the row numbers (so the pk gather is checked)."""
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2**63, size=(distinct, cols), dtype=np.uint64)
    r = keys[rng.integers(0, distinct, size=nr)]
    s = keys[rng.integers(0, distinct, size=ns)]
    s[::4, cols - 1] ^= np.uint64(1)  # misses
    r_pk = workloads.mix64(np.arange(nr, dtype=np.uint64) + np.uint64(seed * 7919 + 1))
    s_pk = workloads.mix64(np.arange(ns, dtype=np.uint64) + np.uint64(seed * 104729 + 2))
    return np.ascontiguousarray(r), r_pk, np.ascontiguousarray(s), s_pk


def _expected(r, r_pk, s, s_pk, kind, bits):
    ri, si = oracle.join(r, s, kind, 64, 64, passes=1, bits1=bits)
    return _sorted_pk_pairs(r_pk[ri.astype(np.int64)], s_pk[si.astype(np.int64)])


def _sorted_pk_pairs(a, b):
    order = np.lexsort((b, a))
    return np.asarray(a)[order], np.asarray(b)[order]


def _check_rs(values, pks, want, what):
    values = np.asarray(values, dtype=np.uint64).reshape(-1, 2)
    assert len(values) == len(want[0]), (what, len(values), len(want[0]))
    assert np.array_equal(np.asarray(pks, dtype=np.uint64), np.arange(1, len(values) + 1, dtype=np.uint64)), what
    got = _sorted_pk_pairs(values[:, 0], values[:, 1])
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), what


CASES = [
    # (build rows, probe rows, key columns, distinct keys, buckets, algorithm)
    (1 << 14, 1 << 17, 1, 1 << 13, 256, "fnv"),
    (5000, 40000, 3, 1500, 16, "custom_xor_hw"),
    (3000, 20000, 4, 2500, 64, "custom_mult_hw"),
    (2000, 9000, 2, 1 << 20, 8, "fnv"),  # almost no matches
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(CASES)))
def test_reference_style_caller_matches_oracle(tmp_path, case):
    nr, ns, cols, distinct, buckets, kind = CASES[case]
    r, r_pk, s, s_pk = _tables(nr, ns, cols, distinct, 50 + case)
    d = str(tmp_path)
    for name, a in (("r_values", r), ("r_pks", r_pk), ("s_values", s), ("s_pks", s_pk)):
        a.astype(np.uint64).tofile(os.path.join(d, name + ".bin"))
    p = subprocess.run([CALLER, d, str(cols), str(buckets), kind], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, (p.stdout, p.stderr)
    want = _expected(r, r_pk, s, s_pk, kind, int(np.ceil(np.log2(buckets))))
    for tag in ("rs", "rs_host"):
        vals = np.fromfile(os.path.join(d, tag + "_values.bin"), dtype=np.uint64)
        pks = np.fromfile(os.path.join(d, tag + "_pks.bin"), dtype=np.uint64)
        _check_rs(vals, pks, want, (case, tag))
    summary = dict(line.split(" ", 1) for line in open(os.path.join(d, "summary.txt")).read().splitlines())
    assert int(summary["rs_elements"]) == len(want[0]) and summary["status"] == "Successful"


@pytest.mark.gpu
@pytest.mark.parametrize("cols", [1, 4])
def test_join_rows_abi_matches_oracle(cols):
    import torch
    r, r_pk, s, s_pk = _tables(7000, 50000, cols, 3000, 90 + cols)
    ctx = vjoin.Context(0)
    try:
        dev = [torch.from_numpy(a.view(np.int64)).cuda() for a in (r, r_pk, s, s_pk)]
        vals, pks = ctx.join_rows(dev[0], dev[1], dev[2], dev[3], vjoin.join_opts("fnv", 64, 64, passes=2))
        want = _expected(r, r_pk, s, s_pk, "fnv", 6)
        _check_rs(vals.cpu().numpy().view(np.uint64), pks.cpu().numpy().view(np.uint64), want, cols)
        # no matches: an empty rs table
        s2 = s.copy()
        s2[:, 0] = ~s2[:, 0]
        d2 = torch.from_numpy(s2.view(np.int64)).cuda()
        vals, pks = ctx.join_rows(dev[0], dev[1], d2, dev[3], vjoin.join_opts())
        assert vals.shape[0] == 0 and pks.shape[0] == 0
    finally:
        ctx.close()


def test_reference_caller_reports_failure_without_a_device(tmp_path):
    """On a machine without a usable GPU the facade's join returns a failed
    JoinStatus (reported by the caller), instead of exiting from inside."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the failure path is exercised on CPU-only hosts")
    assert os.path.exists(CALLER), "reference_caller is not built (run __graft_entry__.build())"
    d = str(tmp_path)
    for name, n in (("r_values", 4), ("r_pks", 4), ("s_values", 4), ("s_pks", 4)):
        np.arange(n, dtype=np.uint64).tofile(os.path.join(d, name + ".bin"))
    p = subprocess.run([CALLER, d, "1", "2", "fnv"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "Join failed" in p.stdout, (p.returncode, p.stdout, p.stderr)
