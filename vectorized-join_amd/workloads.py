"""Seeded synthetic workloads of the benchmark configurations (numpy, host side).

Every generator is deterministic in its seed (numpy PCG64). Keys are returned
as lists of 1-D arrays (one per key column); row ids are implicit (0-based row
index). The configurations' payload columns are not generated: the join
carries row ids, and a payload is fetched afterwards by row id.

  plain_small   build 2^16 uint32 keys = a permutation of [0, 2^16); probe 2^20
                uint32 keys uniform over [0, 2^17) (about 50 % match)
  large_fk      build 2^24 unique int64 keys = mix64(permutation of [0, 2^24));
                probe 2^27 keys drawn uniformly from the build keys (100 % match)
  multicol_skew build 2^22 rows of 4 uint32 columns, 2^20 distinct keys x 4;
                probe 2^25 rows, 80 % Zipf(s=1) over the distinct keys, 20 %
                keys outside the build key space
  batched       256 int32 joins: build n = floor(2^U(10,20)) keys = permutation
                of [0, n); probe n * U(4,16) keys uniform over [0, 2n)

mix64 is the SplitMix64 finalizer (Steele, Lea, Flood 2014), a bijection on
64-bit integers, so distinct inputs stay distinct.
"""
import numpy as np

SEEDS = {"plain_small": 0x5EED0000, "large_fk": 0x5EED0001, "multicol_skew": 0x5EED0002, "batched": 0x5EED0003,
         "served": 0x5EED0004}


def mix64(x):
    z = np.asarray(x, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


def plain_small(log_r=16, log_s=20, seed=None):
    rng = np.random.default_rng(SEEDS["plain_small"] if seed is None else seed)
    r = rng.permutation(1 << log_r).astype(np.uint32)
    s = rng.integers(0, 1 << (log_r + 1), size=1 << log_s, dtype=np.uint64).astype(np.uint32)
    return [r], [s]


def large_fk(log_r=24, log_s=27, seed=None):
    rng = np.random.default_rng(SEEDS["large_fk"] if seed is None else seed)
    r = mix64(rng.permutation(1 << log_r).astype(np.uint64))
    idx = rng.integers(0, 1 << log_r, size=1 << log_s, dtype=np.int64)
    s = r[idx]
    return [r.view(np.int64)], [s.view(np.int64)]


def multicol_skew(log_distinct=20, dup=4, log_s=25, zipf_s=1.0, miss=0.2, seed=None):
    rng = np.random.default_rng(SEEDS["multicol_skew"] if seed is None else seed)
    nd = 1 << log_distinct
    # distinct 4-column keys: a random bijection of the index into 128 bits, low half never 0xFFFFFFFF
    base = mix64(rng.permutation(nd).astype(np.uint64) + np.uint64(1))
    base2 = mix64(base ^ np.uint64(0x9E3779B97F4A7C15))
    cols = [(base & np.uint64(0x7FFFFFFF)).astype(np.uint32), (base >> np.uint64(32)).astype(np.uint32),
            (base2 & np.uint64(0xFFFFFFFF)).astype(np.uint32), (base2 >> np.uint64(32)).astype(np.uint32)]
    order = rng.permutation(nd * dup) % nd
    r = [c[order] for c in cols]
    ns = 1 << log_s
    n_hit = int(round(ns * (1.0 - miss)))
    ranks = np.arange(1, nd + 1, dtype=np.float64)
    cdf = np.cumsum(1.0 / ranks ** zipf_s)
    cdf /= cdf[-1]
    rank_of = np.searchsorted(cdf, rng.random(n_hit), side="right")
    rank_of = np.minimum(rank_of, nd - 1)
    key_of_rank = rng.permutation(nd)
    hit = key_of_rank[rank_of]
    # misses: first column has the top bit set, outside every build key
    miss_cols = [rng.integers(0, 2**32, size=ns - n_hit, dtype=np.uint64).astype(np.uint32) for _ in range(4)]
    miss_cols[0] |= np.uint32(0x80000000)
    s = [np.concatenate([c[hit], m]) for c, m in zip(cols, miss_cols)]
    perm = rng.permutation(ns)
    s = [c[perm] for c in s]
    return r, s


def batched(n_joins=256, log_min=10, log_max=20, fmin=4.0, fmax=16.0, seed=None):
    """List of (build_cols, probe_cols) for independent int32 joins."""
    rng = np.random.default_rng(SEEDS["batched"] if seed is None else seed)
    out = []
    for j in range(n_joins):
        n = int(2.0 ** rng.uniform(log_min, log_max))
        m = int(n * rng.uniform(fmin, fmax))
        jr = np.random.default_rng([SEEDS["batched"] if seed is None else seed, j])
        r = jr.permutation(n).astype(np.int32)
        s = jr.integers(0, 2 * n, size=m, dtype=np.int64).astype(np.int32)
        out.append(([r], [s]))
    return out


def to_device(cols):
    import torch
    out = []
    for c in cols:
        c = np.ascontiguousarray(c)
        if c.dtype == np.uint32:
            c = c.view(np.int32)
        elif c.dtype == np.uint64:
            c = c.view(np.int64)
        out.append(torch.from_numpy(c).cuda())
    return out


def reference_rows(cols):
    """Row-major uint64 buffer of the reference layout (4-byte keys zero-extended)."""
    n = len(cols[0])
    rows = np.empty((n, len(cols)), dtype=np.uint64)
    for i, c in enumerate(cols):
        c = np.asarray(c)
        rows[:, i] = c.view(np.uint32).astype(np.uint64) if c.dtype.itemsize == 4 else c.view(np.uint64)
    return rows
