"""Pins the CPU oracle (oracle/vj_oracle.c) before it is trusted as the checker.

* Hashes: the committed known-answer vectors (tests/golden/hash_vectors.json,
  an independent Python restatement), the published FNV-1a vectors through the
  word/byte equivalence, and invariants of the reference's chunk walk.
* Radix partition: histogram/offset/permutation invariants and brute force.
* Join: the reference's nested loop (join_cpu) on small inputs and an
  independent numpy sort-join on larger ones, over duplicates, multi-column
  keys, empty sides and every pass count.
"""
import json
import os

import numpy as np
import pytest

import oracle
from joinref import fnv1a_bytes, np_join, sorted_pairs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hash_vectors.json")
PINNED_KINDS = [k for k in oracle.HASH_KINDS if k not in oracle.UNPINNED_KINDS]


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("kind", oracle.HASH_KINDS)
def test_oracle_matches_golden_vectors(golden, kind):
    cases = [c for c in golden["cases"] if c["kind"] == kind]
    assert cases
    for c in cases:
        row = np.array([[int(v) for v in c["row"]]], dtype=np.uint64)
        got = int(oracle.hash_rows(row, kind, c["hash_bits"], c["chunk_bits"])[0])
        assert got == int(c["hash"]), (c, hex(got))


def test_fnv_textbook_vectors_in_golden_file(golden):
    for v in golden["fnv_published"]:
        assert fnv1a_bytes(bytes.fromhex(v["data"]), v["bits"]) == int(v["hash"])


@pytest.mark.parametrize("bits", [32, 64])
def test_oracle_fnv_equals_textbook_bytewise(bits):
    # Published vectors pin the textbook form; word-wise FNV-1a over words < 256
    # equals byte-wise FNV-1a over those bytes.
    assert fnv1a_bytes(b"a", 32) == 0xE40C292C and fnv1a_bytes(b"a", 64) == 0xAF63DC4C8601EC8C
    assert fnv1a_bytes(b"foobar", 32) == 0xBF9CF968 and fnv1a_bytes(b"foobar", 64) == 0x85944171F73967E8
    rng = np.random.default_rng(1)
    for ncols in (1, 2, 3, 4):
        lo = rng.integers(0, 256, size=(64, ncols), dtype=np.uint64)
        hi = rng.integers(0, 256, size=(64, ncols), dtype=np.uint64)
        rows = lo | (hi << np.uint64(32))
        for chunk in (32, 64, 128):
            if (ncols * 64) % chunk:
                continue
            got = oracle.hash_rows(rows, "fnv", bits, chunk)
            for i in range(len(rows)):
                words = []
                for c in range(ncols):
                    words += [int(lo[i, c]), int(hi[i, c])]
                assert int(got[i]) == fnv1a_bytes(bytes(words), bits)


def test_fnv_word_equals_bytes_of_a():
    # the row (0x61) is the words (0x61, 0): FNV-1a over bytes "a\0"
    row = np.array([[0x61]], dtype=np.uint64)
    assert int(oracle.hash_rows(row, "fnv", 64, 64)[0]) == fnv1a_bytes(b"a\x00", 64)
    assert int(oracle.hash_rows(row, "fnv", 32, 64)[0]) == fnv1a_bytes(b"a\x00", 32)


@pytest.mark.parametrize("kind", ["fnv", "custom_add", "custom_xor", "custom_mult", "custom_xor_hw",
                                  "custom_mult_hw", "custom_mult_xor_hw"])
def test_chunk_width_invariance(kind):
    # 32/64/128-bit chunks feed the same word sequence to per-word kinds.
    rng = np.random.default_rng(2)
    rows = rng.integers(0, 2**63, size=(200, 2), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    for bits in (32, 64):
        a = oracle.hash_rows(rows, kind, bits, 32)
        assert np.array_equal(a, oracle.hash_rows(rows, kind, bits, 64))
        assert np.array_equal(a, oracle.hash_rows(rows, kind, bits, 128))


def test_add_hw_reverses_per_chunk():
    # custom_add_hw reverses once per chunk (hash.cu:199): widths differ.
    rows = np.array([[0x0000000200000001]], dtype=np.uint64)
    a32 = int(oracle.hash_rows(rows, "custom_add_hw", 64, 32)[0])
    a64 = int(oracle.hash_rows(rows, "custom_add_hw", 64, 64)[0])
    rev = lambda x: int(format(x, "064b")[::-1], 2)
    rot = lambda v, s: int((format(v, "032b")[s:] + format(v, "032b")[:s]), 2)
    assert a32 == rev((rev(1) + rot(2, 1)) % 2**64)
    assert a64 == rev(1 + rot(2, 1))
    assert a32 != a64


def test_xor_and_add_closed_forms():
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 2**64 - 1, size=(100, 3), dtype=np.uint64)
    words = np.concatenate([rows & np.uint64(0xFFFFFFFF), rows >> np.uint64(32)], axis=1)
    for bits in (32, 64):
        x = np.bitwise_xor.reduce(words, axis=1)
        assert np.array_equal(oracle.hash_rows(rows, "custom_xor", bits, 64), x)
        s = (words.astype(object).sum(axis=1) % (1 << bits)).astype(np.uint64)
        assert np.array_equal(oracle.hash_rows(rows, "custom_add", bits, 64), s)


def test_signed_char_chunks_sign_extend():
    # HASH_CHUNK_BITS 8 uses char (types.hpp:16): byte 0x80 XORs as 0xFF..80.
    row = np.array([[0x80]], dtype=np.uint64)
    got = int(oracle.hash_rows(row, "custom_xor", 64, 8)[0])
    assert got == 0xFFFFFFFFFFFFFF80


def test_oracle_rejects_bad_hash_arguments():
    row = np.zeros((1, 1), dtype=np.uint64)
    with pytest.raises(ValueError):
        oracle.hash_rows(row, 13, 64, 64)
    with pytest.raises(ValueError):
        oracle.hash_rows(row, "fnv", 48, 64)
    with pytest.raises(ValueError):
        oracle.hash_rows(row, "fnv", 64, 128)  # 8-byte row is not a whole 16-byte chunk
    with pytest.raises(ValueError):
        oracle.kind_id("xor_hw")  # dead name in the reference TOMLs


@pytest.mark.parametrize("bits,shift", [(1, 0), (4, 3), (8, 0), (8, 8), (7, 57)])
def test_radix_partition_invariants(bits, shift):
    rng = np.random.default_rng(bits * 100 + shift)
    h = rng.integers(0, 2**64 - 1, size=5000, dtype=np.uint64)
    hist, off, perm = oracle.radix_partition(h, shift, bits)
    d = (h >> np.uint64(shift)) & np.uint64((1 << bits) - 1)
    assert np.array_equal(hist, np.bincount(d.astype(np.int64), minlength=1 << bits))
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(hist)[:-1]]))
    assert np.array_equal(np.sort(perm), np.arange(len(h)))
    dp = d[perm]
    assert np.all(np.diff(dp.astype(np.int64)) >= 0)
    for b in range(1 << bits):  # stable inside a digit
        seg = perm[off[b]:off[b] + hist[b]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)


def _tables(rng, nr, ns, ncols, key_space, dup=1):
    base = rng.integers(0, key_space, size=(nr // dup, ncols), dtype=np.uint64)
    r = np.repeat(base, dup, axis=0)
    r = r[rng.permutation(len(r))]
    s = rng.integers(0, key_space, size=(ns, ncols), dtype=np.uint64)
    return r, s


@pytest.mark.parametrize("ncols", [1, 2, 4])
@pytest.mark.parametrize("passes,b1,b2", [(0, 0, 0), (1, 3, 0), (2, 2, 3)])
def test_oracle_join_equals_brute_force(ncols, passes, b1, b2):
    rng = np.random.default_rng(ncols * 10 + passes)
    r, s = _tables(rng, 300, 700, ncols, 40 if ncols == 1 else 4, dup=3)
    rr = rng.permutation(1000)[:len(r)].astype(np.uint32)
    got = oracle.join(r, s, passes=passes, bits1=b1, bits2=b2, r_rowids=rr)
    want = oracle.join_brute(r, s, r_rowids=rr)
    assert len(got[0]) == len(want[0]) > 0
    g, w = sorted_pairs(*got), sorted_pairs(*want)
    assert np.array_equal(g[0], w[0]) and np.array_equal(g[1], w[1])


def test_brute_force_order_is_s_major():
    r = np.array([[5], [7], [5]], dtype=np.uint64)
    s = np.array([[5], [9], [7]], dtype=np.uint64)
    rr, ss = oracle.join_brute(r, s)
    assert list(rr) == [0, 2, 1] and list(ss) == [0, 0, 2]


@pytest.mark.parametrize("kind,hash_bits", [("fnv", 64), ("fnv", 32), ("custom_xor_hw", 64), ("custom_add", 32)])
def test_oracle_join_equals_numpy_join(kind, hash_bits):
    rng = np.random.default_rng(11)
    r, s = _tables(rng, 20000, 100000, 1, 30000)
    got = sorted_pairs(*oracle.join(r, s, kind=kind, hash_bits=hash_bits, passes=2, bits1=5, bits2=4))
    want = np_join(r, s)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_oracle_join_multicolumn_skew():
    rng = np.random.default_rng(12)
    distinct = rng.integers(0, 2**32, size=(500, 4), dtype=np.uint64)
    r = np.repeat(distinct, 4, axis=0)
    ranks = np.minimum(rng.zipf(1.5, size=20000) - 1, 499)
    s = distinct[ranks]
    got = sorted_pairs(*oracle.join(r, s, passes=1, bits1=6))
    want = np_join(r, s)
    assert len(want[0]) == 4 * 20000
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_oracle_join_empty_and_disjoint():
    r = np.arange(10, dtype=np.uint64).reshape(-1, 1)
    s = np.arange(10, 20, dtype=np.uint64).reshape(-1, 1)
    assert len(oracle.join(r, s)[0]) == 0
    assert len(oracle.join(r[:0], s)[0]) == 0
    assert len(oracle.join(r, s[:0])[0]) == 0
