"""GPU parity of vj_hash against the oracle: bit-exact, every kind, hash width,
chunk width, key width and column count, on seeded inputs; plus a large
input checked by a size-independent property."""
import numpy as np
import pytest

import oracle
import vjoin

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).cuda()


def _host(t, bits):
    a = t.cpu().numpy()
    return a.view(np.uint32).astype(np.uint64) if bits == 32 else a.view(np.uint64)


def _random_cols(rng, n, ncols, kw):
    if kw == 4:
        cols = [rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32) for _ in range(ncols)]
        # include values with the top bit set and small values
        for c in cols:
            c[:8] = [0, 1, 0x61, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0x80, 0xFF]
    else:
        cols = [rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64) for _ in range(ncols)]
        for c in cols:
            c[:6] = [0, 1, 0x61, 0x8000000000000000, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF]
    return cols


@pytest.mark.parametrize("kind", vjoin.HASH_KINDS)
@pytest.mark.parametrize("hash_bits", [32, 64])
def test_hash_all_kinds_match_oracle(kind, hash_bits):
    rng = np.random.default_rng(abs(hash(kind)) % 1000 + hash_bits)
    n = 3001
    for kw, ncols in [(8, 1), (4, 1), (4, 2), (8, 3), (4, 4)]:
        cols = _random_cols(rng, n, ncols, kw)
        rows = oracle.reference_rows(cols)
        for chunk_bits in (8, 32, 64, 128):
            if (ncols * 64) % chunk_bits:
                continue
            want = oracle.hash_rows(rows, kind, hash_bits, chunk_bits)
            got = _host(vjoin.hash_table([_dev(c) for c in cols], kind, hash_bits, chunk_bits), hash_bits)
            bad = np.nonzero(got != want)[0]
            assert len(bad) == 0, (kind, hash_bits, kw, ncols, chunk_bits, bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("kind", ["fnv", "custom_xor_hw", "custom_mult_xor_hw"])
def test_hash_signed_4byte_keys(kind):
    rng = np.random.default_rng(5)
    k = rng.integers(-2**31, 2**31, size=2000, dtype=np.int64).astype(np.int32)
    rows = oracle.reference_rows([k.view(np.uint32)], signed=True)
    want = oracle.hash_rows(rows, kind, 64, 64)
    got = _host(vjoin.hash_table([_dev(k)], kind, 64, 64, key_signed=True), 64)
    assert np.array_equal(got, want)


def test_hash_native_words_equal_packed_reference_rows():
    # native mode on 4-byte keys (k0, k1) is the reference layout of one 64-bit column k0 | k1 << 32
    rng = np.random.default_rng(6)
    k0 = rng.integers(0, 2**32, size=4000, dtype=np.uint64).astype(np.uint32)
    k1 = rng.integers(0, 2**32, size=4000, dtype=np.uint64).astype(np.uint32)
    packed = (k0.astype(np.uint64) | (k1.astype(np.uint64) << np.uint64(32))).reshape(-1, 1)
    for kind in ("fnv", "custom_xor_hw", "custom_add_hw"):
        want = oracle.hash_rows(packed, kind, 64, 64)
        got = _host(vjoin.hash_table([_dev(k0), _dev(k1)], kind, 64, 64, word_mode=vjoin.WORDS_NATIVE), 64)
        assert np.array_equal(got, want), kind


def test_hash_large_input_sample_matches_oracle():
    import torch
    n = 1 << 25
    g = torch.Generator(device="cuda").manual_seed(7)
    keys = torch.randint(-2**62, 2**62, (n,), device="cuda", dtype=torch.int64, generator=g)
    h = vjoin.hash_table([keys], "fnv", 64, 64)
    torch.cuda.synchronize()
    idx = torch.randint(0, n, (20000,), device="cuda", generator=g)
    idx[0] = n - 1
    ks = keys[idx].cpu().numpy().view(np.uint64).reshape(-1, 1)
    want = oracle.hash_rows(ks, "fnv", 64, 64)
    assert np.array_equal(h[idx].cpu().numpy().view(np.uint64), want)
