"""Writes tests/golden/hash_vectors.json: known-answer vectors for the hash family.

An independent restatement of the reference hash kernels (src/hash/hash.cu:84-505)
in plain Python integers, written differently from the C oracle on purpose:

* FNV-1a is computed byte-wise over a byte string, the textbook form, and checked
  here against the published vectors FNV-1a("a") and FNV-1a("foobar"). The
  reference's word-wise form equals it whenever every 32-bit word is < 256
  (XOR of a word < 256 only touches the low byte), which the script asserts on
  the small-word rows; rows with large words use the word-wise closed form.
* Rotate-left and bit reversal are done on binary strings (slicing / reversing),
  not with shifts and masks.
* C integer conversions (short truncation, unsigned wrap, C's truncating `%`)
  are written as explicit modular arithmetic.

The `*_shift` kinds depend on shifts the reference leaves undefined; they are
written with the PTX clamp rule (amount taken as unsigned, >= width gives 0),
the same documented assumption the C oracle makes, so their vectors are not an
independent pin (parity unpinned for those five kinds).

Run: python tests/golden/gen_golden.py   (rewrites the JSON next to this file)
"""
import json
import os
import random

M32 = (1 << 32) - 1

# Published FNV-1a test vectors (Noll, "FNV hash" reference page).
PUBLISHED = {
    (32, b""): 0x811C9DC5,
    (32, b"a"): 0xE40C292C,
    (32, b"foobar"): 0xBF9CF968,
    (64, b""): 0xCBF29CE484222325,
    (64, b"a"): 0xAF63DC4C8601EC8C,
    (64, b"foobar"): 0x85944171F73967E8,
}
FNV = {32: (16777619, 2166136261), 64: (1099511628211, 14695981039346656037)}


def fnv1a_bytes(data, bits):
    prime, h = FNV[bits]
    for byte in data:
        h = ((h ^ byte) * prime) % (1 << bits)
    return h


def rotl32_str(x, s):
    b = format(x & M32, "032b")
    s %= 32
    return int(b[s:] + b[:s], 2)


def rev_str(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def to_short(x):
    x &= 0xFFFF
    return x - 0x10000 if x & 0x8000 else x


def to_int32(x):
    x &= M32
    return x - (1 << 32) if x & 0x80000000 else x


def c_rem(a, m):
    """C's % for int operands (truncates toward zero)."""
    r = abs(a) % m
    return r if a >= 0 else -r


def ptx_shl(x, s, bits):
    s &= M32
    return 0 if s >= bits else (x << s) % (1 << bits)


def max_shift(hash_bytes, chunk_bytes):
    return (((hash_bytes - chunk_bytes * 8) % (1 << 64)) % 65536 + 1) % 65536


def values(row, chunk_bits):
    """The reference's per-chunk values of a row of 64-bit columns.

    Returns a list of chunks; each chunk is a list of (value, kind) where kind is
    'u32' for 32-bit words and 'c8' for signed chars."""
    raw = b"".join(int(v).to_bytes(8, "little") for v in row)
    if chunk_bits == 8:
        return [[((b - 256) if b >= 128 else b, "c8")] for b in raw]
    words = [int.from_bytes(raw[i:i + 4], "little") for i in range(0, len(raw), 4)]
    wpc = chunk_bits // 32
    return [[(w, "u32") for w in words[i:i + wpc]] for i in range(0, len(words), wpc)]


def ref_hash(kind, row, bits, chunk_bits):
    mask = (1 << bits) - 1
    chunks = values(row, chunk_bits)
    ms = max_shift(bits // 8, 1 if chunk_bits == 8 else chunk_bits // 8)

    def ht(v):  # conversion of a chunk value to hash_t
        return v % (1 << bits)

    def u32(v):  # conversion to unsigned int
        return v % (1 << 32)

    def offset_step(off, v, k):
        # hash_offset = (hash_offset + v) % get_max_shift(); assigned to short
        if k == "u32":
            return to_short(((off % (1 << 32)) + v) % (1 << 32) % ms)
        return to_short(c_rem(off + v, ms))

    if kind == "fnv":
        prime, h = FNV[bits]
        for ch in chunks:
            for v, _ in ch:
                h = ((h ^ ht(v)) * prime) & mask
        return h
    if kind == "custom_add":
        return sum(ht(v) for ch in chunks for v, _ in ch) & mask
    if kind == "custom_xor":
        h = 0
        for ch in chunks:
            for v, _ in ch:
                h ^= ht(v)
        return h
    if kind == "custom_mult":
        h = 1
        for ch in chunks:
            for v, _ in ch:
                h = (1 + h * ((ht(v) + 1) & mask)) & mask
        return h
    if kind == "custom_xor_hw":
        h, shift = 0, 0
        for ch in chunks:
            for v, _ in ch:
                shift = (shift + v) % (1 << 32)
                h ^= rotl32_str(u32(v), shift)
                h = rev_str(h, bits)
        return h
    if kind == "custom_add_hw":
        h, shift = 0, 0
        for ch in chunks:
            for v, _ in ch:
                h = (h + rotl32_str(u32(v), shift)) & mask
                shift += 1
            h = rev_str(h, bits)
        return h
    if kind == "custom_mult_hw":
        h, shift = 1, 0
        for ch in chunks:
            for v, _ in ch:
                shift = to_int32(shift + v)
                h = (1 + h * ((rotl32_str(u32(v), shift) + 1) & mask)) & mask
                h = rev_str(h, bits)
        return h
    if kind == "custom_mult_xor_hw":
        h, off, op = 0, 0, 0
        for ch in chunks:
            for v, _ in ch:
                c = ht(v)
                if op & 1:
                    h = (1 + h * rotl32_str(u32(c + 1), off)) & mask
                else:
                    h ^= rotl32_str(u32(c), off)
                op += 1
                off = to_short(off + c)
                h = rev_str(h, bits)
        return h
    if kind in ("custom_add_shift", "custom_xor_shift", "custom_mult_shift"):
        h = 1 if kind == "custom_mult_shift" else 0
        off = 0
        for ch in chunks:
            for v, k in ch:
                off = offset_step(off, v, k)
                if kind == "custom_add_shift":
                    h = (h + ptx_shl(ht(v), off, bits)) & mask
                elif kind == "custom_xor_shift":
                    h ^= ptx_shl(ht(v), off, bits)
                else:
                    h = (1 + ptx_shl((h * ((ht(v) + 1) & mask)) & mask, off, bits)) & mask
        return h
    if kind in ("custom_mult_xor_shift", "custom_n_mult_xor_shift"):
        rounds = 2 if kind == "custom_n_mult_xor_shift" else 1
        h, off, op = 0, 0, 0
        for _ in range(rounds):
            for ch in chunks:
                for v, _ in ch:
                    c = ht(v)
                    off = to_short(((off % (1 << bits)) + c) % (1 << bits) % ms)
                    if op & 1:
                        h = (1 + h * ptx_shl((c + 1) & mask, off, bits)) & mask
                    else:
                        h = h ^ ptx_shl(c, off, bits)
                    op += 1
        return h
    raise ValueError(kind)


KINDS = ["fnv", "custom_add", "custom_add_shift", "custom_add_hw", "custom_xor", "custom_xor_shift",
         "custom_xor_hw", "custom_mult", "custom_mult_shift", "custom_mult_hw", "custom_mult_xor_shift",
         "custom_n_mult_xor_shift", "custom_mult_xor_hw"]


def self_check():
    for (bits, data), want in PUBLISHED.items():
        assert fnv1a_bytes(data, bits) == want, (bits, data)
    # word-wise FNV-1a over words < 256 == byte-wise FNV-1a over those bytes
    rng = random.Random(7)
    for _ in range(200):
        ncols = rng.randint(1, 4)
        row = [rng.randrange(256) | (rng.randrange(256) << 32) for _ in range(ncols)]
        words = []
        for v in row:
            words += [v & M32, v >> 32]
        for bits in (32, 64):
            assert ref_hash("fnv", row, bits, 64) == fnv1a_bytes(bytes(words), bits)
    # rotate / reverse string forms agree with their definitions on fixed points
    assert rotl32_str(0x80000001, 1) == 0x00000003
    assert rev_str(1, 32) == 0x80000000 and rev_str(1, 64) == 1 << 63


def main():
    self_check()
    rng = random.Random(20240607)
    cases = []
    special = [[0], [1], [0x61], [0xFFFFFFFF], [0xFFFFFFFFFFFFFFFF], [0x8000000080000000], [0x7F80FF01],
               [1, 2], [0x61, 0x62, 0x63, 0x64], [0xDEADBEEFCAFEBABE, 3]]
    rows = list(special)
    for _ in range(24):
        ncols = rng.choice([1, 2, 3, 4])
        mode = rng.randrange(3)
        if mode == 0:
            row = [rng.randrange(1 << 16) for _ in range(ncols)]
        elif mode == 1:
            row = [rng.randrange(1 << 32) for _ in range(ncols)]
        else:
            row = [rng.randrange(1 << 64) for _ in range(ncols)]
        rows.append(row)
    for row in rows:
        for kind in KINDS:
            for bits in (32, 64):
                for chunk_bits in (8, 32, 64, 128):
                    if (len(row) * 64) % chunk_bits:
                        continue
                    cases.append({"row": [str(v) for v in row], "kind": kind, "hash_bits": bits,
                                  "chunk_bits": chunk_bits, "hash": str(ref_hash(kind, row, bits, chunk_bits))})
    fnv_published = [{"bits": b, "data": d.hex(), "hash": str(h)} for (b, d), h in PUBLISHED.items()]
    out = {"generator": "tests/golden/gen_golden.py", "unpinned_kinds": ["custom_add_shift", "custom_xor_shift",
           "custom_mult_shift", "custom_mult_xor_shift", "custom_n_mult_xor_shift"],
           "fnv_published": fnv_published, "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))
    print("wrote %d cases to %s" % (len(cases), path))


if __name__ == "__main__":
    main()
