// Device-side building blocks shared by every kernel of the library:
// key columns in HBM, the 32-bit word stream of a key, and the hash family.
//
// Hash family: the 13 kernels of src/hash/hash.cu:84-505, restated per value
// with the reference's own integer types (hash_t, short, unsigned short, int,
// uint32_t, char), so the C++ promotions are the reference's. Two places need
// care on gfx950:
//  * `x << s` with s >= width or s < 0 is undefined in C++; the reference's
//    nvcc build gets PTX shl semantics (shift taken as unsigned, clamped to
//    the width -> 0). AMD's v_lshlrev masks the amount instead, so the clamp
//    is written out (shl_clamp).
//  * __funnelshift_l(v, v, s) (hash.cu:72-74) is a 32-bit rotate by s & 31.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vj {

enum HashKind : int {
    HK_FNV = 0,
    HK_ADD = 1,
    HK_ADD_SHIFT = 2,
    HK_ADD_HW = 3,
    HK_XOR = 4,
    HK_XOR_SHIFT = 5,
    HK_XOR_HW = 6,
    HK_MULT = 7,
    HK_MULT_SHIFT = 8,
    HK_MULT_HW = 9,
    HK_MULT_XOR_SHIFT = 10,
    HK_N_MULT_XOR_SHIFT = 11,
    HK_MULT_XOR_HW = 12,
    HK_COUNT = 13
};

// Kernel specialisations of the hash: the two kinds the join configurations
// use get straight-line code; everything else goes through the generic walker.
enum HashPath : int { HP_FNV = 0, HP_XOR_HW = 1, HP_GENERIC = 2 };

struct HashCfg {
    int kind;
    int bits;        // 32 / 64
    int chunk_bits;  // 8 / 32 / 64 / 128
    int native;      // word mode: 0 reference (columns widened to 64 bit), 1 native
    int key_signed;  // 4-byte keys sign-extended when widened
};

// A table in HBM: SoA key columns and optional uint32 row ids.
struct TView {
    const void *cols[4];
    const uint32_t *rowids;  // nullptr: row id = index
    uint64_t n;
};

template <int KW> struct KeyT;
template <> struct KeyT<4> { using T = uint32_t; };
template <> struct KeyT<8> { using T = uint64_t; };

template <int KW, int NC> struct Key {
    typename KeyT<KW>::T c[NC];
};

template <int KW, int NC>
__device__ __forceinline__ Key<KW, NC> load_key(const TView &t, uint64_t i) {
    using T = typename KeyT<KW>::T;
    Key<KW, NC> k;
#pragma unroll
    for (int c = 0; c < NC; c++) k.c[c] = static_cast<const T *>(t.cols[c])[i];
    return k;
}

// Pointers in the global address space. Memory operations through them
// compile to global_* instructions, which count only on vmcnt; through
// generic pointers they compile to flat_* instructions, which also count on
// lgkmcnt, so every wait for an LDS result would also wait for the loads in
// flight (and the compiler must assume a flat store may alias LDS).
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<T> to_global(T *p) { return (gptr<T>)p; }

// A table's key columns as global pointers (copied out of a TView once, so a
// loop does not re-read them from its descriptor).
template <int KW, int NC> struct GKeys {
    gptr<const typename KeyT<KW>::T> c[NC];
};

// A TView with global pointers.
struct GView {
    gptr<const void> cols[4];
    gptr<const uint32_t> rowids;  // null: row id = index
};

template <int KW, int NC>
__device__ __forceinline__ Key<KW, NC> load_key(const GView &t, uint64_t i) {
    using T = typename KeyT<KW>::T;
    Key<KW, NC> k;
#pragma unroll
    for (int c = 0; c < NC; c++) k.c[c] = static_cast<gptr<const T>>(t.cols[c])[i];
    return k;
}

template <int KW, int NC>
__device__ __forceinline__ Key<KW, NC> load_key(const GKeys<KW, NC> &g, uint64_t i) {
    Key<KW, NC> k;
#pragma unroll
    for (int c = 0; c < NC; c++) k.c[c] = g.c[c][i];
    return k;
}

template <int KW, int NC>
__device__ __forceinline__ bool key_eq(const Key<KW, NC> &a, const Key<KW, NC> &b) {
    bool eq = true;
#pragma unroll
    for (int c = 0; c < NC; c++) eq &= (a.c[c] == b.c[c]);
    return eq;
}

// The key's 32-bit words in reference order: per column low then high word
// of its 64-bit column_t value (types.hpp:7, hash.cu:99-106). Returns count.
template <int KW, int NC>
__device__ __forceinline__ int key_words(const Key<KW, NC> &k, const HashCfg &h, uint32_t (&w)[2 * NC]) {
    int n = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (KW == 8) {
            uint64_t v = (uint64_t)k.c[c];
            w[n++] = (uint32_t)v;
            w[n++] = (uint32_t)(v >> 32);
        } else {
            uint32_t v = (uint32_t)k.c[c];
            w[n++] = v;
            if (!h.native) w[n++] = h.key_signed ? (uint32_t)((int32_t)v >> 31) : 0u;
        }
    }
    return n;
}

// ---------------------------------------------------------------------------
// Primitives (hash.cu:44-74)
// ---------------------------------------------------------------------------

template <typename HT> __device__ __forceinline__ HT bit_reverse(HT x);
template <> __device__ __forceinline__ uint32_t bit_reverse<uint32_t>(uint32_t x) { return __builtin_bitreverse32(x); }
template <> __device__ __forceinline__ uint64_t bit_reverse<uint64_t>(uint64_t x) { return __builtin_bitreverse64(x); }

__device__ __forceinline__ uint32_t permutate(uint32_t value, int shift) {
    return __builtin_rotateleft32(value, (uint32_t)shift & 31u);
}

template <typename HT> __device__ __forceinline__ HT shl_clamp(HT x, int s) {
    return (uint32_t)s >= (uint32_t)(sizeof(HT) * 8) ? (HT)0 : (HT)(x << (uint32_t)s);
}

template <typename HT> struct Fnv;
template <> struct Fnv<uint32_t> {
    static constexpr uint32_t prime = 16777619u, offset = 2166136261u;
    static constexpr uint32_t prime2 = prime * prime;  // mod 2^32
};
template <> struct Fnv<uint64_t> {
    static constexpr uint64_t prime = 1099511628211ull, offset = 14695981039346656037ull;
    static constexpr uint64_t prime2 = prime * prime;  // mod 2^64
};

// get_max_shift(), hash.cu:60-62, with sizeof(chunk_t) from chunk_bits.
template <typename HT> __device__ __forceinline__ unsigned short max_shift(int chunk_bits) {
    size_t chunk_bytes = chunk_bits == 8 ? 1 : (size_t)chunk_bits / 8;
    return (unsigned short)((int)(unsigned short)(sizeof(HT) - chunk_bytes * 8) + 1);
}

// ---------------------------------------------------------------------------
// Fast paths
// ---------------------------------------------------------------------------

// hash_fnv (hash.cu:84-116) over the word stream; the same for every chunk width >= 32.
template <typename HT, int NW> __device__ __forceinline__ HT fnv_words(const uint32_t (&w)[NW], int n) {
    HT h = Fnv<HT>::offset;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        if (i < n) {
            h ^= w[i];
            h *= Fnv<HT>::prime;
        }
    }
    return h;
}

// hash_custom_xor_hw (hash.cu:264-299), chunk width >= 32.
template <typename HT, int NW> __device__ __forceinline__ HT xor_hw_words(const uint32_t (&w)[NW], int n) {
    HT h = 0;
    uint32_t shift = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        if (i < n) {
            shift += w[i];
            h ^= (HT)permutate(w[i], (int)shift);
            h = bit_reverse<HT>(h);
        }
    }
    return h;
}

// ---------------------------------------------------------------------------
// Generic walker: every kind, every chunk width (including signed 8-bit chunks)
// ---------------------------------------------------------------------------

template <typename HT> struct Walker {
    int kind;
    unsigned short ms;
    HT hash;
    short hash_offset;
    short operation;
    uint32_t ushift;  // custom_xor_hw's uint32_t shift
    int ishift;       // custom_add_hw / custom_mult_hw's int shift

    __device__ __forceinline__ void init(int k, int chunk_bits) {
        kind = k;
        ms = max_shift<HT>(chunk_bits);
        hash_offset = 0;
        operation = 0;
        ushift = 0;
        ishift = 0;
        if (k == HK_FNV) hash = Fnv<HT>::offset;
        else if (k == HK_MULT || k == HK_MULT_SHIFT || k == HK_MULT_HW) hash = 1;
        else hash = 0;
    }

    // One chunk value; V is uint32_t (a word) or signed char (an 8-bit chunk).
    template <typename V> __device__ __forceinline__ void step(V v) {
        typedef HT hash_t;
        switch (kind) {
        case HK_FNV:  // hash.cu:96-111
            hash ^= v;
            hash *= Fnv<HT>::prime;
            break;
        case HK_ADD:  // hash.cu:128-138
            hash += ((hash_t)v);
            break;
        case HK_ADD_SHIFT:  // hash.cu:155-170
            hash_offset = (hash_offset + v) % ms;
            hash += shl_clamp<HT>(((hash_t)v), hash_offset);
            break;
        case HK_ADD_HW:  // hash.cu:187-197 (reverse per chunk in chunk_end)
            hash += (hash_t)permutate((uint32_t)v, ishift++);
            break;
        case HK_XOR:  // hash.cu:215-225
            hash ^= ((hash_t)v);
            break;
        case HK_XOR_SHIFT:  // hash.cu:242-257
            hash_offset = (hash_offset + v) % ms;
            hash ^= shl_clamp<HT>(((hash_t)v), hash_offset);
            break;
        case HK_XOR_HW:  // hash.cu:274-294
            ushift += v;
            hash ^= (hash_t)permutate((uint32_t)v, (int)ushift);
            hash = bit_reverse<HT>(hash);
            break;
        case HK_MULT:  // hash.cu:311-321
            hash = 1 + (hash * ((hash_t)v + 1));
            break;
        case HK_MULT_SHIFT:  // hash.cu:339-354, (hash * (w + 1)) << off as written
            hash_offset = (hash_offset + v) % ms;
            hash = 1 + shl_clamp<HT>((hash_t)(hash * ((hash_t)v + 1)), hash_offset);
            break;
        case HK_MULT_HW:  // hash.cu:371-391
            ishift += v;
            hash = 1 + (hash * ((hash_t)permutate((uint32_t)v, ishift) + 1));
            hash = bit_reverse<HT>(hash);
            break;
        case HK_MULT_XOR_SHIFT:
        case HK_N_MULT_XOR_SHIFT: {  // mult_xor_shift, hash.cu:398-406
            hash_t chunk = (hash_t)v;
            hash_offset = (hash_offset + chunk) % ms;
            if (operation & 1) hash = 1 + (hash * shl_clamp<HT>((hash_t)(chunk + 1), hash_offset));
            else hash = (hash ^ shl_clamp<HT>((hash_t)chunk, hash_offset));
            operation++;
            break;
        }
        case HK_MULT_XOR_HW: {  // mult_xor_hw, hash.cu:408-417
            hash_t chunk = (hash_t)v;
            if (operation & 1) hash = 1 + (hash * ((hash_t)permutate((uint32_t)(chunk + 1), hash_offset)));
            else hash = (hash ^ ((hash_t)permutate((uint32_t)chunk, hash_offset)));
            operation++;
            hash_offset += chunk;
            hash = bit_reverse<HT>(hash);
            break;
        }
        default:
            break;
        }
    }

    __device__ __forceinline__ void chunk_end() {
        if (kind == HK_ADD_HW) hash = bit_reverse<HT>(hash);  // hash.cu:199
    }
};

template <typename HT, int NW>
__device__ __noinline__ HT generic_words(const uint32_t (&w)[NW], int n, int kind, int chunk_bits) {
    Walker<HT> s;
    s.init(kind, chunk_bits);
    const int rounds = kind == HK_N_MULT_XOR_SHIFT ? 2 : 1;  // hash_func passes rounds = 2, hash.cu:526
    for (int r = 0; r < rounds; r++) {
        if (chunk_bits == 8) {
            for (int i = 0; i < n; i++) {
                uint32_t word = w[i];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    s.step((signed char)(word >> (8 * b)));
                    s.chunk_end();
                }
            }
        } else {
            const int wpc = chunk_bits / 32;
            for (int i = 0; i < n; i += wpc) {
                for (int j = 0; j < wpc; j++) s.step(w[i + j]);
                s.chunk_end();
            }
        }
    }
    return s.hash;
}

// The same two fast paths for 4-byte key columns widened by zero extension
// (reference words, unsigned keys): every column contributes the words (k, 0).
// FNV-1a: ((h ^ k) * P ^ 0) * P = (h ^ k) * P^2 (mod 2^bits), one multiply per
// column. custom_xor_hw: the zero word adds 0 to the shift, xors in
// rotl(0, s) = 0 and reverses the bits again, undoing the reversal of the
// column's own word, so each column is h ^= rotl(k, s) with s += k. Both equal
// the word-by-word results exactly.
template <typename HT, int NC> __device__ __forceinline__ HT fnv_zext_cols(const uint32_t (&k)[NC]) {
    HT h = Fnv<HT>::offset;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        h ^= k[c];
        h *= Fnv<HT>::prime2;
    }
    return h;
}
template <typename HT, int NC> __device__ __forceinline__ HT xor_hw_zext_cols(const uint32_t (&k)[NC]) {
    HT h = 0;
    uint32_t shift = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        shift += k[c];
        h ^= (HT)permutate(k[c], (int)shift);
    }
    return h;
}

// FNV-1a with the given constants in 32-bit arithmetic. With the truncated
// 64-bit offset basis and prime this is the low half of FNV-1a-64 (a product's
// low 32 bits depend only on its factors' low 32 bits, and every word is 32
// bits wide), at one 32-bit multiply per word instead of a 64-bit one.
template <int NW>
__device__ __forceinline__ uint32_t fnv_words_lo(const uint32_t (&w)[NW], int n, uint32_t off, uint32_t prime) {
    uint32_t x = off;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        if (i < n) {
            x ^= w[i];
            x *= prime;
        }
    }
    return x;
}
template <int NC>
__device__ __forceinline__ uint32_t fnv_zext_cols_lo(const uint32_t (&k)[NC], uint32_t off, uint32_t prime2) {
    uint32_t x = off;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        x ^= k[c];
        x *= prime2;
    }
    return x;
}

// Hash of a key as a uint64 (32-bit hashes zero-extended).
template <int HP, int KW, int NC>
__device__ __forceinline__ uint64_t hash_key(const Key<KW, NC> &k, const HashCfg &h) {
    // 4-byte keys on the fast paths: the word count is fixed by the (wave-uniform)
    // word mode, so the words are not assembled at run-time positions
    if (KW == 4 && HP != HP_GENERIC && !h.key_signed) {
        uint32_t c[NC];
#pragma unroll
        for (int i = 0; i < NC; i++) c[i] = (uint32_t)k.c[i];
        if (!h.native) {  // reference words (k, 0) per column
            if (HP == HP_FNV)
                return h.bits == 32 ? (uint64_t)fnv_zext_cols<uint32_t, NC>(c) : fnv_zext_cols<uint64_t, NC>(c);
            return h.bits == 32 ? (uint64_t)xor_hw_zext_cols<uint32_t, NC>(c) : xor_hw_zext_cols<uint64_t, NC>(c);
        }
        // native words: one per column
        if (HP == HP_FNV)
            return h.bits == 32 ? (uint64_t)fnv_words<uint32_t, NC>(c, NC) : fnv_words<uint64_t, NC>(c, NC);
        return h.bits == 32 ? (uint64_t)xor_hw_words<uint32_t, NC>(c, NC) : xor_hw_words<uint64_t, NC>(c, NC);
    }
    uint32_t w[2 * NC];
    int n = key_words<KW, NC>(k, h, w);
    if (HP == HP_FNV) {
        return h.bits == 32 ? (uint64_t)fnv_words<uint32_t, 2 * NC>(w, n) : fnv_words<uint64_t, 2 * NC>(w, n);
    } else if (HP == HP_XOR_HW) {
        return h.bits == 32 ? (uint64_t)xor_hw_words<uint32_t, 2 * NC>(w, n) : xor_hw_words<uint64_t, 2 * NC>(w, n);
    } else {
        return h.bits == 32 ? (uint64_t)generic_words<uint32_t, 2 * NC>(w, n, h.kind, h.chunk_bits)
                            : generic_words<uint64_t, 2 * NC>(w, n, h.kind, h.chunk_bits);
    }
}

// The low 32 bits of hash_key (for digits and table places below bit 32):
// FNV-1a in 32-bit arithmetic, the other kinds by truncation.
template <int HP, int KW, int NC>
__device__ __forceinline__ uint32_t hash_key_lo(const Key<KW, NC> &k, const HashCfg &h) {
    if constexpr (HP == HP_FNV) {
        const bool b32 = h.bits == 32;
        const uint32_t off = b32 ? Fnv<uint32_t>::offset : (uint32_t)Fnv<uint64_t>::offset;
        const uint32_t prime = b32 ? Fnv<uint32_t>::prime : (uint32_t)Fnv<uint64_t>::prime;
        if (KW == 4 && !h.key_signed) {
            uint32_t c[NC];
#pragma unroll
            for (int i = 0; i < NC; i++) c[i] = (uint32_t)k.c[i];
            if (!h.native) return fnv_zext_cols_lo<NC>(c, off, b32 ? Fnv<uint32_t>::prime2 : (uint32_t)Fnv<uint64_t>::prime2);
            return fnv_words_lo<NC>(c, NC, off, prime);
        }
        uint32_t w[2 * NC];
        const int n = key_words<KW, NC>(k, h, w);
        return fnv_words_lo<2 * NC>(w, n, off, prime);
    } else {
        return (uint32_t)hash_key<HP, KW, NC>(k, h);
    }
}

// Kernel-path selection for a hash configuration.
inline int hash_path(const HashCfg &h) {
    if (h.chunk_bits != 8 && h.kind == HK_FNV) return HP_FNV;
    if (h.chunk_bits != 8 && h.kind == HK_XOR_HW) return HP_XOR_HW;
    return HP_GENERIC;
}

}  // namespace vj
