/*
 * vj_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
 *
 * A plain C11 + OpenMP restatement of the reference's hashing and hash-join
 * path (vgerber/vectorized-join). Every function cites the reference lines it
 * follows. The product (the HIP library) never links or calls this file.
 *
 * Integer semantics: the reference kernels are C++ with the types spelled out
 * below (hash_t, short, unsigned short, int, uint32_t, char); C applies the
 * same integer promotions, so each step is written with the same types and
 * yields the same bits. The one place where the reference relies on undefined
 * behaviour is the `*_shift` family (64/32-bit shifts by amounts >= the width
 * or negative, hash.cu:60-62,151-156,398-404). There the PTX `shl` rule
 * ("shift amounts greater than the register width N are clamped to N", so the
 * result is 0) is emulated explicitly; no reference binary is available to
 * confirm what nvcc emitted, so those five kinds are "parity unpinned".
 */
#include "vj_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Hash helpers (src/hash/hash.cu:44-74)                                     */
/* ------------------------------------------------------------------------ */

/* get_fnv_prime / get_fnv_offset, hash.cu:44-58 */
#define FNV_PRIME_32 ((uint32_t)16777619u)
#define FNV_OFFSET_32 ((uint32_t)2166136261u)
#define FNV_PRIME_64 ((uint64_t)1099511628211ull)
#define FNV_OFFSET_64 ((uint64_t)14695981039346656037ull)

static inline uint32_t rd32(const unsigned char *p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

/* reverse(), hash.cu:64-70 (__brev / __brevll) */
static inline uint32_t rev_32(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; i++) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
static inline uint64_t rev_64(uint64_t x) {
    uint64_t r = 0;
    for (int i = 0; i < 64; i++) r |= ((x >> i) & 1ull) << (63 - i);
    return r;
}
#define REV(x) _Generic((x), uint32_t: rev_32, uint64_t: rev_64)(x)

/* permutate(), hash.cu:72-74: __funnelshift_l(v, v, s) == rotate-left by s & 31 */
static inline uint32_t permutate(uint32_t value, int shift) {
    int s = shift & 31;
    return s ? (uint32_t)((value << s) | (value >> (32 - s))) : value;
}

/* PTX shl.b32 / shl.b64 with the shift amount taken as unsigned and clamped
 * to the width (see the file header). */
static inline uint32_t shl_32(uint32_t x, int s) { return (uint32_t)s >= 32u ? 0u : (uint32_t)(x << (uint32_t)s); }
static inline uint64_t shl_64(uint64_t x, int s) { return (uint32_t)s >= 64u ? 0ull : (uint64_t)(x << (uint32_t)s); }
#define SHL(x, s) _Generic((x), uint32_t: shl_32, uint64_t: shl_64)((x), (int)(s))

/* get_max_shift(), hash.cu:60-62: (unsigned short)(sizeof(hash_t) - sizeof(chunk_t) * 8) + 1 */
static inline unsigned short max_shift(size_t hash_bytes, size_t chunk_bytes) {
    return (unsigned short)((int)(unsigned short)(hash_bytes - chunk_bytes * 8) + 1);
}

/*
 * One kind = (state declaration, per-value step, per-chunk end). The row is
 * walked exactly as the reference kernels walk `element_chunks` chunks of a
 * row-major buffer: chunk_t = char (signed) for 8-bit chunks, else uint1 /
 * uint2 / uint4 whose .x .y .z .w are the row's consecutive 32-bit words
 * (types.hpp:15-23, hash.cu:92-113).
 */
#define KIND_FN(W, NAME, DECL, STEP, CHUNK_END, ROUNDS)                                                   \
    static uint##W##_t NAME##_##W(const unsigned char *row, int nbytes, int chunk_bits) {                 \
        typedef uint##W##_t hash_t;                                                                       \
        const unsigned short MS = max_shift(sizeof(hash_t), chunk_bits == 8 ? 1 : (size_t)chunk_bits / 8); \
        (void)MS;                                                                                         \
        DECL;                                                                                             \
        for (int round_ = 0; round_ < (ROUNDS); round_++) {                                               \
            if (chunk_bits == 8) {                                                                        \
                for (int b = 0; b < nbytes; b++) {                                                        \
                    signed char v = (signed char)row[b];                                                  \
                    STEP;                                                                                 \
                    CHUNK_END;                                                                            \
                }                                                                                         \
            } else {                                                                                      \
                int wpc = chunk_bits / 32, nw = nbytes / 4;                                               \
                for (int c = 0; c < nw; c += wpc) {                                                       \
                    for (int j = 0; j < wpc; j++) {                                                       \
                        uint32_t v = rd32(row + 4 * (c + j));                                             \
                        STEP;                                                                             \
                    }                                                                                     \
                    CHUNK_END;                                                                            \
                }                                                                                         \
            }                                                                                             \
        }                                                                                                 \
        return hash;                                                                                      \
    }

#define NOTHING ((void)0)

/* hash_fnv, hash.cu:84-116 */
#define FNV_DECL hash_t hash = (sizeof(hash_t) == 4) ? (hash_t)FNV_OFFSET_32 : (hash_t)FNV_OFFSET_64; \
    const hash_t prime = (sizeof(hash_t) == 4) ? (hash_t)FNV_PRIME_32 : (hash_t)FNV_PRIME_64
#define FNV_STEP hash ^= v; hash *= prime
/* hash_custom_add, hash.cu:118-143 */
#define ADD_DECL hash_t hash = 0
#define ADD_STEP hash += ((hash_t)v)
/* hash_custom_add_shift, hash.cu:145-175 */
#define ADDS_DECL hash_t hash = 0; short hash_offset = 0
#define ADDS_STEP hash_offset = (hash_offset + v) % MS; hash += SHL(((hash_t)v), hash_offset)
/* hash_custom_add_hw, hash.cu:177-203 (reverse once per chunk, :199) */
#define ADDHW_DECL hash_t hash = 0; int shift = 0
#define ADDHW_STEP hash += (hash_t)permutate(v, shift++)
#define ADDHW_END hash = REV(hash)
/* hash_custom_xor, hash.cu:205-230 */
#define XOR_DECL hash_t hash = 0
#define XOR_STEP hash ^= ((hash_t)v)
/* hash_custom_xor_shift, hash.cu:232-262 */
#define XORS_DECL hash_t hash = 0; short hash_offset = 0
#define XORS_STEP hash_offset = (hash_offset + v) % MS; hash ^= SHL(((hash_t)v), hash_offset)
/* hash_custom_xor_hw, hash.cu:264-299 */
#define XORHW_DECL hash_t hash = 0; uint32_t shift = 0
#define XORHW_STEP shift += v; hash ^= (hash_t)permutate(v, shift); hash = REV(hash)
/* hash_custom_mult, hash.cu:301-326 */
#define MULT_DECL hash_t hash = 1
#define MULT_STEP hash = 1 + (hash * ((hash_t)v + 1))
/* hash_custom_mult_shift, hash.cu:328-359; `(hash * (w + 1)) << off` precedence as written (:340,:344) */
#define MULTS_DECL hash_t hash = 1; short hash_offset = 0
#define MULTS_STEP hash_offset = (hash_offset + v) % MS; hash = 1 + SHL((hash_t)(hash * ((hash_t)v + 1)), hash_offset)
/* hash_custom_mult_hw, hash.cu:361-396 */
#define MULTHW_DECL hash_t hash = 1; int shift = 0
#define MULTHW_STEP shift += v; hash = 1 + (hash * ((hash_t)permutate(v, shift) + 1)); hash = REV(hash)
/* mult_xor_shift helper, hash.cu:398-406; chunk passed as hash_t */
#define MXS_DECL hash_t hash = 0; short hash_offset = 0; short operation = 0
#define MXS_STEP                                                                      \
    do {                                                                              \
        hash_t chunk = (hash_t)v;                                                     \
        hash_offset = (hash_offset + chunk) % MS;                                     \
        if (operation & 1) {                                                          \
            hash = 1 + (hash * SHL((hash_t)(chunk + 1), hash_offset));                \
        } else {                                                                      \
            hash = (hash ^ SHL((hash_t)chunk, hash_offset));                          \
        }                                                                             \
        operation++;                                                                  \
    } while (0)
/* mult_xor_hw helper, hash.cu:408-417 */
#define MXHW_DECL hash_t hash = 0; short hash_offset = 0; short operation = 0
#define MXHW_STEP                                                                     \
    do {                                                                              \
        hash_t chunk = (hash_t)v;                                                     \
        if (operation & 1) {                                                          \
            hash = 1 + (hash * ((hash_t)permutate((uint32_t)(chunk + 1), hash_offset))); \
        } else {                                                                      \
            hash = (hash ^ ((hash_t)permutate((uint32_t)chunk, hash_offset)));        \
        }                                                                             \
        operation++;                                                                  \
        hash_offset += chunk;                                                         \
        hash = REV(hash);                                                             \
    } while (0)

#define DEFINE_ALL(W)                                                          \
    KIND_FN(W, k_fnv, FNV_DECL, FNV_STEP, NOTHING, 1)                          \
    KIND_FN(W, k_add, ADD_DECL, ADD_STEP, NOTHING, 1)                          \
    KIND_FN(W, k_add_shift, ADDS_DECL, ADDS_STEP, NOTHING, 1)                  \
    KIND_FN(W, k_add_hw, ADDHW_DECL, ADDHW_STEP, ADDHW_END, 1)                 \
    KIND_FN(W, k_xor, XOR_DECL, XOR_STEP, NOTHING, 1)                          \
    KIND_FN(W, k_xor_shift, XORS_DECL, XORS_STEP, NOTHING, 1)                  \
    KIND_FN(W, k_xor_hw, XORHW_DECL, XORHW_STEP, NOTHING, 1)                   \
    KIND_FN(W, k_mult, MULT_DECL, MULT_STEP, NOTHING, 1)                       \
    KIND_FN(W, k_mult_shift, MULTS_DECL, MULTS_STEP, NOTHING, 1)               \
    KIND_FN(W, k_mult_hw, MULTHW_DECL, MULTHW_STEP, NOTHING, 1)                \
    KIND_FN(W, k_mult_xor_shift, MXS_DECL, MXS_STEP, NOTHING, 1)               \
    KIND_FN(W, k_n_mult_xor_shift, MXS_DECL, MXS_STEP, NOTHING, 2)             \
    KIND_FN(W, k_mult_xor_hw, MXHW_DECL, MXHW_STEP, NOTHING, 1)

DEFINE_ALL(32)
DEFINE_ALL(64)

typedef uint32_t (*fn32)(const unsigned char *, int, int);
typedef uint64_t (*fn64)(const unsigned char *, int, int);
static const fn32 FN32[OR_HASH_KINDS] = {k_fnv_32, k_add_32, k_add_shift_32, k_add_hw_32, k_xor_32,
                                          k_xor_shift_32, k_xor_hw_32, k_mult_32, k_mult_shift_32, k_mult_hw_32,
                                          k_mult_xor_shift_32, k_n_mult_xor_shift_32, k_mult_xor_hw_32};
static const fn64 FN64[OR_HASH_KINDS] = {k_fnv_64, k_add_64, k_add_shift_64, k_add_hw_64, k_xor_64,
                                          k_xor_shift_64, k_xor_hw_64, k_mult_64, k_mult_shift_64, k_mult_hw_64,
                                          k_mult_xor_shift_64, k_n_mult_xor_shift_64, k_mult_xor_hw_64};

/* Names of the string dispatch in hash_func, hash.cu:515-541. */
static const char *KIND_NAMES[OR_HASH_KINDS] = {
    "fnv", "custom_add", "custom_add_shift", "custom_add_hw", "custom_xor", "custom_xor_shift", "custom_xor_hw",
    "custom_mult", "custom_mult_shift", "custom_mult_hw", "custom_mult_xor_shift", "custom_n_mult_xor_shift",
    "custom_mult_xor_hw"};

int or_hash_kind_from_name(const char *name) {
    if (!name) return -1;
    for (int k = 0; k < OR_HASH_KINDS; k++)
        if (strcmp(name, KIND_NAMES[k]) == 0) return k;
    return -1;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* JoinProvider::hash_table + hash_func, join_provider.cuh:913-946 and hash.cu:507-557 */
int or_hash_rows(const uint64_t *rows, uint64_t n, int ncols, int kind, int hash_bits, int chunk_bits,
                 uint64_t *out, int nthreads) {
    if (kind < 0 || kind >= OR_HASH_KINDS) return -1;
    if (hash_bits != 32 && hash_bits != 64) return -1;
    if (chunk_bits != 8 && chunk_bits != 32 && chunk_bits != 64 && chunk_bits != 128) return -1;
    if (ncols < 1) return -1;
    int nbytes = ncols * 8;
    if ((nbytes * 8) % chunk_bits) return -1; /* join_provider.cuh:926-928 */
    set_threads(nthreads);
    const unsigned char *base = (const unsigned char *)rows;
    if (hash_bits == 32) {
        fn32 f = FN32[kind];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; i++) out[i] = f(base + (size_t)i * nbytes, nbytes, chunk_bits);
    } else {
        fn64 f = FN64[kind];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; i++) out[i] = f(base + (size_t)i * nbytes, nbytes, chunk_bits);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Radix partition (histogram_kernel, prefix_sum_kernel, element_swap_kernel */
/* join_provider.cu:315-380, digit get_radix_mask :303-305)                  */
/* ------------------------------------------------------------------------ */

void or_radix_partition(const uint64_t *hashes, uint64_t n, int shift, int bits, uint64_t *hist, uint64_t *offsets,
                        uint32_t *perm) {
    uint64_t bins = 1ull << bits, mask = bins - 1;
    memset(hist, 0, bins * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) hist[(hashes[i] >> shift) & mask]++;
    uint64_t acc = 0;
    for (uint64_t b = 0; b < bins; b++) {
        offsets[b] = acc;
        acc += hist[b];
    }
    if (perm) {
        uint64_t *cur = (uint64_t *)malloc(bins * sizeof(uint64_t));
        memcpy(cur, offsets, bins * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) perm[cur[(hashes[i] >> shift) & mask]++] = (uint32_t)i;
        free(cur);
    }
}

/* ------------------------------------------------------------------------ */
/* Partitioned join                                                          */
/* ------------------------------------------------------------------------ */

/* A side of the join in partitioned form: tuples (hash, rowid, key row). */
typedef struct {
    uint64_t *h;
    uint32_t *id;
    uint64_t *k; /* ncols per tuple */
} tuples_t;

static int alloc_tuples(tuples_t *t, uint64_t n, int ncols) {
    size_t m = n ? n : 1;
    t->h = (uint64_t *)malloc(m * sizeof(uint64_t));
    t->id = (uint32_t *)malloc(m * sizeof(uint32_t));
    t->k = (uint64_t *)malloc(m * ncols * sizeof(uint64_t));
    return t->h && t->id && t->k;
}
static void free_tuples(tuples_t *t) {
    free(t->h);
    free(t->id);
    free(t->k);
}

/*
 * One LSB radix pass over [lo, hi) of src into dst (element_swap_kernel,
 * join_provider.cu:356-380: (hash, pk, columns) moved together). The pass is
 * stable here; the reference's is not, which only changes the order.
 * sub_off receives bins+1 absolute offsets of the children.
 */
static void radix_pass(const tuples_t *src, tuples_t *dst, uint64_t lo, uint64_t hi, int ncols, int shift, int bits,
                       uint64_t *sub_off) {
    uint64_t bins = 1ull << bits, mask = bins - 1;
    uint64_t *cnt = (uint64_t *)calloc(bins, sizeof(uint64_t));
    for (uint64_t i = lo; i < hi; i++) cnt[(src->h[i] >> shift) & mask]++;
    uint64_t acc = lo;
    for (uint64_t b = 0; b < bins; b++) {
        sub_off[b] = acc;
        acc += cnt[b];
        cnt[b] = sub_off[b];
    }
    sub_off[bins] = hi;
    for (uint64_t i = lo; i < hi; i++) {
        uint64_t d = (src->h[i] >> shift) & mask;
        uint64_t o = cnt[d]++;
        dst->h[o] = src->h[i];
        dst->id[o] = src->id[i];
        memcpy(&dst->k[o * ncols], &src->k[i * ncols], ncols * sizeof(uint64_t));
    }
    free(cnt);
}

/* Parallel first pass over the whole side: per-thread histograms. */
static void radix_pass_par(const tuples_t *src, tuples_t *dst, uint64_t n, int ncols, int shift, int bits,
                           uint64_t *sub_off) {
    uint64_t bins = 1ull << bits, mask = bins - 1;
    int T = 1;
#ifdef _OPENMP
    T = omp_get_max_threads();
#endif
    uint64_t *cnt = (uint64_t *)calloc((size_t)T * bins, sizeof(uint64_t));
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        uint64_t lo = n * t / T, hi = n * (t + 1) / T;
        uint64_t *c = cnt + (size_t)t * bins;
        for (uint64_t i = lo; i < hi; i++) c[(src->h[i] >> shift) & mask]++;
#pragma omp barrier
#pragma omp single
        {
            uint64_t acc = 0;
            for (uint64_t b = 0; b < bins; b++) {
                sub_off[b] = acc;
                for (int u = 0; u < T; u++) {
                    uint64_t v = cnt[(size_t)u * bins + b];
                    cnt[(size_t)u * bins + b] = acc;
                    acc += v;
                }
            }
            sub_off[bins] = n;
        }
        for (uint64_t i = lo; i < hi; i++) {
            uint64_t d = (src->h[i] >> shift) & mask;
            uint64_t o = c[d]++;
            dst->h[o] = src->h[i];
            dst->id[o] = src->id[i];
            memcpy(&dst->k[o * ncols], &src->k[i * ncols], ncols * sizeof(uint64_t));
        }
    }
    free(cnt);
}

typedef struct {
    uint32_t *r, *s;
    uint64_t n, cap;
} pairbuf_t;

static void pb_push(pairbuf_t *pb, uint32_t r, uint32_t s) {
    if (pb->n == pb->cap) {
        pb->cap = pb->cap ? pb->cap * 2 : 1024;
        pb->r = (uint32_t *)realloc(pb->r, pb->cap * sizeof(uint32_t));
        pb->s = (uint32_t *)realloc(pb->s, pb->cap * sizeof(uint32_t));
    }
    pb->r[pb->n] = r;
    pb->s[pb->n] = s;
    pb->n++;
}

/*
 * Build + probe of one leaf pair (build_hash_table / probe_kernel,
 * join_provider.cu:382-442): chained table, slots = max(1, load * n_r)
 * (join_provider.cu:238-240), slot = (h >> key_offset) % slots, links are
 * 1-based with 0 ending a chain; a pair matches iff the hashes are equal and
 * every key column is equal.
 */
static void build_probe(const tuples_t *R, uint64_t rlo, uint64_t rhi, const tuples_t *S, uint64_t slo, uint64_t shi,
                        int ncols, int key_offset, float load, pairbuf_t *out) {
    uint64_t nr = rhi - rlo;
    if (nr == 0 || shi == slo) return;
    float fs = load * (float)nr;
    uint32_t slots = (uint32_t)(fs > 1.f ? fs : 1.f);
    uint32_t *head = (uint32_t *)calloc(slots, sizeof(uint32_t));
    uint32_t *link = (uint32_t *)malloc(nr * sizeof(uint32_t));
    for (uint64_t i = 0; i < nr; i++) {
        uint64_t slot = (R->h[rlo + i] >> key_offset) % slots;
        link[i] = head[slot];
        head[slot] = (uint32_t)i + 1;
    }
    for (uint64_t j = slo; j < shi; j++) {
        uint64_t sh = S->h[j];
        uint32_t t = head[(sh >> key_offset) % slots];
        while (t) {
            t--;
            uint64_t ri = rlo + t;
            if (R->h[ri] == sh) {
                int eq = 0;
                for (int c = 0; c < ncols; c++) eq += (R->k[ri * ncols + c] == S->k[j * ncols + c]);
                if (eq == ncols) pb_push(out, R->id[ri], S->id[j]);
            }
            t = link[t];
        }
    }
    free(head);
    free(link);
}

static int load_side(tuples_t *t, const uint64_t *rows, uint64_t n, const uint32_t *rowids, int ncols,
                     const or_join_cfg *cfg) {
    if (!alloc_tuples(t, n, ncols)) return -1;
    if (or_hash_rows(rows, n, ncols, cfg->kind, cfg->hash_bits, cfg->chunk_bits, t->h, cfg->nthreads)) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) t->id[i] = rowids ? rowids[i] : (uint32_t)i;
    memcpy(t->k, rows, n * ncols * sizeof(uint64_t));
    return 0;
}

/* JoinProvider::join_gpu pipeline restated on the CPU (join_provider.cuh:286-421):
 * hash -> partition_recursive (same digit for R and S, :566-715) -> probe -> merge. */
int64_t or_join(const uint64_t *r_rows, uint64_t nr, const uint32_t *r_rowids, const uint64_t *s_rows, uint64_t ns,
                const uint32_t *s_rowids, int ncols, const or_join_cfg *cfg, uint32_t *out_r, uint32_t *out_s,
                uint64_t cap) {
    if (!cfg || ncols < 1 || cfg->passes < 0 || cfg->passes > 2) return -1;
    if (cfg->passes >= 1 && (cfg->bits1 < 1 || cfg->bits1 > 16)) return -1;
    if (cfg->passes == 2 && (cfg->bits2 < 1 || cfg->bits2 > 16)) return -1;
    set_threads(cfg->nthreads);
    tuples_t R[2], S[2];
    memset(R, 0, sizeof(R));
    memset(S, 0, sizeof(S));
    if (load_side(&R[0], r_rows, nr, r_rowids, ncols, cfg) || load_side(&S[0], s_rows, ns, s_rowids, ncols, cfg)) {
        free_tuples(&R[0]);
        free_tuples(&S[0]);
        return -1;
    }
    int P = cfg->passes;
    int b1 = P >= 1 ? cfg->bits1 : 0, b2 = P >= 2 ? cfg->bits2 : 0;
    uint64_t nparts = 1ull << (b1 + b2);
    uint64_t *roff = (uint64_t *)malloc((nparts + 1) * sizeof(uint64_t));
    uint64_t *soff = (uint64_t *)malloc((nparts + 1) * sizeof(uint64_t));
    int cur = 0;
    if (P == 0) {
        roff[0] = 0;
        roff[1] = nr;
        soff[0] = 0;
        soff[1] = ns;
    } else {
        alloc_tuples(&R[1], nr, ncols);
        alloc_tuples(&S[1], ns, ncols);
        uint64_t *r1 = (uint64_t *)malloc(((1ull << b1) + 1) * sizeof(uint64_t));
        uint64_t *s1 = (uint64_t *)malloc(((1ull << b1) + 1) * sizeof(uint64_t));
        radix_pass_par(&R[0], &R[1], nr, ncols, 0, b1, r1);
        radix_pass_par(&S[0], &S[1], ns, ncols, 0, b1, s1);
        cur = 1;
        if (P == 1) {
            memcpy(roff, r1, ((1ull << b1) + 1) * sizeof(uint64_t));
            memcpy(soff, s1, ((1ull << b1) + 1) * sizeof(uint64_t));
        } else {
            uint64_t bins2 = 1ull << b2;
            /* second level: each first-level bucket is split again at shift b1
             * (partition_recursive recursion, join_provider.cuh:575,694-699) */
#pragma omp parallel for schedule(dynamic, 1)
            for (int64_t p = 0; p < (int64_t)(1ull << b1); p++) {
                uint64_t *sub = (uint64_t *)malloc((bins2 + 1) * sizeof(uint64_t));
                radix_pass(&R[1], &R[0], r1[p], r1[p + 1], ncols, b1, b2, sub);
                memcpy(&roff[p * bins2], sub, bins2 * sizeof(uint64_t));
                radix_pass(&S[1], &S[0], s1[p], s1[p + 1], ncols, b1, b2, sub);
                memcpy(&soff[p * bins2], sub, bins2 * sizeof(uint64_t));
                free(sub);
            }
            roff[nparts] = nr;
            soff[nparts] = ns;
            cur = 0;
        }
        free(r1);
        free(s1);
    }
    int key_offset = b1 + b2; /* radix_width * depth, join_provider.cuh:817 */
    int T = 1;
#ifdef _OPENMP
    T = omp_get_max_threads();
#endif
    pairbuf_t *pbs = (pairbuf_t *)calloc(T, sizeof(pairbuf_t));
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 16)
        for (int64_t q = 0; q < (int64_t)nparts; q++)
            build_probe(&R[cur], roff[q], roff[q + 1], &S[cur], soff[q], soff[q + 1], ncols, key_offset, cfg->load,
                        &pbs[t]);
    }
    /* merge_joined_tables, join_provider.cuh:948-996 */
    uint64_t total = 0;
    for (int t = 0; t < T; t++) total += pbs[t].n;
    uint64_t o = 0;
    for (int t = 0; t < T; t++) {
        for (uint64_t i = 0; i < pbs[t].n && o < cap; i++, o++) {
            out_r[o] = pbs[t].r[i];
            out_s[o] = pbs[t].s[i];
        }
        free(pbs[t].r);
        free(pbs[t].s);
    }
    free(pbs);
    free(roff);
    free(soff);
    for (int i = 0; i < 2; i++) {
        free_tuples(&R[i]);
        free_tuples(&S[i]);
    }
    return (int64_t)total;
}

/* JoinProvider::join_cpu, join_provider.cuh:241-284 */
int64_t or_join_brute(const uint64_t *r_rows, uint64_t nr, const uint32_t *r_rowids, const uint64_t *s_rows,
                      uint64_t ns, const uint32_t *s_rowids, int ncols, uint32_t *out_r, uint32_t *out_s,
                      uint64_t cap) {
    uint64_t m = 0;
    for (uint64_t s = 0; s < ns; s++) {
        for (uint64_t r = 0; r < nr; r++) {
            int match_counter = 0;
            for (int c = 0; c < ncols; c++) match_counter += (r_rows[r * ncols + c] == s_rows[s * ncols + c]);
            if (match_counter == ncols) {
                if (m < cap) {
                    out_r[m] = r_rowids ? r_rowids[r] : (uint32_t)r;
                    out_s[m] = s_rowids ? s_rowids[s] : (uint32_t)s;
                }
                m++;
            }
        }
    }
    return (int64_t)m;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

void or_sort_pairs(uint32_t *r, uint32_t *s, uint64_t n) {
    uint64_t *tmp = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) tmp[i] = ((uint64_t)r[i] << 32) | s[i];
    qsort(tmp, n, sizeof(uint64_t), cmp_u64);
    for (uint64_t i = 0; i < n; i++) {
        r[i] = (uint32_t)(tmp[i] >> 32);
        s[i] = (uint32_t)tmp[i];
    }
    free(tmp);
}
