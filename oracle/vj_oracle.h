/*
 * vj_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of vgerber/vectorized-join's hashing and single-GPU hash-join
 * path. It is the parity checker for the HIP library and the CPU baseline that
 * bench.py times; nothing in the product links or calls it.
 *
 * Data model follows the reference (src/base/types.hpp:128-266): key columns
 * are given row-major as 64-bit values (`column_t`), one row after the other,
 * exactly the buffer the reference hashes (src/join/join_provider.cuh:930-931).
 *
 * Hash kinds are numbered in the order below; the names match the string
 * dispatch of hash_func (src/hash/hash.cu:515-541).
 */
#ifndef VJ_ORACLE_H
#define VJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_FNV = 0,
    OR_CUSTOM_ADD = 1,
    OR_CUSTOM_ADD_SHIFT = 2,
    OR_CUSTOM_ADD_HW = 3,
    OR_CUSTOM_XOR = 4,
    OR_CUSTOM_XOR_SHIFT = 5,
    OR_CUSTOM_XOR_HW = 6,
    OR_CUSTOM_MULT = 7,
    OR_CUSTOM_MULT_SHIFT = 8,
    OR_CUSTOM_MULT_HW = 9,
    OR_CUSTOM_MULT_XOR_SHIFT = 10,
    OR_CUSTOM_N_MULT_XOR_SHIFT = 11,
    OR_CUSTOM_MULT_XOR_HW = 12,
    OR_HASH_KINDS = 13
};

/* Returns the kind id for a reference algorithm name, or -1 if unknown. */
int or_hash_kind_from_name(const char *name);

/*
 * Hash n rows of the row-major 64-bit key buffer `rows` (ncols values per
 * row). hash_bits in {32,64}; chunk_bits in {8,32,64,128} is the reference's
 * HASH_CHUNK_BITS. Output: n uint64 (32-bit hashes zero-extended).
 * Returns 0 on success, -1 on bad arguments (unknown kind, widths, or a row
 * width that is not a whole number of chunks, join_provider.cuh:926-928).
 */
int or_hash_rows(const uint64_t *rows, uint64_t n, int ncols, int kind, int hash_bits, int chunk_bits,
                 uint64_t *out, int nthreads);

/*
 * Stable LSB radix partition of n hashes on digit (h >> shift) & (2^bits-1).
 * hist[b] = count, offsets[b] = exclusive prefix, perm[k] = source index of
 * the k-th element of the partitioned order (stable).
 */
void or_radix_partition(const uint64_t *hashes, uint64_t n, int shift, int bits, uint64_t *hist, uint64_t *offsets,
                        uint32_t *perm);

/* Join configuration for or_join (the reference's JoinConfig subset). */
typedef struct {
    int kind;        /* hash kind */
    int hash_bits;   /* 32 / 64 */
    int chunk_bits;  /* 8 / 32 / 64 / 128 */
    int passes;      /* radix passes: 0, 1, 2 */
    int bits1;       /* radix width of pass 1 */
    int bits2;       /* radix width of pass 2 */
    float load;      /* build table load (slots = max(1, load * n_r)), join_provider.cu:238-240 */
    int nthreads;    /* OpenMP threads (<= 0: runtime default) */
} or_join_cfg;

/*
 * Partitioned hash join with the reference semantics: hash both tables,
 * LSB radix partition them with the same digits, build a chained table per
 * leaf partition of R and probe it with the leaf partition of S; a pair
 * matches iff the hashes are equal and every key column is equal
 * (join_provider.cu:382-442). Row ids are r_rowids[i] / s_rowids[j] when
 * given, else the 0-based row index.
 * Writes up to `cap` pairs into out_r/out_s, returns the total match count
 * (may exceed cap), or -1 on bad arguments.
 */
int64_t or_join(const uint64_t *r_rows, uint64_t nr, const uint32_t *r_rowids, const uint64_t *s_rows, uint64_t ns,
                const uint32_t *s_rowids, int ncols, const or_join_cfg *cfg, uint32_t *out_r, uint32_t *out_s,
                uint64_t cap);

/*
 * Brute-force nested-loop join (JoinProvider::join_cpu,
 * join_provider.cuh:241-284): pairs in s-major / r-minor order.
 */
int64_t or_join_brute(const uint64_t *r_rows, uint64_t nr, const uint32_t *r_rowids, const uint64_t *s_rows,
                      uint64_t ns, const uint32_t *s_rowids, int ncols, uint32_t *out_r, uint32_t *out_s,
                      uint64_t cap);

/* Sort pairs lexicographically by (r, s) in place. */
void or_sort_pairs(uint32_t *r, uint32_t *s, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
