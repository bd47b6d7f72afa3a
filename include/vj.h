/*
 * vj.h -- C ABI of the MI355X (gfx950) hash-join and hashing library.
 *
 * This is the drop-in boundary for vgerber/vectorized-join's single-GPU hash
 * join and hashing path. The reference has no ABI of its own (its executables
 * #include the .cu files); each entry point below names the reference
 * interface it replaces.
 *
 * Conventions
 *  - Every function returns a vj_status; nothing calls exit() or throws across
 *    the ABI (the reference's JoinStatus{successful,message},
 *    src/join/join_provider.cu:11-35, becomes a status code plus
 *    vj_last_error()).
 *  - All table and output pointers are device pointers (HBM) unless stated.
 *  - `stream` is a hipStream_t passed as void*; NULL is the default stream.
 *  - A vj_ctx is not thread-safe; distinct contexts are independent.
 *  - Tables hold at most VJ_MAX_ROWS rows; row ids are uint32. (The kernels
 *    index rows, tiles and probe chunks in 32-bit arithmetic; the margin below
 *    2^32 keeps the last tile's and chunk's end inside that range.)
 */
#ifndef VJ_H
#define VJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VJ_ABI_VERSION 4
#define VJ_MAX_KEY_COLS 4
#define VJ_MAX_ROWS 4294901760ull /* 2^32 - 2^16 */

typedef enum vj_status {
    VJ_OK = 0,
    VJ_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad width, bad count */
    VJ_ERR_UNKNOWN_HASH = 2,     /* unknown hash kind or name (hash.cu:515-541 silently did nothing) */
    VJ_ERR_UNSUPPORTED = 3,      /* a valid request this build does not support */
    VJ_ERR_BUDGET = 4,           /* admission refused: the HBM budget would be exceeded */
    VJ_ERR_OUT_OF_MEMORY = 5,    /* hipMalloc failed */
    VJ_ERR_CAPACITY = 6,         /* caller's output buffer too small; the count holds the size needed */
    VJ_ERR_HIP = 7,              /* HIP runtime error */
    VJ_ERR_NO_DEVICE = 8,        /* no usable gfx950 device */
    VJ_ERR_NOT_FOUND = 9,        /* unknown handle / ticket */
    VJ_ERR_INTERNAL = 10
} vj_status;

/* Hash algorithms, in the order of the reference's string dispatch
 * (hash_func, src/hash/hash.cu:515-541; kernels hash.cu:84-505). */
typedef enum vj_hash_kind {
    VJ_HASH_FNV = 0,
    VJ_HASH_CUSTOM_ADD = 1,
    VJ_HASH_CUSTOM_ADD_SHIFT = 2,
    VJ_HASH_CUSTOM_ADD_HW = 3,
    VJ_HASH_CUSTOM_XOR = 4,
    VJ_HASH_CUSTOM_XOR_SHIFT = 5,
    VJ_HASH_CUSTOM_XOR_HW = 6,
    VJ_HASH_CUSTOM_MULT = 7,
    VJ_HASH_CUSTOM_MULT_SHIFT = 8,
    VJ_HASH_CUSTOM_MULT_HW = 9,
    VJ_HASH_CUSTOM_MULT_XOR_SHIFT = 10,
    VJ_HASH_CUSTOM_N_MULT_XOR_SHIFT = 11,
    VJ_HASH_CUSTOM_MULT_XOR_HW = 12,
    VJ_HASH_KIND_COUNT = 13
} vj_hash_kind;

/* How the 32-bit hash word stream is formed from the key columns. */
typedef enum vj_word_mode {
    /* The reference layout: every key column is a 64-bit column_t
     * (src/base/types.hpp:7) and contributes its low then high 32-bit word
     * (src/hash/hash.cu:92-113 with chunk_t = uint2). 4-byte keys are widened
     * (zero- or sign-extended). Bit-exact with the reference. */
    VJ_WORDS_REFERENCE = 0,
    /* Native: a 4-byte key column contributes one word, an 8-byte column two. */
    VJ_WORDS_NATIVE = 1
} vj_word_mode;

/* A table of key columns in HBM, column-major (SoA). Replaces db_table
 * (src/base/types.hpp:128-266), whose key columns are row-major uint64. */
typedef struct vj_table {
    uint64_t n_rows;
    int32_t n_cols;      /* 1..VJ_MAX_KEY_COLS */
    int32_t key_width;   /* 4 or 8 bytes per key column */
    int32_t key_signed;  /* 4-byte keys: widen by sign extension (else zero extension) */
    int32_t reserved;
    const void *cols[VJ_MAX_KEY_COLS]; /* device pointers, n_rows values each */
    /* Optional row ids carried with each row (the reference's primary_keys
     * column, types.hpp:130). NULL: the 0-based row index is used. */
    const uint32_t *rowids;
} vj_table;

typedef struct vj_hash_opts {
    int32_t kind;       /* vj_hash_kind */
    int32_t hash_bits;  /* 32 or 64 (reference HASH_BITS, src/base/config.h:28) */
    int32_t chunk_bits; /* 8, 32, 64 or 128 (reference HASH_CHUNK_BITS, config.h:31) */
    int32_t word_mode;  /* vj_word_mode */
} vj_hash_opts;

/* Defaults are the reference's compiled configuration: fnv, 64, 64, reference words. */
void vj_hash_opts_default(vj_hash_opts *opts);

const char *vj_status_str(vj_status status);

/* Maps a reference algorithm name ("fnv", "custom_xor_hw", ...) to its kind.
 * Replaces the string compare chain of hash_func (src/hash/hash.cu:515-541);
 * an unknown name is VJ_ERR_UNKNOWN_HASH instead of a silent no-op. */
vj_status vj_hash_kind_from_name(const char *name, int32_t *kind);
const char *vj_hash_kind_name(int32_t kind);

/*
 * Hash every row of `table` into `out` (n_rows values of hash_bits/8 bytes).
 * Replaces hash_func (src/hash/hash.cu:507-557) as used by
 * JoinProvider::hash_table (src/join/join_provider.cuh:913-946): the word
 * stream is all key columns from offset 0. Asynchronous on `stream`.
 * Errors: VJ_ERR_UNKNOWN_HASH, VJ_ERR_INVALID_ARGUMENT (including a row width
 * that is not a whole number of chunks, join_provider.cuh:926-928).
 */
vj_status vj_hash(const vj_table *table, const vj_hash_opts *opts, void *out, void *stream);

/* ------------------------------------------------------------------------ */
/* Context (JoinProvider, src/join/join_provider.cuh:222-500)                */
/* ------------------------------------------------------------------------ */

typedef struct vj_ctx vj_ctx;

/* Creates a context on HIP device `device`. Replaces
 * JoinProvider::JoinProvider(JoinConfig) (join_provider.cuh:225-228). */
vj_status vj_ctx_create(int device, vj_ctx **ctx);
/* Frees all device memory of the context (JoinProvider::~JoinProvider, :427-429). */
vj_status vj_ctx_destroy(vj_ctx *ctx);
/* HBM admission budget in bytes for scratch + outputs the context allocates
 * (0 = unlimited). The reference's DeviceConfig::is_out_of_memory /
 * MEMORY_TOLERANCE (join_provider.cuh:140-150, config.h:25). */
vj_status vj_ctx_set_budget(vj_ctx *ctx, uint64_t bytes);
/* Bytes of device memory the context currently holds. */
vj_status vj_ctx_memory(vj_ctx *ctx, uint64_t *bytes);
/* Detail message of the last failure on this context ("" if none). */
const char *vj_last_error(const vj_ctx *ctx);

typedef struct vj_join_opts {
    vj_hash_opts hash;
    int32_t passes;     /* radix passes: -1 = planner, 0, 1 or 2 */
    int32_t bits1;      /* radix bits of pass 1 (0 = planner; 1..8) */
    int32_t bits2;      /* radix bits of pass 2 (0 = planner; 1..8) */
    int32_t one_based;  /* emit row ids + 1 (the reference's primary_keys = row + 1) */
    uint64_t out_capacity_hint; /* expected pair count for vj_join (0 = probe rows) */
    int32_t tile;       /* radix-pass tile: 0 = planner / autotuned, VJ_TILE_STANDARD, VJ_TILE_SMALL or VJ_TILE_LARGE */
    int32_t table;      /* build+probe block: 0 = planner / autotuned, VJ_TABLE_STANDARD or VJ_TABLE_WIDE */
    int32_t layout;     /* partition layout: 0 = planner / autotuned, VJ_LAYOUT_EXACT or VJ_LAYOUT_REGIONS */
    int32_t reserved;   /* 0 */
} vj_join_opts;

/* Kernel shapes (vj_join_opts.tile / .table). All joins of one pipeline (a
 * batch) share them. The planner takes the standard tile (the large one in
 * the region layout), and the wide table
 * when the standard one would need more than 14 radix bits (build sides over
 * 2^24 rows of 8-byte keys); vj_tune times the alternatives. */
#define VJ_TILE_STANDARD 1   /* 512 threads x 6 rows (keys <= 8 bytes), x 4 (<= 16), x 2 (wider) */
#define VJ_TILE_SMALL 2      /* 512 threads x 2 rows: more blocks for small tables */
#define VJ_TILE_LARGE 3      /* 512 threads x 8 rows (keys <= 8 bytes), x 6 (<= 16), x 4 (wider): longer runs
                                per digit; the planner's tile in the region layout */
#define VJ_TABLE_STANDARD 1  /* 512 threads (8 waves), up to four blocks per CU, LDS table of 2048 rows
                                (keys <= 8 bytes) or 1024 rows */
#define VJ_TABLE_WIDE 2      /* 1024 threads (16 waves), one block per CU, LDS table of 8192 rows
                                (keys <= 8 bytes), 4096 (<= 24 bytes) or 2048 */

/* Partition layouts (vj_join_opts.layout). The planner takes the exact one;
 * vj_tune times both for two-pass plans.
 * EXACT: a histogram read of the keys gives every row its exact destination
 * (partitions back to back; no device atomics).
 * REGIONS: no histogram read. Each partition gets a region of fixed capacity
 * (its mean size plus a margin of 32 standard deviations of a uniform spread,
 * about 1.2-2x the mean) and the radix passes reserve their runs in it with
 * device atomics. Faster where keys spread evenly over the partitions (the
 * histogram is about a tenth of a large join); when a region overflows (heavy
 * key skew) the join is run again with the exact layout, the summary reports
 * it (layout_fallback), and the context uses the exact layout for that size
 * class from then on. Two-pass plans with at least 3 bits in the first pass
 * only, and not for batches that hold resident or kept build sides, served
 * batches or vj_join_into_noalloc: these take the exact layout. */
#define VJ_LAYOUT_EXACT 1
#define VJ_LAYOUT_REGIONS 2

void vj_join_opts_default(vj_join_opts *opts);

/* Join result: (build_rowid, probe_rowid) pairs in device memory owned by the
 * caller after return; release with vj_pairs_free. */
typedef struct vj_pairs {
    uint64_t count;
    uint32_t *build_rowids;
    uint32_t *probe_rowids;
} vj_pairs;

/*
 * Equi-join on all key columns: every (r, s) with build row r and probe row s
 * whose key columns are all equal. Output order is unspecified (as in the
 * reference, whose atomics make it nondeterministic); the pair set is exact.
 * Replaces JoinStatus JoinProvider::join(db_table r, db_table s, db_table &rs)
 * (join_provider.cuh:230-239, join_gpu :286-421). Like the reference it is
 * synchronous; unlike it, inputs are never modified (the reference reuses
 * them as swap space, join_provider.cuh:652-663).
 * Output memory is allocated by the callee (as join_provider.cuh:975-976).
 */
vj_status vj_join(vj_ctx *ctx, const vj_table *build, const vj_table *probe, const vj_join_opts *opts,
                  vj_pairs *out, void *stream);

/*
 * Same join into caller-provided device buffers of `capacity` pairs.
 * *count receives the total number of matches. If it exceeds capacity the
 * first `capacity` pairs (in unspecified order) are written and
 * VJ_ERR_CAPACITY is returned. No allocation happens when the context's
 * scratch is already large enough (the case for repeated joins of a shape).
 * When per-stage profiling is off, the second call of one launch shape
 * captures the pipeline in a hipGraph and later calls of that shape replay it.
 */
vj_status vj_join_into(vj_ctx *ctx, const vj_table *build, const vj_table *probe, const vj_join_opts *opts,
                       uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count, void *stream);

/*
 * No-allocation variant of vj_join_into. It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc, hipFree, hipHostMalloc or other allocating routine:
 *  - the pipeline's scratch is a fixed array of VJ_NOALLOC_SCRATCH_BYTES per
 *    device, which the library holds on every GPU it runs on until the
 *    process ends; its address is looked up once per device by vj_ctx_create;
 *  - descriptors and match counts pass through a small pinned block the
 *    context allocated in vj_ctx_create.
 * A join whose scratch would exceed the array fails with VJ_ERR_UNSUPPORTED
 * before anything is launched. Two radix passes need about
 * (build rows + probe rows) x (key bytes + 4) x 2 bytes, one pass half that.
 * Every call on one GPU uses that GPU's array, whichever context it comes
 * from: calls that overlap on one GPU race (run them in order). Nothing is
 * captured into or replayed from a hipGraph. Results, capacity handling and
 * the summary are those of vj_join_into.
 */
#define VJ_NOALLOC_SCRATCH_BYTES (64ull << 20)
vj_status vj_join_into_noalloc(vj_ctx *ctx, const vj_table *build, const vj_table *probe, const vj_join_opts *opts,
                               uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count,
                               void *stream);

vj_status vj_pairs_free(vj_pairs *pairs);

/* ------------------------------------------------------------------------ */
/* The reference's own table layout (db_table, src/base/types.hpp:128-266)   */
/* ------------------------------------------------------------------------ */

/* A table as the reference stores it: every key column a 64-bit column_t,
 * row-major (column_values[row * n_cols + c]), and a 64-bit primary key per
 * row. Device pointers. */
typedef struct vj_rows {
    uint64_t n_rows;
    int32_t n_cols;          /* 1..VJ_MAX_KEY_COLS key columns per row */
    int32_t reserved;
    uint64_t *column_values; /* n_rows * n_cols values, row-major */
    uint64_t *primary_keys;  /* n_rows values */
} vj_rows;

/*
 * JoinProvider::join on GPU tables in the reference layout (join_gpu,
 * src/join/join_provider.cuh:286-421), with the reference's result: `rs` gets
 * n_cols = 2, column_values[2i] = r->primary_keys[r_row], column_values[2i+1]
 * = s->primary_keys[s_row] for every matching pair, and primary_keys[i] = i + 1
 * (copy_probe_results_kernel, src/join/join_provider.cu:593-608;
 * merge_joined_tables and generate_primary_key_kernel, join_provider.cuh:948-996,
 * join_provider.cu:307-313). Keys hash as the reference's word stream (each
 * column's low then high 32-bit word). `rs` is allocated by the callee (as
 * join_provider.cuh:975-976); release it with vj_rows_free. Pair order is
 * unspecified. Synchronous; the inputs are not modified. Multi-column rows
 * are transposed into temporary key columns first (freed before return).
 */
vj_status vj_join_rows(vj_ctx *ctx, const vj_rows *r, const vj_rows *s, const vj_join_opts *opts, vj_rows *rs,
                       void *stream);
/* Frees a table vj_join_rows allocated (db_table::free, types.hpp:235-261). */
vj_status vj_rows_free(vj_rows *rows);

/* Per-join statistics of the last join on the context (JoinSummary,
 * join_provider.cuh:54-106; times from hipEvents around each stage). */
typedef struct vj_join_summary {
    uint64_t build_rows, probe_rows, matches;
    int32_t passes, bits1, bits2, retries;
    uint64_t partitions, tasks;
    double ms_total;     /* all stages, device time */
    double ms_histogram, ms_pass1, ms_pass2, ms_build_probe;
    uint64_t scratch_bytes;
    int32_t graph_replay;  /* 1: the join ran as a replay of a captured hipGraph (same launch shape as before) */
    int32_t layout_fallback; /* 1: a region of VJ_LAYOUT_REGIONS overflowed; the join ran again with the exact layout */
    /* Algorithmic HBM bytes of each stage -- the key bytes and row ids of every
     * tuple the stage reads and writes, plus 8 bytes per output pair -- and the
     * resulting GB/s over the stage's device time (0 while per-stage timing is
     * off; the total is always timed). The reference reports per-kernel GB/s in
     * PartitionSummary / ProbeSummary (src/join/join_provider.cu:37-93) and
     * per-phase tuples/s in JoinSummary (join_provider.cuh:54-106). */
    uint64_t bytes_histogram, bytes_pass1, bytes_pass2, bytes_build_probe, bytes_total;
    double gbps_histogram, gbps_pass1, gbps_pass2, gbps_build_probe, gbps_total;
    double probe_tuples_per_s; /* probe rows / ms_total */
    int32_t tile_rows;         /* rows per radix-pass tile */
    int32_t probe_threads;     /* threads per build+probe block */
    int32_t table_rows;        /* rows per build chunk (LDS table capacity) */
    int32_t layout;            /* VJ_LAYOUT_EXACT or VJ_LAYOUT_REGIONS: the layout the join's result came from */
} vj_join_summary;

vj_status vj_join_summary_get(vj_ctx *ctx, vj_join_summary *summary);
/* Turn per-stage event timing on or off (off by default; the total is always timed). */
vj_status vj_ctx_set_profiling(vj_ctx *ctx, int enabled);

/*
 * One radix-partition pass of a table on digit (hash >> shift) & (2^bits - 1):
 * the rows are moved into `out` grouped by digit (digit 0 first), and
 * hist[b] / offsets[b] receive the count and exclusive prefix per digit
 * (uint32, device pointers, 2^bits entries each). Row order inside a digit
 * is unspecified. Replaces partition_gpu (src/join/join_provider.cu:610-661);
 * `out` must have n_rows capacity in every key column and in out->rowids.
 */
vj_status vj_partition(vj_ctx *ctx, const vj_table *table, const vj_hash_opts *hash, int32_t shift, int32_t bits,
                       const vj_table *out, uint32_t *hist, uint32_t *offsets, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched joins (the reference runs bucket pairs as concurrent host threads */
/* over streams, JoinProvider::probe, join_provider.cuh:792-911)            */
/* ------------------------------------------------------------------------ */

/*
 * Runs n_joins independent joins as one pipeline: every stage is one launch
 * over tiles of all joins, and build+probe is one persistent launch drawing
 * (join, partition, chunk) tasks from a device work queue. All joins share
 * opts (key schema and hash). Join i writes into out_build[i]/out_probe[i]
 * (capacity caps[i]); counts[i] receives its match count. Returns
 * VJ_ERR_CAPACITY if any join overflowed (its count tells the size needed).
 * With use_graph != 0 the pipeline is captured once per shape in a hipGraph
 * and replayed.
 */
vj_status vj_join_batch(vj_ctx *ctx, int32_t n_joins, const vj_table *builds, const vj_table *probes,
                        const vj_join_opts *opts, uint32_t *const *out_build, uint32_t *const *out_probe,
                        const uint64_t *caps, uint64_t *counts, int32_t use_graph, void *stream);

/*
 * Autotuner (no counterpart in the reference, whose radix width per level is
 * a configuration value, join_provider.cuh:296-300, as are its threads and
 * rows per thread, ProbeConfig, join_provider.cu:177-192). Times every candidate plan near the
 * planner's choice for each table shape (each pass count; radix bits -1..+1;
 * both splits of two passes, each in both partition layouts; standard and wide
 * build+probe blocks), then the
 * other pass tiles with the fastest, on the caller's tables, `reps` timed runs
 * each after a warm-up, and keeps the fastest in the context for the size
 * class of `build` (build rows rounded up to a power of two, key bytes, hash
 * path). Later joins on this context with opts->passes == -1 and no bits set
 * use the tuned plan for their size class (vj_join, vj_join_into,
 * vj_join_batch; a batch takes its shapes from its largest build's class).
 * With opts->passes >= 0 only plans of that pass count are tried; a nonzero
 * opts->tile, opts->table or opts->layout fixes that shape or layout.
 * *best receives opts with the chosen passes/bits/tile/table/layout, *best_ms its
 * median time.
 * Allocates a temporary output; synchronous.
 */
vj_status vj_tune(vj_ctx *ctx, const vj_table *build, const vj_table *probe, const vj_join_opts *opts, int32_t reps,
                  vj_join_opts *best, double *best_ms, void *stream);
/* Forgets every tuned plan of the context (and which size classes overflowed
 * a region of VJ_LAYOUT_REGIONS). */
vj_status vj_tune_clear(vj_ctx *ctx);

/* ------------------------------------------------------------------------ */
/* Resident build sides. No counterpart in the reference, whose partitioned  */
/* tables live for one join (join_gpu, join_provider.cuh:286-421): a served  */
/* table is partitioned once and probed by many later requests.             */
/* ------------------------------------------------------------------------ */

typedef struct vj_built vj_built;

/* Hashes and radix-partitions `build` (planner bits, at least one pass) into
 * a new device allocation that outlives the call. Synchronous. */
vj_status vj_build(vj_ctx *ctx, const vj_table *build, const vj_join_opts *opts, vj_built **out, void *stream);
/* Joins `probe` against a resident build side: only the probe side is hashed
 * and partitioned (with the table's bits). opts->hash must equal the hash the
 * table was built with; passes / bits are ignored. Same output rules as
 * vj_join_into. */
vj_status vj_probe_into(vj_ctx *ctx, const vj_built *built, const vj_table *probe, const vj_join_opts *opts,
                        uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count, void *stream);
vj_status vj_built_bytes(const vj_built *built, uint64_t *bytes);
/* Frees the table (hipFree: waits for the device). No join may still use it. */
vj_status vj_built_free(vj_built *built);

/* ------------------------------------------------------------------------ */
/* Serving engine: continuous batching of join requests under an HBM budget. */
/* The reference's nearest analogue is JoinProvider::probe, which runs       */
/* bucket pairs on host threads and admits them against free device memory  */
/* (join_provider.cuh:792-911, DeviceConfig :124-220).                       */
/* ------------------------------------------------------------------------ */

typedef struct vj_server vj_server;

typedef struct vj_request {
    const vj_table *build; /* build side, or NULL to probe the resident table `table` */
    uint64_t table;        /* resident table id (from an earlier response's `table`) when build is NULL */
    const vj_table *probe;
    uint32_t *out_build, *out_probe; /* caller's device buffers of `capacity` pairs */
    uint64_t capacity;
    int32_t keep_build;    /* keep this request's partitioned build side resident (its id is returned) */
    int32_t reserved;
} vj_request;

typedef struct vj_response {
    int32_t status;        /* vj_status of the request */
    uint32_t batch_size;   /* requests in the launch that ran it */
    uint64_t count;        /* matches (if > capacity: VJ_ERR_CAPACITY, first `capacity` pairs written) */
    uint64_t table;        /* id of the kept build side (keep_build), else 0 */
    uint64_t t_submit_ns, t_launch_ns, t_done_ns; /* steady clock */
} vj_response;

typedef struct vj_server_stats {
    uint64_t submitted, completed, failed, batches;
    uint64_t overlapped_batches; /* launched while another batch was still on the device */
    uint64_t max_batch;          /* most requests in one launch */
    uint64_t resident_tables, resident_bytes;
    uint64_t peak_bytes;         /* most device memory held (arenas + resident tables) */
    uint64_t budget_waits;       /* times the head request waited for memory */
    uint64_t budget_deferred;    /* launches that left queued requests behind because the budget was reached */
} vj_server_stats;

/* Starts the engine on `device` with a worker thread and two launch lanes
 * (each its own context, arenas and stream), so a batch can be launched
 * while the previous one runs. hbm_budget caps arenas + resident tables
 * (0 = none); max_batch caps requests per launch; opts (hash, passes, bits,
 * one_based) apply to every request. */
vj_status vj_server_create(int device, uint64_t hbm_budget, const vj_join_opts *opts, int32_t max_batch,
                           vj_server **out);
/* Queues a request (its tables and outputs must stay valid until it completes).
 * Requests are admitted in arrival order as lanes and memory free up; all
 * admitted requests of a launch run in one pipeline. A request that can
 * never fit the budget completes with VJ_ERR_BUDGET. */
vj_status vj_server_submit(vj_server *srv, const vj_request *req, uint64_t *ticket);
/* Blocks until the request completes, fills *resp and forgets the ticket. */
vj_status vj_server_wait(vj_server *srv, uint64_t ticket, vj_response *resp);
/* Non-blocking: *done = 1 and *resp filled (ticket forgotten) if complete. */
vj_status vj_server_poll(vj_server *srv, uint64_t ticket, vj_response *resp, int32_t *done);
/* Releases a resident table once no queued or running request uses it. */
vj_status vj_server_drop_table(vj_server *srv, uint64_t table);
vj_status vj_server_stats_get(vj_server *srv, vj_server_stats *stats);
/* Finishes queued work, then frees everything. */
vj_status vj_server_destroy(vj_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* VJ_H */
