// Descriptors of one join pipeline (a batch of one or more joins), shared by
// the host planner (vj_context.cpp) and the kernels (vj_join_kernels.hip).
//
// A join has two jobs: its build side (R) and its probe side (S). Each job is
// hashed and radix-partitioned on the same digits; the build+probe stage then
// draws (join, partition, R-chunk, S-chunk) tasks from a device work queue.
//
// Partitioning is count-then-scatter with exact destinations, so scatter
// blocks never wait on each other:
//  * A job's rows are cut into pass-1 tiles (T rows) grouped into segments,
//    and each segment into hsplit histogram rows of consecutive tiles (more
//    histogram blocks than segments when the segments alone would leave CUs
//    idle). The histogram kernel walks each row's tiles in order and records,
//    per tile, the exclusive prefix of its pass-1 digit counts inside the row
//    (tilepre), and per row the pass-1 digit counts (m1 row) and the
//    (pass-1, pass-2) digit counts (m2 row).
//  * Column scans over the rows turn m1 / m2 into per-row offsets, and a scan
//    over the column totals gives the partition bases.
//  * Pass 1 writes tile t's rows of digit d1 at
//      base[d1 << b2] + m1[row(t)][d1] + tilepre[t][d1] + rank.
//  * Pass 2 reads chunk (d1, g) = segment g's rows of pass-1 partition d1 (a
//    contiguous range of the pass-1 output, from row g * hsplit to the next
//    segment's) and writes its rows of digit d2 at
//      base[(d1 << b2) | d2] + m2[g * hsplit][(d1 << b2) | d2] + rank.
//
// HBM layout per partitioned job: SoA copies of the key columns plus a uint32
// row-id column, once after pass 1 ("mid", only for 2-pass jobs) and once
// after the last pass ("fin"). Partition q of a job occupies rows
// [base[q], base[q+1]) of "fin"; q = d1 * 2^b2 + d2 (pass-1 digit major).
//
// Region layout (2-pass pipelines that ask for it): no histogram and no scans.
// Every destination is a region of fixed capacity, and blocks reserve their
// runs in it with one device atomic per digit:
//  * pass 1 writes tile t's rows of digit d1 into region (d1, g) of "mid",
//    g = t / seg_tiles (the segment: segments spread the atomics over
//    2^b1 * nseg counters), rows [(d1 * nseg + g) * region1, + region1);
//  * pass 2 reads that region in tiles (chunk = (d1, g, tile k)) and writes
//    its rows of digit d2 into region q = (d1 << b2) | d2 of "fin", rows
//    [q * region2, + region2), counter m2[q].
// A run that does not fit its region goes to a one-tile trash area after the
// regions, and the counter keeps the full count, so an overflow is visible
// after the passes (k_tasks then schedules no task for the join and sets
// Ctl::overflow; the host runs the join again with the exact layout).
#pragma once
#include <stdint.h>

#include "vj_device.hpp"

namespace vj {

constexpr int kPassThreads = 512;   // radix pass kernels
constexpr int kPassWaves = kPassThreads / 64;
constexpr int kHistThreads = 1024;  // histogram kernel
constexpr int kMaxProbeThreads = 1024;  // build+probe kernel (512 or 1024 threads, see probe_threads)
constexpr int kMaxProbeWaves = kMaxProbeThreads / 64;
constexpr int kMaxSegments = 512;  // segments (histogram blocks) per job
constexpr int kMaxPassBits = 8;     // digits per pass <= 256
constexpr int kMaxHistBits = 14;    // 2D histogram in LDS: 2^14 x 4 B = 64 KiB
constexpr uint32_t kXcds = 8;       // accelerator dies (XCDs) of an MI355X
constexpr uint32_t kXcdGroup = 32;  // consecutive pass-2 chunks kept on one XCD

// Kernel shapes, chosen per pipeline by the planner or the autotuner.
//  tile  0: standard pass tile, 512 threads x pass_items rows (6 / 4 / 2 rows by
//           key bytes); 1: small tile, 512 threads x 2 rows (more blocks for
//           small tables); 2: large tile, 512 threads x 8 / 6 / 4 rows (longer
//           runs per digit; the region layout's tile).
//  table 0: standard build+probe block, 512 threads (up to four blocks per
//           CU) around an LDS table of build_cap rows; 1: wide block, 1024
//           threads (one block per CU) around a table of wide_cap rows, so a
//           partition four to eight times larger fits one build chunk.
__host__ __device__ constexpr int pass_items(int key_bytes, int tile = 0) {
    return tile == 1 ? 2
                     : (tile == 2 ? (key_bytes <= 8 ? 8 : (key_bytes <= 16 ? 6 : 4))
                                  : (key_bytes <= 8 ? 6 : (key_bytes <= 16 ? 4 : 2)));
}
constexpr int kTileShapes = 3;
__host__ __device__ constexpr int pass_tile(int key_bytes, int tile = 0) { return kPassThreads * pass_items(key_bytes, tile); }
__host__ __device__ constexpr int build_cap(int key_bytes) { return key_bytes <= 8 ? 2048 : 1024; }
// Build rows' ids are kept in LDS next to the table for keys wider than 8
// bytes (whose tables are small), so duplicate chains are written without a
// global load per entry; narrower keys read them from the partitioned table.
__host__ __device__ constexpr bool table_keeps_ids(int key_bytes) { return key_bytes > 8; }
// LDS bytes per build row of a table: the key, two 4-byte slots, a 2-byte chain
// link, and the row id where the table keeps it.
__host__ __device__ constexpr int table_row_bytes(int key_bytes) {
    return key_bytes + 10 + (table_keeps_ids(key_bytes) ? 4 : 0);
}
constexpr int kWideTableBytes = 156 * 1024;  // of the 160 KiB of LDS per CU
// Largest power of two <= 8192 whose table fits kWideTableBytes (row indexes
// stay below 2^16 - 1, the no-match marker).
__host__ __device__ constexpr int wide_cap(int key_bytes) {
    int c = 8192;
    while (c > 1024 && c * table_row_bytes(key_bytes) > kWideTableBytes) c >>= 1;
    return c;
}
__host__ __device__ constexpr int probe_cap(int key_bytes, int table) { return table ? wide_cap(key_bytes) : build_cap(key_bytes); }
__host__ __device__ constexpr int probe_threads(int table) { return table ? 1024 : 512; }
constexpr int kProbeChunk = 16384;  // probe rows per build+probe task
static_assert((kProbeChunk & (kProbeChunk - 1)) == 0 && (build_cap(4) & (build_cap(4) - 1)) == 0 &&
                  (build_cap(8) & (build_cap(8) - 1)) == 0 && (build_cap(16) & (build_cap(16) - 1)) == 0 &&
                  (wide_cap(8) & (wide_cap(8) - 1)) == 0 && (wide_cap(12) & (wide_cap(12) - 1)) == 0,
              "chunk capacities are powers of two (the task counts are shifts)");
static_assert(wide_cap(4) == 8192 && wide_cap(8) == 8192 && wide_cap(16) == 4096 && wide_cap(32) == 2048 &&
                  wide_cap(32) * table_row_bytes(32) <= kWideTableBytes,
              "wide tables");

struct JobDesc {
    TView in;                 // the caller's table
    void *mid_cols[4];        // pass-1 output (2-pass jobs)
    uint32_t *mid_ids;
    void *fin_cols[4];        // last-pass output (partitioned jobs)
    uint32_t *fin_ids;
    uint32_t n;
    uint32_t b1, b2;          // radix bits of pass 1 / pass 2 (b1 == 0: not partitioned)
    uint32_t shift0;          // digit of pass 1 = (h >> shift0) & (2^b1 - 1)
    uint32_t base_off;        // u32 offset of the 2^(b1+b2)+1 partition bases in the base arena
    uint32_t p1_tiles;        // pass-1 tiles
    uint32_t seg_tiles;       // pass-1 tiles per segment
    uint32_t nseg;            // segments (pass-2 chunks of one pass-1 digit)
    uint32_t hsplit;          // histogram rows (blocks) per segment
    uint32_t hrow_tiles;      // pass-1 tiles per histogram row (ceil(seg_tiles / hsplit))
    uint32_t nrow;            // histogram rows = nseg * hsplit (= histogram blocks)
    // Region layout (0: exact layout): rows per pass-1 region (d1, segment),
    // rows per final region (partition), pass-2 chunks (tiles) per pass-1 region
    uint32_t region1, region2, region_tiles;
    // u32 offset of the partition ends in the base arena: partition q is rows
    // [base[base_off + q], base[end_off + q]) of fin (exact layout: end_off =
    // base_off + 1; region layout: a second array, both written by k_tasks)
    uint32_t end_off;
    uint32_t pad_[2];
    uint64_t tilepre_off;     // u32 offset of tilepre [p1_tiles][2^b1] in the count arena
    uint64_t m1_off;          // u32 offset of m1 [nrow + 1][2^b1] (region layout: counters [2^b1][nseg])
    uint64_t m2_off;          // u32 offset of m2 [nrow + 1][2^(b1+b2)] (2-pass jobs; region layout: counters [2^(b1+b2)])
    // Resident build sides. res_base != null: the job is already partitioned
    // (fin_* point at a resident table, res_base at its 2^(b1+b2)+1 bases) and
    // no hist / pass work is scheduled for it. keep_base != null: the bases
    // are also written there (the job's fin_* point at a table being kept).
    const uint32_t *res_base;
    uint32_t *keep_base;
};

struct JoinDesc {
    uint32_t job_r, job_s;
    uint32_t *out_r, *out_s;
    uint64_t cap;             // output capacity (pairs)
    uint32_t cap_r, cap_s;    // tuples per build chunk / probe chunk (powers of two)
    uint32_t ptask_off;       // u64 offset of the per-partition task prefix (nparts + 1)
    uint32_t nparts;          // 2^(b1+b2)
    uint32_t slot_shift;      // hash bits consumed by partitioning (slot uses h >> slot_shift)
    uint32_t one_based;
    uint32_t part_off;        // u64 offset of the task-prefix blocks' partials [ceil(nparts / kTaskBlock)][4]
    uint32_t pad_;
};

constexpr uint32_t kTaskBlock = 1024;  // partitions per block of the task-prefix kernels

constexpr int kColScanCols = 64;     // columns per block of the column-scan kernel

// One block of the column-scan kernel: kColScanCols columns of one job's m1 or m2.
struct ColScanDesc {
    uint32_t job;
    uint32_t which;  // 1: m1, 2: m2
    uint32_t col0;
    uint32_t pad;
};

// Control words: the ticket is reset by k_bases (k_region_reset in the region
// layout), the total written by k_tasks (one join) or k_scan2, before the
// build+probe kernel reads them. overflow: a region of the region layout
// received more rows than it holds (set by k_tasks).
struct Ctl {
    unsigned long long task_ticket;
    unsigned long long task_total;
    unsigned long long overflow;
};


struct PipeArgs {
    const JobDesc *jobs;
    const JoinDesc *joins;
    const ColScanDesc *colscan;     // [n_colscan]
    const uint32_t *hist_block0;    // [n_jobs + 1] prefix of nrow (host)
    const uint32_t *p1_tile0;       // [n_jobs + 1] prefix of p1_tiles (host)
    const uint32_t *p2_chunk0;      // [n_jobs + 1] prefix of pass-2 chunks (host)
    unsigned long long *join_tasks;   // [n_joins] (device)
    unsigned long long *task0;        // [n_joins + 1] (device)
    unsigned long long *out_count;    // [n_joins] (device)
    Ctl *ctl;
    uint32_t *counts;               // count arena: tilepre / m1 / m2 of every job
    uint32_t *base;
    unsigned long long *ptask;
    unsigned long long *taskmap;    // [taskmap_cap] task -> join << 32 | partition (tasks past it are searched)
    unsigned long long taskmap_cap;
    uint32_t max_nparts;            // partitions of the largest join
    uint32_t n_jobs, n_joins, n_colscan;
    uint32_t total_hist_blocks, total_p1_tiles, total_p2_chunks;
    uint32_t max_hist_bins;         // LDS bins of the histogram kernel
    uint32_t pass_tile;             // tuples per radix-pass tile (pass_tile(key bytes, tile shape))
    uint32_t regions;               // region layout (every job: two passes, 2^b1 a multiple of kXcds)
    HashCfg hash;
};

// Kernel launchers (vj_join_kernels.hip); all asynchronous on `s`.
struct LaunchShape {
    int hp, kw, nc;  // hash path, key width, key columns
    int cus;         // compute units of the device
    int tile, table; // kernel shapes (see pass_items / probe_cap)
};

// The hashing stages (histogram, radix pass, build+probe), compiled once per
// hash path in its own translation unit (vj_join_fnv.hip, vj_join_xorhw.hip,
// vj_join_generic.hip) from vj_join_stages.hpp.
template <int HP> struct StageLaunch {
    static hipError_t hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s);
    static hipError_t pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s);
    static hipError_t probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s);
    static int probe_blocks_per_cu(const LaunchShape &sh);
};
extern template struct StageLaunch<HP_FNV>;
extern template struct StageLaunch<HP_XOR_HW>;
extern template struct StageLaunch<HP_GENERIC>;
hipError_t launch_hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s);
hipError_t launch_scan(const PipeArgs &a, hipStream_t s);
hipError_t launch_region_reset(const PipeArgs &a, hipStream_t s);
hipError_t launch_tasks(const PipeArgs &a, hipStream_t s);
hipError_t launch_reset_probe(const PipeArgs &a, hipStream_t s);
hipError_t launch_stage_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s);
hipError_t launch_probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s);
int probe_blocks_per_cu(const LaunchShape &sh);
hipError_t launch_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets, hipStream_t s);

}  // namespace vj
