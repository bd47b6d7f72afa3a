// Descriptors of one join pipeline (a batch of one or more joins), shared by
// the host planner (vj_context.cpp) and the kernels (vj_join_kernels.hip).
//
// A join has two jobs: its build side (R) and its probe side (S). Each job is
// hashed and radix-partitioned on the same digits; the build+probe stage then
// draws (join, partition, R-chunk, S-chunk) tasks from a device work queue.
//
// HBM layout per partitioned job: SoA copies of the key columns plus a uint32
// row-id column, once after pass 1 ("mid", only for 2-pass jobs) and once
// after the last pass ("fin"). Partition q of a job occupies rows
// [base[q], base[q+1]) of "fin"; q = d1 * 2^b2 + d2 (pass-1 digit major).
#pragma once
#include <stdint.h>

#include "vj_device.hpp"

namespace vj {

constexpr int kPassThreads = 512;   // radix pass kernel
constexpr int kPassWaves = kPassThreads / 64;
constexpr int kProbeThreads = 512;  // build+probe kernel
constexpr int kProbeWaves = kProbeThreads / 64;
constexpr int kMaxPassBits = 8;   // digits per pass <= 256
constexpr int kMaxHistBits = 14;  // 2D histogram in LDS: 2^14 x 4 B = 64 KiB
constexpr int kOutBuf = 128;      // per-wave LDS output staging (pairs)

// Tuples per pass tile, per build chunk and per probe chunk, by key bytes per row.
__host__ __device__ constexpr int pass_items(int key_bytes) { return key_bytes <= 8 ? 8 : (key_bytes <= 16 ? 4 : 2); }
__host__ __device__ constexpr int pass_tile(int key_bytes) { return kPassThreads * pass_items(key_bytes); }
__host__ __device__ constexpr int build_cap(int key_bytes) { return key_bytes <= 8 ? 2048 : 1024; }
constexpr int kProbeChunk = 16384;  // probe rows per build+probe task

// Status word of the decoupled look-back: flag in bits 62..63, count in 0..31.
constexpr unsigned long long kFlagAgg = 1ull << 62;
constexpr unsigned long long kFlagIncl = 2ull << 62;

struct JobDesc {
    TView in;                 // the caller's table
    void *mid_cols[4];        // pass-1 output (2-pass jobs)
    uint32_t *mid_ids;
    void *fin_cols[4];        // last-pass output (partitioned jobs)
    uint32_t *fin_ids;
    uint32_t n;
    uint32_t b1, b2;          // radix bits of pass 1 / pass 2 (b1 == 0: not partitioned)
    uint32_t shift0;          // digit of pass 1 = (h >> shift0) & (2^b1 - 1)
    uint32_t hist_off;        // u32 offset of the 2^(b1+b2) histogram in the hist arena
    uint32_t base_off;        // u32 offset of the 2^(b1+b2)+1 partition bases in the base arena
    uint32_t p2pre_off;       // u32 offset of the 2^b1+1 pass-2 tile prefix (2-pass jobs)
    uint32_t hist_blocks;     // histogram blocks = pass-1 segments of this job
    uint32_t seg_tiles;       // pass-1 tiles per segment (segment k = tiles [k*seg_tiles, (k+1)*seg_tiles))
    uint32_t seg_off;         // offset of this job's segments in the per-segment ticket array
    uint32_t segc_off;        // offset of this job's (digit, segment) counts / bases, digit-major
    uint32_t p1_tiles;        // pass-1 tiles of this job
    uint32_t p2_tiles_max;    // upper bound of pass-2 tiles
    uint64_t p1_stat_off;     // u64 offset of pass-1 tile status (p1_tiles x 2^b1)
    uint64_t p2_stat_off;     // u64 offset of pass-2 tile status (p2_tiles_max x 2^b2)
};

struct JoinDesc {
    uint32_t job_r, job_s;
    uint32_t *out_r, *out_s;
    uint64_t cap;             // output capacity (pairs)
    uint32_t cap_r, cap_s;    // tuples per build chunk / probe chunk
    uint32_t ptask_off;       // u64 offset of the per-partition task prefix (nparts + 1)
    uint32_t nparts;          // 2^(b1+b2)
    uint32_t slot_shift;      // hash bits consumed by partitioning (slot uses h >> slot_shift)
    uint32_t one_based;
};

// Control words, reset by k_prepare at the start of every pipeline.
struct Ctl {
    unsigned int p2_total;
    unsigned int pad[3];
    unsigned long long task_ticket;
    unsigned long long task_total;
};

struct PipeArgs {
    const JobDesc *jobs;
    const JoinDesc *joins;
    const uint32_t *hist_block0;   // [n_jobs + 1] prefix of hist_blocks (host)
    const uint32_t *p1_tile0;      // [n_jobs + 1] prefix of p1_tiles (host)
    uint32_t *job_p2_tiles;        // [n_jobs] (device)
    uint32_t *p2_tile0;            // [n_jobs + 1] (device)
    unsigned long long *join_tasks;  // [n_joins] (device)
    unsigned long long *task0;       // [n_joins + 1] (device)
    unsigned long long *out_count;   // [n_joins] (device)
    uint32_t *seg_ticket;          // [total segments] pass-1 ticket per segment (device)
    uint32_t *p2_ticket;           // [p2pre words] pass-2 ticket per pass-1 partition (device)
    uint32_t *segcnt;              // [total segments x 2^b1] digit counts per segment, digit-major
    uint32_t *segbase;             // same shape: output base of each (digit, segment)
    void *zero_begin;              // per-pipeline counters cleared by k_prepare
    uint64_t zero_words;
    Ctl *ctl;
    uint32_t *hist;
    uint32_t *base;
    uint32_t *p2pre;
    unsigned long long *ptask;
    unsigned long long *status;
    uint64_t status_words;         // zeroed by the histogram kernel
    uint32_t n_jobs, n_joins;
    uint32_t total_hist_blocks, total_p1_tiles, total_p2_tiles_max;
    uint32_t max_hist_bins;        // LDS bins of the histogram kernel
    uint32_t pass_tile;            // tuples per radix-pass tile (pass_tile(key bytes))
    HashCfg hash;
};

// Kernel launchers (vj_join_kernels.hip); all asynchronous on `s`.
struct LaunchShape {
    int hp, kw, nc;  // hash path, key width, key columns
    int cus;         // compute units of the device
};
hipError_t launch_prepare(const PipeArgs &a, hipStream_t s);
hipError_t launch_hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s);
hipError_t launch_scan(const PipeArgs &a, hipStream_t s);
hipError_t launch_reset_probe(const PipeArgs &a, hipStream_t s);
hipError_t launch_pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s);
hipError_t launch_probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s);
int probe_blocks_per_cu(const LaunchShape &sh);
hipError_t launch_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets, hipStream_t s);

}  // namespace vj
