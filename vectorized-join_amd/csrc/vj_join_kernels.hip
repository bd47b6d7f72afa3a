// The join pipeline's kernels (gfx950). One pipeline runs any number of joins;
// every kernel draws its tiles from all of them.
//
//   k_hist      per segment of a job: per-tile pass-1 digit prefixes, the
//               segment's pass-1 and (pass-1, pass-2) digit counts (LDS)
//   k_colscan   exclusive scans down the segment columns of those counts
//   k_bases     per job: partition bases from the column totals; block 0
//               also resets the task queue and the output counters
//   k_tasks, k_task_offsets  per join, one block per 1024 partitions:
//               build/probe task prefix over partition pairs (one join: also
//               the queue totals and the task map); in the region layout also
//               the overflow check
//   k_region_reset  region layout: zeroes the region counters (replaces
//               k_hist, k_colscan and k_bases)
//   k_scan2     task prefix across joins                 (1 block, >1 join)
//   k_taskmap   task -> (join, partition) for the first tasks  (>1 join)
//   k_pass      one radix pass over a tile (pass 1) or a (digit, segment)
//               chunk (pass 2): per-wave LDS digit ranks -> LDS staging
//               ordered by digit -> contiguous runs per digit written at the
//               precomputed destinations (region layout: at runs reserved with
//               one device atomic per digit); no block waits on another
//   k_probe     persistent; (join, partition, R-chunk, S-chunk) tasks from a
//               device queue: open-addressing table of the R chunk in LDS,
//               probe with the S chunk; matches counted first, one output
//               reservation per task, then written with 64-lane ballot
//               compaction straight into contiguous output rows
//
// The reference's counterparts are histogram_kernel / prefix_sum_kernel /
// element_swap_kernel (src/join/join_provider.cu:315-380) and the build /
// probe / extract kernels (:382-608). Digits are LSB-first on the hash as in
// the reference (:303-305, :326, :362); the table slot uses the hash bits
// above the consumed digits, like the reference's key_offset (:394-395, :409);
// a pair matches iff every key column is equal (:422-437).//
// This file holds the scan kernels and the launchers; the hashing stages
// (k_hist, k_pass, k_probe) are in vj_join_stages.hpp, compiled once per hash
// path (vj_join_fnv.hip, vj_join_xorhw.hip, vj_join_generic.hip).
#include "vj_kernel_common.hpp"

namespace vj {

namespace {

// Exclusive scan of n <= 64 * kScanRounds * (blockDim.x / 64) values by one
// block. Wave w owns a contiguous span of the elements; in round j lane l
// holds element span * w + 64 * j + l (Spans::elem), so every load and store
// of a round is one contiguous run. The caller fills v[j] (0 where !ok(j));
// put(e, prefix, value) consumes the results. Returns the total; block-wide.
constexpr int kScanRounds = 16;
struct Spans {
    uint32_t e0, span, n;
    __device__ explicit Spans(uint32_t n_) : n(n_) {
        const uint32_t nw = blockDim.x >> 6;
        span = ((n + nw - 1) / nw + 63) & ~63u;
        e0 = (threadIdx.x >> 6) * span + (threadIdx.x & 63);
    }
    __device__ uint32_t elem(int j) const { return e0 + (uint32_t)j * 64; }
    __device__ bool ok(int j) const { return (uint32_t)j * 64 < span && elem(j) < n; }
};
template <typename T, typename Put>
__device__ T block_scan_spans(const Spans &sp, const T (&v)[kScanRounds], T *sh, Put put) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T sum = 0;
#pragma unroll
    for (int j = 0; j < kScanRounds; j++) sum += v[j];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if (lane == 0) sh[w] = sum;
    __syncthreads();
    T carry = 0, total = 0;
    for (uint32_t k = 0; k < nw; k++) {
        const T x = sh[k];
        carry += k < w ? x : T(0);
        total += x;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanRounds; j++) {
        if ((uint32_t)j * 64 >= sp.span) break;  // wave-uniform
        const T inc = wave_incl_scan(v[j]);
        if (sp.ok(j)) put(sp.elem(j), carry + inc - v[j], v[j]);
        carry += __shfl(inc, 63, 64);
    }
    return total;
}

// Exclusive scan of in[0..n) into out[0..n], total into out[n]; in may equal
// out. Each thread takes a contiguous run of elements, read in register
// chunks of kScanChunk so the loads of a chunk are in flight together.
constexpr int kScanChunk = 16;
template <typename T> __device__ void block_scan_array(const T *in_, T *out_, uint32_t n, T *sh) {
    const gptr<const T> in = to_global(in_);
    const gptr<T> out = to_global(out_);
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = threadIdx.x * per < n ? threadIdx.x * per : n;
    const uint32_t hi = lo + per < n ? lo + per : n;
    T sum = 0;
    for (uint32_t b = lo; b < hi; b += kScanChunk) {
        T v[kScanChunk];
#pragma unroll
        for (int r = 0; r < kScanChunk; r++) v[r] = b + r < hi ? in[b + r] : T(0);
#pragma unroll
        for (int r = 0; r < kScanChunk; r++) sum += v[r];
    }
    T total;
    T pre = block_excl_scan<T>(sum, sh, &total);
    for (uint32_t b = lo; b < hi; b += kScanChunk) {
        T v[kScanChunk];  // the whole chunk is read before any of it is written (in == out)
#pragma unroll
        for (int r = 0; r < kScanChunk; r++) v[r] = b + r < hi ? in[b + r] : T(0);
#pragma unroll
        for (int r = 0; r < kScanChunk; r++) {
            if (b + r < hi) out[b + r] = pre;
            pre += v[r];
        }
    }
    if (threadIdx.x == 0) out[n] = total;
    __syncthreads();
}


// Exclusive scans down kColScanCols columns of a [nrow + 1][ncols] count
// matrix (the last row receives the column totals). 1024 threads = 64 columns
// x 16 row groups: each thread reads its rows once into registers (all loads
// in flight together), the groups' sums are combined in LDS, and the rows are
// rewritten with their prefixes.
constexpr int kColScanGroups = 1024 / kColScanCols;
constexpr int kColScanRows = (kMaxSegments + kColScanGroups - 1) / kColScanGroups;
__global__ __launch_bounds__(1024) void k_colscan(PipeArgs a) {
    __shared__ uint32_t part[kColScanGroups][kColScanCols];
    const ColScanDesc d = a.colscan[blockIdx.x];
    const JobDesc &job = a.jobs[d.job];
    const uint32_t ncols = d.which == 1 ? (1u << job.b1) : (1u << (job.b1 + job.b2));
    const gptr<uint32_t> m = to_global(a.counts + (d.which == 1 ? job.m1_off : job.m2_off));
    const uint32_t cc = threadIdx.x % kColScanCols, rg = threadIdx.x / kColScanCols;
    const uint32_t c = d.col0 + cc;
    const uint32_t rows = job.nrow, per = (rows + kColScanGroups - 1) / kColScanGroups;  // <= kColScanRows
    const uint32_t lo = rg * per < rows ? rg * per : rows, hi = lo + per < rows ? lo + per : rows;
    const bool col_ok = c < ncols;
    const uint32_t cl = col_ok ? c : 0u;
    uint32_t v[kColScanRows];
    uint32_t sum = 0;
#pragma unroll
    for (int r = 0; r < kColScanRows; r++) {  // rows past hi read row lo (<= rows: in bounds) and count 0
        const bool ok = col_ok && lo + r < hi;
        const uint32_t x = m[(uint64_t)(ok ? lo + r : lo) * ncols + cl];
        v[r] = ok ? x : 0u;
        sum += v[r];
    }
    part[rg][cc] = sum;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < rg; k++) pre += part[k][cc];
    if (col_ok) {
#pragma unroll
        for (int r = 0; r < kColScanRows; r++) {
            if (lo + r < hi) m[(uint64_t)(lo + r) * ncols + c] = pre;
            pre += v[r];
        }
        if (rg == kColScanGroups - 1) m[(uint64_t)rows * ncols + c] = pre;
    }
}

// One block of 1024 threads per job: partition bases = exclusive scan of the
// column totals (the last row of m2, or of m1 for one pass).
// Block 0 also clears the task queue ticket and the output counters for the
// build+probe kernel, which runs after this one.
__global__ __launch_bounds__(1024) void k_bases(PipeArgs a) {
    __shared__ uint32_t sh32[20];
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            a.ctl->task_ticket = 0;
            a.ctl->overflow = 0;
        }
        for (uint32_t j = threadIdx.x; j < a.n_joins; j += blockDim.x) a.out_count[j] = 0;
    }
    if (blockIdx.x >= a.n_jobs) return;
    const JobDesc &job = a.jobs[blockIdx.x];
    const gptr<uint32_t> base = to_global(a.base + job.base_off);
    const gptr<uint32_t> keep = to_global(job.keep_base);  // null: not kept
    if (job.b1 == 0) {
        if (threadIdx.x == 0) {
            base[0] = 0;
            base[1] = job.n;
        }
        return;
    }
    const uint32_t nb = 1u << (job.b1 + job.b2);
    if (job.res_base) {  // resident table: its bases were computed when it was built
        const gptr<const uint32_t> rb = to_global(job.res_base);
        uint32_t x[kScanChunk + 1];  // nb + 1 <= (kScanChunk + 1) * blockDim.x; loads first, then stores
#pragma unroll
        for (int r = 0; r <= kScanChunk; r++) {
            const uint32_t q = threadIdx.x + r * blockDim.x;
            x[r] = rb[q <= nb ? q : nb];
        }
#pragma unroll
        for (int r = 0; r <= kScanChunk; r++) {
            const uint32_t q = threadIdx.x + r * blockDim.x;
            if (q <= nb) base[q] = x[r];
        }
        return;
    }
    if (job.n == 0) {  // an empty job has no count rows
        for (uint32_t q = threadIdx.x; q <= nb; q += blockDim.x) {
            base[q] = 0;
            if (keep) keep[q] = 0;
        }
    } else {
        const gptr<const uint32_t> tot = to_global(a.counts + (job.b2 ? job.m2_off : job.m1_off) + (uint64_t)job.nrow * nb);
        const Spans sp(nb);  // nb <= 2^14
        uint32_t v[kScanRounds];
#pragma unroll
        for (int j = 0; j < kScanRounds; j++) {
            const uint32_t x = tot[sp.ok(j) ? sp.elem(j) : 0u];
            v[j] = sp.ok(j) ? x : 0u;
        }
        const uint32_t total = block_scan_spans<uint32_t>(sp, v, sh32, [&](uint32_t e, uint32_t pre, uint32_t) {
            base[e] = pre;
            if (keep) keep[e] = pre;
        });
        if (threadIdx.x == 0) {
            base[nb] = total;
            if (keep) keep[nb] = total;
        }
    }
}

// One block of 1024 threads per join: build/probe tasks = R chunks x S chunks
// of every partition pair, as an exclusive prefix over partitions.
// The chunk capacities are powers of two (cap = 1 << log2), so the chunk
// counts are shifts.
__device__ __forceinline__ unsigned long long part_tasks(uint32_t nr, uint32_t ns, uint32_t lr, uint32_t ls) {
    return nr && ns ? (unsigned long long)(((nr - 1) >> lr) + 1) * (((ns - 1) >> ls) + 1) : 0ull;
}

// Task prefix, first kernel: one block per kTaskBlock partitions of a join.
// Per partition the task count (R chunks x S chunks), its exclusive prefix
// inside the block (into ptask), and per block the total and, in the region
// layout, the build and probe row sums and whether a counter passed its
// region (the block's partials); the region layout's partition first / end
// rows are written here too.
__global__ __launch_bounds__(1024) void k_tasks(PipeArgs a) {
    __shared__ unsigned long long sh64[kTaskBlock / 64];
    __shared__ unsigned long long rows[2];
    __shared__ int bad;
    const uint32_t jn = blockIdx.x, blk = blockIdx.y, tid = threadIdx.x;
    const JoinDesc &jd = a.joins[jn];
    const uint32_t n = jd.nparts;
    if (blk * kTaskBlock >= n) return;  // block-uniform
    const JobDesc &jr = a.jobs[jd.job_r], &js = a.jobs[jd.job_s];
    const uint32_t q = blk * kTaskBlock + tid;
    const bool ok = q < n;
    const uint32_t qi = ok ? q : 0u;
    const bool regions = jr.region1 != 0;  // both jobs of a join share the layout
    uint32_t nr, ns;
    if (!regions) {  // exact layout: n + 1 bases per job
        const gptr<const uint32_t> rb = to_global(a.base + jr.base_off), sb = to_global(a.base + js.base_off);
        const uint32_t r0 = rb[qi], r1 = rb[qi + 1], s0 = sb[qi], s1 = sb[qi + 1];
        nr = r1 - r0;
        ns = s1 - s0;
    } else {  // region layout: the final region counters
        nr = to_global(a.counts + jr.m2_off)[qi];
        ns = to_global(a.counts + js.m2_off)[qi];
    }
    if (!ok) nr = ns = 0;
    const gptr<unsigned long long> part = to_global(a.ptask + jd.part_off + (uint64_t)blk * 4);
    if (regions) {
        if (tid == 0) {
            rows[0] = rows[1] = 0;
            bad = 0;
        }
        __syncthreads();
        unsigned long long sr = nr, ss = ns;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            sr += __shfl_xor(sr, d, 64);
            ss += __shfl_xor(ss, d, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&rows[0], sr);
            atomicAdd(&rows[1], ss);
        }
        if (nr > jr.region2 || ns > js.region2) bad = 1;
        __syncthreads();
        if (tid == 0) {
            part[1] = rows[0];
            part[2] = rows[1];
            part[3] = (unsigned long long)bad;
        }
        nr = nr < jr.region2 ? nr : jr.region2;
        ns = ns < js.region2 ? ns : js.region2;
        if (ok) {
            to_global(a.base + jr.base_off)[q] = q * jr.region2;
            to_global(a.base + jr.end_off)[q] = q * jr.region2 + nr;
            to_global(a.base + js.base_off)[q] = q * js.region2;
            to_global(a.base + js.end_off)[q] = q * js.region2 + ns;
        }
    }
    const uint32_t lr = (uint32_t)__builtin_ctz(jd.cap_r), ls = (uint32_t)__builtin_ctz(jd.cap_s);
    const unsigned long long t = ok ? part_tasks(nr, ns, lr, ls) : 0ull;
    unsigned long long total;
    const unsigned long long pre = block_excl_scan<unsigned long long, kTaskBlock / 64>(t, sh64, &total);
    if (ok) to_global(a.ptask + jd.ptask_off)[q] = pre;
    if (tid == 0) part[0] = total;
}

// Task prefix, second kernel (same grid): adds the totals of the blocks
// before each block, writes the join's task total (one join: also the queue
// totals and the task -> partition map). In the region layout the join
// overflowed when a counter passed its region or the counters miss rows (a
// run that did not fit its pass-1 region never reached pass 2); it then gets
// no tasks, Ctl::overflow is set, and the host runs it again with the exact
// layout.
__global__ __launch_bounds__(1024) void k_task_offsets(PipeArgs a) {
    const uint32_t jn = blockIdx.x, blk = blockIdx.y, tid = threadIdx.x;
    const JoinDesc &jd = a.joins[jn];
    const uint32_t n = jd.nparts, nblk = (n + kTaskBlock - 1) / kTaskBlock;
    if (blk >= nblk) return;  // block-uniform
    const JobDesc &jr = a.jobs[jd.job_r], &js = a.jobs[jd.job_s];
    const gptr<const unsigned long long> part = to_global(a.ptask + jd.part_off);
    unsigned long long off = 0, total = 0, sr = 0, ss = 0, bad = 0;
    for (uint32_t b = 0; b < nblk; b++) {  // at most 16 blocks' partials
        const unsigned long long x = part[b * 4];
        off += b < blk ? x : 0ull;
        total += x;
        sr += part[b * 4 + 1];
        ss += part[b * 4 + 2];
        bad |= part[b * 4 + 3];
    }
    const bool over = jr.region1 != 0 && (bad != 0 || sr != jr.n || ss != js.n);
    const bool one = a.n_joins == 1;
    const gptr<unsigned long long> pt = to_global(a.ptask + jd.ptask_off);
    const uint32_t q = blk * kTaskBlock + tid;
    // this partition's in-block prefix and the next one's (the block total after the last)
    unsigned long long pre = 0, next = 0;
    if (q < n) {
        pre = pt[q];
        next = tid + 1 < kTaskBlock && q + 1 < n ? pt[q + 1] : part[blk * 4];
    }
    __syncthreads();  // every in-block prefix is read before any is rewritten
    if (q < n) {
        pt[q] = over ? 0ull : pre + off;
        if (one && !over) {
            const gptr<unsigned long long> tmap = to_global(a.taskmap);
            for (unsigned long long k = pre + off; k < next + off && k < a.taskmap_cap; k++) tmap[k] = q;
        }
    }
    if (blk == 0 && tid == 0) {
        const unsigned long long tasks = over ? 0ull : total;
        pt[n] = tasks;
        a.join_tasks[jn] = tasks;
        if (over) a.ctl->overflow = 1;
        if (one) {
            a.task0[0] = 0;
            a.task0[1] = tasks;
            a.ctl->task_total = tasks;
        }
    }
}

// Region layout, before pass 1: zeroes every job's region counters, the task
// queue ticket, the overflow flag and the output counters (one block per job).
__global__ __launch_bounds__(1024) void k_region_reset(PipeArgs a) {
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            a.ctl->task_ticket = 0;
            a.ctl->overflow = 0;
        }
        for (uint32_t j = threadIdx.x; j < a.n_joins; j += blockDim.x) a.out_count[j] = 0;
    }
    if (blockIdx.x >= a.n_jobs) return;
    const JobDesc &job = a.jobs[blockIdx.x];
    if (!job.region1) return;
    const gptr<uint32_t> m1 = to_global(a.counts + job.m1_off), m2 = to_global(a.counts + job.m2_off);
    const uint32_t n1 = job.nseg << job.b1, nb = 1u << (job.b1 + job.b2);
    for (uint32_t i = threadIdx.x; i < n1; i += blockDim.x) m1[i] = 0;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) m2[i] = 0;
}

// Copies 8-byte words between device memory and the context's pinned staging
// buffer (device-accessible host memory). A captured pipeline uploads its
// descriptors and downloads its counts with it instead of copy nodes, so a
// replay never leaves the compute queue.
__global__ __launch_bounds__(256) void k_stage_copy(unsigned long long *dst, const unsigned long long *src,
                                                    uint64_t words) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Resets the queue and the output counters to run build+probe again (after
// an output buffer was found too small).
__global__ void k_reset_probe(PipeArgs a) {
    if (threadIdx.x == 0) a.ctl->task_ticket = 0;
    for (uint32_t j = threadIdx.x; j < a.n_joins; j += blockDim.x) a.out_count[j] = 0;
}

// Task -> (join, partition) for the first taskmap_cap tasks, so the
// build+probe kernel finds a task's partition with one load.
__global__ __launch_bounds__(1024) void k_taskmap(PipeArgs a) {
    const uint32_t jn = blockIdx.x, q = blockIdx.y * blockDim.x + threadIdx.x;
    const JoinDesc &jd = a.joins[jn];
    if (q >= jd.nparts) return;
    const unsigned long long *pt = a.ptask + jd.ptask_off;
    const unsigned long long t0 = a.task0[jn];
    const unsigned long long lo = t0 + pt[q], hi = t0 + pt[q + 1];
    const unsigned long long e = ((unsigned long long)jn << 32) | q;
    for (unsigned long long t = lo; t < hi && t < a.taskmap_cap; t++) a.taskmap[t] = e;
}

__global__ __launch_bounds__(1024) void k_scan2(PipeArgs a) {
    __shared__ unsigned long long sh64[20];
    block_scan_array<unsigned long long>(a.join_tasks, a.task0, a.n_joins, sh64);
    if (threadIdx.x == 0) a.ctl->task_total = a.task0[a.n_joins];
}


__global__ void k_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
        offsets[i] = base[i];
        hist[i] = base[i + 1] - base[i];
    }
}

}  // namespace

hipError_t launch_hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s) {
    if (a.total_hist_blocks == 0) return hipSuccess;
    switch (sh.hp) {
    case HP_FNV: return StageLaunch<HP_FNV>::hist(a, sh, s);
    case HP_XOR_HW: return StageLaunch<HP_XOR_HW>::hist(a, sh, s);
    case HP_GENERIC: return StageLaunch<HP_GENERIC>::hist(a, sh, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_scan(const PipeArgs &a, hipStream_t s) {
    hipError_t e;
    if (a.n_colscan) {
        hipLaunchKernelGGL(k_colscan, dim3(a.n_colscan), dim3(1024), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bases, dim3(a.n_jobs ? a.n_jobs : 1), dim3(1024), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_tasks(a, s);
}

hipError_t launch_region_reset(const PipeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_region_reset, dim3(a.n_jobs ? a.n_jobs : 1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tasks(const PipeArgs &a, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (a.n_joins) {
        const dim3 grid(a.n_joins, (a.max_nparts + kTaskBlock - 1) / kTaskBlock);
        hipLaunchKernelGGL(k_tasks, grid, dim3(kTaskBlock), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_task_offsets, grid, dim3(kTaskBlock), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (a.n_joins == 1) return hipSuccess;  // k_tasks wrote the queue totals and the task map
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.n_joins && a.taskmap_cap) {
        hipLaunchKernelGGL(k_taskmap, dim3(a.n_joins, (a.max_nparts + 1023) / 1024), dim3(1024), 0, s, a);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_stage_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes % 8) return hipErrorInvalidValue;
    const uint64_t words = bytes / 8;
    if (!words) return hipSuccess;
    const uint64_t blocks = (words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_stage_copy, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<unsigned long long *>(dst),
                       static_cast<const unsigned long long *>(src), words);
    return hipGetLastError();
}

hipError_t launch_reset_probe(const PipeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_probe, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s) {
    if ((pass == 1 ? a.total_p1_tiles : a.total_p2_chunks) == 0) return hipSuccess;
    switch (sh.hp) {
    case HP_FNV: return StageLaunch<HP_FNV>::pass(a, sh, pass, s);
    case HP_XOR_HW: return StageLaunch<HP_XOR_HW>::pass(a, sh, pass, s);
    case HP_GENERIC: return StageLaunch<HP_GENERIC>::pass(a, sh, pass, s);
    }
    return hipErrorInvalidValue;
}

int probe_blocks_per_cu(const LaunchShape &sh) {
    switch (sh.hp) {
    case HP_FNV: return StageLaunch<HP_FNV>::probe_blocks_per_cu(sh);
    case HP_XOR_HW: return StageLaunch<HP_XOR_HW>::probe_blocks_per_cu(sh);
    case HP_GENERIC: return StageLaunch<HP_GENERIC>::probe_blocks_per_cu(sh);
    }
    return 1;
}

hipError_t launch_probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s) {
    switch (sh.hp) {
    case HP_FNV: return StageLaunch<HP_FNV>::probe(a, sh, blocks, s);
    case HP_XOR_HW: return StageLaunch<HP_XOR_HW>::probe(a, sh, blocks, s);
    case HP_GENERIC: return StageLaunch<HP_GENERIC>::probe(a, sh, blocks, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets, hipStream_t s) {
    unsigned blocks = (nb + 255) / 256;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_emit_hist, dim3(blocks), dim3(256), 0, s, base, nb, hist, offsets);
    return hipGetLastError();
}

}  // namespace vj
