// The join pipeline's kernels (gfx950). One pipeline runs any number of joins;
// every kernel draws its tiles from all of them.
//
//   k_hist      per segment of a job: per-tile pass-1 digit prefixes, the
//               segment's pass-1 and (pass-1, pass-2) digit counts (LDS)
//   k_colscan   exclusive scans down the segment columns of those counts
//   k_bases     per job: partition bases from the column totals; block 0
//               also resets the task queue and the output counters
//   k_tasks     per join: build/probe task prefix over partition pairs (one
//               join: also the queue totals and the task map)
//   k_scan2     task prefix across joins                 (1 block, >1 join)
//   k_taskmap   task -> (join, partition) for the first tasks  (>1 join)
//   k_pass      one radix pass over a tile (pass 1) or a (digit, segment)
//               chunk (pass 2): per-wave LDS digit ranks -> LDS staging
//               ordered by digit -> contiguous runs per digit written at the
//               precomputed destinations; no block waits on another
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
// a pair matches iff every key column is equal (:422-437).
#include "vj_pipeline.hpp"

namespace vj {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Wave-uniform values read from LDS, moved into scalar registers so address
// arithmetic and loop bounds use SGPRs. readfirstlane returns a signed int:
// it is taken as uint32_t before any widening, so a 64-bit value's halves are
// zero-extended (sign extension would corrupt the high word).
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t uniform(uint64_t v) {
    return (uint64_t)uniform((uint32_t)v) | ((uint64_t)uniform((uint32_t)(v >> 32)) << 32);
}
template <typename P> __device__ __forceinline__ P *uniform(P *p) {
    return reinterpret_cast<P *>(uniform((uint64_t)reinterpret_cast<uintptr_t>(p)));
}
// A wave-uniform generic pointer as a global one in scalar registers.
template <typename P> __device__ __forceinline__ gptr<P> uniform_global(P *p) { return to_global(uniform(p)); }
__device__ __forceinline__ GView uniform_view(const TView &t) {
    GView g;
#pragma unroll
    for (int c = 0; c < 4; c++) g.cols[c] = uniform_global(t.cols[c]);
    g.rowids = uniform_global(t.rowids);
    return g;
}
template <int KW, int NC> __device__ __forceinline__ GKeys<KW, NC> uniform_keys(const TView &t) {
    GKeys<KW, NC> g;
#pragma unroll
    for (int c = 0; c < NC; c++) g.c[c] = uniform_global(static_cast<const typename KeyT<KW>::T *>(t.cols[c]));
    return g;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Largest i in [0, n) with a[i] <= x, for non-decreasing a with a[0] <= x.
template <typename T> __device__ __forceinline__ uint32_t upper_index(const T *a, uint32_t n, T x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Same search done by a whole wave: 64 probes per round, so an array of
// 2^16 entries needs three dependent loads instead of sixteen.
template <typename T> __device__ __forceinline__ uint32_t wave_upper_index(const T *a, uint32_t n, T x) {
    const uint32_t lane = lane_id();
    uint32_t lo = 0, len = n;
    while (len > 1) {
        const uint32_t step = (len + 63) >> 6;
        const uint32_t i = lo + lane * step;
        const bool ok = lane == 0 || (i < lo + len && a[i] <= x);
        const uint64_t m = __ballot(ok);
        const uint32_t last = 63 - __builtin_clzll(m);
        const uint32_t end = lo + len;
        lo = lo + last * step;
        len = (end - lo) < step ? (end - lo) : step;
    }
    return lo;
}

template <typename T> __device__ __forceinline__ T wave_incl_scan(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += o;
    }
    return v;
}

// Exclusive scan over the block; `sh` holds >= blockDim/64 + 1 values.
template <typename T> __device__ T block_excl_scan(T v, T *sh, T *total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    T inc = wave_incl_scan(v);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (uint32_t i = 0; i < nw; i++) {
            T t = sh[i];
            sh[i] = acc;
            acc += t;
        }
        sh[nw] = acc;
    }
    __syncthreads();
    T r = sh[w] + inc - v;
    *total = sh[nw];
    __syncthreads();
    return r;
}

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

// ---------------------------------------------------------------------------

// One block per segment. The block sweeps the segment one pass-1 tile at a
// time (all 1024 threads on one tile, coalesced), with one LDS atomic per row
// into the tile's pass-1 digit counters and, for 2-pass jobs, one into the
// segment's (pass-1, pass-2) histogram. Tile counters live in a ring of
// kHistRing tiles; every kHistRing tiles the ring is folded into the running
// in-segment prefix, which is stored per tile (tilepre).
constexpr int kHistRing = kHistThreads / 64;

template <int HP, int KW, int NC>
__global__ __launch_bounds__(kHistThreads) void k_hist(PipeArgs a) {
    constexpr int KB = KW * NC;
    constexpr int T = pass_tile(KB);
    constexpr int U = T / kHistThreads;
    static_assert(U * kHistThreads == T, "a tile is a whole number of block sweeps");
    extern __shared__ uint32_t lds[];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint32_t j = upper_index(a.hist_block0, a.n_jobs, b);
    const JobDesc &job = a.jobs[j];
    const uint32_t g = b - a.hist_block0[j];
    const uint32_t n1 = 1u << job.b1, nb = 1u << (job.b1 + job.b2);
    const bool two = job.b2 > 0;
    uint32_t *h2 = lds;                    // [nb] (2-pass jobs)
    uint32_t *run = lds + (two ? nb : 0);  // [n1] running in-segment prefix
    uint32_t *th = run + n1;               // [kHistRing][n1] tile counts
    for (uint32_t i = tid; i < (two ? nb : 0) + n1 * (1 + kHistRing); i += kHistThreads) lds[i] = 0;
    __syncthreads();
    const uint32_t m1 = n1 - 1, m2 = (1u << job.b2) - 1, b2 = job.b2;
    const uint32_t s1 = job.shift0, s2 = job.shift0 + job.b1;
    const HashCfg hc = a.hash;
    // histogram row g: the hrow_tiles tiles of segment g / hsplit starting at
    // its (g % hsplit)-th row (the last rows of a segment may be short or empty)
    const uint32_t sg = g / job.hsplit;
    const uint32_t s_end0 = (sg + 1) * job.seg_tiles;
    const uint32_t s_end = s_end0 < job.p1_tiles ? s_end0 : job.p1_tiles;
    const uint32_t r0 = sg * job.seg_tiles + (g - sg * job.hsplit) * job.hrow_tiles;
    const uint32_t t_begin = r0 < s_end ? r0 : s_end;
    const uint32_t t_end = t_begin + job.hrow_tiles < s_end ? t_begin + job.hrow_tiles : s_end;
    const gptr<uint32_t> tilepre = to_global(a.counts + job.tilepre_off);
    const GKeys<KW, NC> in = uniform_keys<KW, NC>(job.in);
    const uint32_t n = job.n;
    for (uint32_t t0 = t_begin; t0 < t_end; t0 += kHistRing) {
        const uint32_t kt = t_end - t0 < (uint32_t)kHistRing ? t_end - t0 : (uint32_t)kHistRing;
        for (uint32_t k = 0; k < kt; k++) {
            uint32_t *thk = th + k * n1;
            const uint32_t row_lo = (t0 + k) * T + tid;
            Key<KW, NC> key[U];
#pragma unroll
            for (int u = 0; u < U; u++) {  // all loads in flight at once (past the end: row 0 again)
                const uint32_t i = row_lo + u * kHistThreads;
                key[u] = load_key<KW, NC>(in, i < n ? i : 0u);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = row_lo + u * kHistThreads;
                if (i < n) {
                    const uint64_t h = hash_key<HP, KW, NC>(key[u], hc);
                    const uint32_t d1 = (uint32_t)(h >> s1) & m1;
                    atomicAdd(&thk[d1], 1u);
                    if (two) atomicAdd(&h2[(d1 << b2) | ((uint32_t)(h >> s2) & m2)], 1u);
                }
            }
        }
        __syncthreads();
        if (tid < n1) {
            uint32_t acc = run[tid];
            for (uint32_t k = 0; k < kt; k++) {
                tilepre[(uint64_t)(t0 + k) * n1 + tid] = acc;
                acc += th[k * n1 + tid];
                th[k * n1 + tid] = 0;
            }
            run[tid] = acc;
        }
        __syncthreads();
    }
    const gptr<uint32_t> m1row = to_global(a.counts + job.m1_off + (uint64_t)g * n1);
    for (uint32_t d = tid; d < n1; d += kHistThreads) m1row[d] = run[d];
    if (two) {
        const gptr<uint32_t> m2row = to_global(a.counts + job.m2_off + (uint64_t)g * nb);
        for (uint32_t q = tid; q < nb; q += kHistThreads) m2row[q] = h2[q];
    }
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
        if (threadIdx.x == 0) a.ctl->task_ticket = 0;
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

__global__ __launch_bounds__(1024) void k_tasks(PipeArgs a) {
    __shared__ unsigned long long sh64[20];
    const uint32_t jn = blockIdx.x;
    const JoinDesc &jd = a.joins[jn];
    const gptr<const uint32_t> rb = to_global(a.base + a.jobs[jd.job_r].base_off);
    const gptr<const uint32_t> sb = to_global(a.base + a.jobs[jd.job_s].base_off);
    const gptr<unsigned long long> pt = to_global(a.ptask + jd.ptask_off);
    const gptr<unsigned long long> tmap = to_global(a.taskmap);
    const uint32_t n = jd.nparts;  // <= 2^14; rb, sb hold n + 1 bases
    // One join: its task prefix is the whole queue, so the task -> partition
    // map and the queue totals are written here.
    const bool one = a.n_joins == 1;
    const unsigned long long map_cap = a.taskmap_cap;
    const Spans sp(n);
    // partition sizes: every base is loaded before the first division
    uint32_t nr[kScanRounds], ns[kScanRounds];
#pragma unroll
    for (int j = 0; j < kScanRounds; j++) {
        const uint32_t i = sp.ok(j) ? sp.elem(j) : 0u;
        nr[j] = rb[i + 1] - rb[i];
        ns[j] = sb[i + 1] - sb[i];
    }
    unsigned long long v[kScanRounds];
    const uint32_t lr = (uint32_t)__builtin_ctz(jd.cap_r), ls = (uint32_t)__builtin_ctz(jd.cap_s);
#pragma unroll
    for (int j = 0; j < kScanRounds; j++) v[j] = sp.ok(j) ? part_tasks(nr[j], ns[j], lr, ls) : 0ull;
    const unsigned long long total =
        block_scan_spans<unsigned long long>(sp, v, sh64, [&](uint32_t e, unsigned long long pre, unsigned long long t) {
            pt[e] = pre;
            if (one)
                for (unsigned long long q = pre; q < pre + t && q < map_cap; q++) tmap[q] = e;
        });
    if (threadIdx.x == 0) {
        pt[n] = total;
        a.join_tasks[jn] = total;
        if (one) {
            a.task0[0] = 0;
            a.task0[1] = total;
            a.ctl->task_total = total;
        }
    }
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

// ---------------------------------------------------------------------------
// Radix pass
// ---------------------------------------------------------------------------

struct PassTile {
    TView src;
    void *dst_cols[4];
    uint32_t *dst_ids;
    uint32_t lo, hi;        // rows of src (pass 2: the whole chunk)
    uint32_t bins, shift;
    uint32_t job, local;    // job and tile (pass 1) / chunk (pass 2) index inside it
    uint32_t valid;
};

// Waves per SIMD the pass kernel is compiled for: 8 (64 VGPRs, four blocks per
// CU) for single 8-byte keys with FNV / XOR hashing; wider keys and the
// generic hash path get 6 (80 VGPRs, three blocks) to limit spills.
template <int HP, int KB> constexpr int pass_waves_per_simd() { return KB <= 8 && HP != HP_GENERIC ? 8 : 6; }

// Pass 1: one tile of a job. Pass 2: one (pass-1 digit, segment) chunk, in
// tile-sized steps; a running count per digit carries the destination.
template <int HP, int KW, int NC>
__global__ __launch_bounds__(kPassThreads, (pass_waves_per_simd<HP, KW * NC>())) void k_pass(PipeArgs a, int pass) {
    using KT = typename KeyT<KW>::T;
    constexpr int KB = KW * NC;
    constexpr int ITEMS = pass_items(KB);
    constexpr int T = kPassThreads * ITEMS;
    constexpr int STAGE = (KB > 4 ? KB : 4) * T;  // key columns, then row ids, staged in turn
    extern __shared__ __align__(16) unsigned char smem[];
    KT *skeys = reinterpret_cast<KT *>(smem);                             // [NC][T]
    uint32_t *sids = reinterpret_cast<uint32_t *>(smem);                   // [T] (after the keys are out)
    uint32_t *wcnt = reinterpret_cast<uint32_t *>(smem + STAGE);           // [kPassWaves][256]
    uint32_t *loc = wcnt + kPassWaves * 256;                               // [256]
    uint32_t *gdst = loc + 256;                                            // [256] next destination row
    uint32_t *gdo = gdst + 256;                                            // [256] gdst - loc of the current tile
    uint8_t *sdig = reinterpret_cast<uint8_t *>(gdo + 256);                // [T]
    __shared__ PassTile ti;
    __shared__ uint32_t sh_scan[kPassWaves + 1];

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (w == 0) {
        // Workgroups are dispatched round-robin over the XCDs. Consecutive
        // tiles go to one XCD, so that the runs adjacent tiles write to one
        // partition (which share cache lines) meet in the same L2: pass 1
        // gives each XCD one contiguous range of tiles; pass 2 deals out
        // groups of kXcdGroup consecutive chunks round-robin, so that a run of
        // heavy chunks (a skewed partition) is spread over all XCDs (block k
        // of XCD x takes chunk (k / G * 8 + x) * G + k % G).
        const uint32_t *pfx = pass == 1 ? a.p1_tile0 : a.p2_chunk0;
        const uint32_t total = pass == 1 ? a.total_p1_tiles : a.total_p2_chunks;
        const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
        const uint32_t b = pass == 1 ? x * ((total + kXcds - 1) / kXcds) + k
                                     : ((k / kXcdGroup) * kXcds + x) * kXcdGroup + k % kXcdGroup;
        const uint32_t valid = b < total;
        const uint32_t j = valid ? wave_upper_index(pfx, a.n_jobs, b) : 0;
        if (lane == 0) {
            ti.valid = valid;
            if (valid) {
                const JobDesc &job = a.jobs[j];
                const uint32_t local = b - pfx[j];
                ti.job = j;
                ti.local = local;
                if (pass == 1) {
                    ti.src = job.in;
                    for (int c = 0; c < 4; c++) ti.dst_cols[c] = job.b2 ? job.mid_cols[c] : job.fin_cols[c];
                    ti.dst_ids = job.b2 ? job.mid_ids : job.fin_ids;
                    ti.lo = local * T;
                    ti.hi = ti.lo + T < job.n ? ti.lo + T : job.n;
                    ti.bins = 1u << job.b1;
                    ti.shift = job.shift0;
                } else {
                    const uint32_t n1 = 1u << job.b1;
                    const uint32_t d1 = local / job.nseg, g = local - d1 * job.nseg;
                    const uint32_t p = a.base[job.base_off + (d1 << job.b2)];
                    const uint32_t *m1 = a.counts + job.m1_off;
                    for (int c = 0; c < 4; c++) ti.src.cols[c] = job.mid_cols[c];
                    ti.src.rowids = job.mid_ids;
                    ti.src.n = job.n;
                    for (int c = 0; c < 4; c++) ti.dst_cols[c] = job.fin_cols[c];
                    ti.dst_ids = job.fin_ids;
                    ti.lo = p + m1[(uint64_t)g * job.hsplit * n1 + d1];
                    ti.hi = p + m1[(uint64_t)(g + 1) * job.hsplit * n1 + d1];
                    ti.bins = 1u << job.b2;
                    ti.shift = job.shift0 + job.b1;
                }
            }
        }
    }
    __syncthreads();
    if (!ti.valid) return;
    if (tid < ti.bins) {  // first destination row of every digit
        const JobDesc &job = a.jobs[ti.job];
        const uint32_t *base = a.base + job.base_off;
        if (pass == 1) {
            const uint32_t n1 = 1u << job.b1, t = ti.local, sg = t / job.seg_tiles;
            const uint32_t g = sg * job.hsplit + (t - sg * job.seg_tiles) / job.hrow_tiles;  // histogram row
            gdst[tid] = base[tid << job.b2] + a.counts[job.m1_off + (uint64_t)g * n1 + tid] +
                        a.counts[job.tilepre_off + (uint64_t)t * n1 + tid];
        } else {
            const uint32_t nb = 1u << (job.b1 + job.b2);
            const uint32_t d1 = ti.local / job.nseg, g = ti.local - d1 * job.nseg;
            const uint32_t q = (d1 << job.b2) | tid;
            gdst[tid] = base[q] + a.counts[job.m2_off + (uint64_t)g * job.hsplit * nb + q];
        }
    }

    const uint32_t bins = ti.bins, shift = ti.shift, mask = bins - 1;
    const GKeys<KW, NC> skeys_g = uniform_keys<KW, NC>(ti.src);  // source columns, in scalar registers
    const gptr<const uint32_t> sids_g = uniform_global(ti.src.rowids);
    gptr<KT> dcols[NC];  // destination columns
#pragma unroll
    for (int c = 0; c < NC; c++) dcols[c] = uniform_global(static_cast<KT *>(ti.dst_cols[c]));
    const gptr<uint32_t> dids = uniform_global(ti.dst_ids);
    const HashCfg hc = a.hash;
    for (uint32_t lo = ti.lo; lo < ti.hi; lo += T) {
        const uint32_t hi = lo + T < ti.hi ? lo + T : ti.hi;
        for (uint32_t i = tid; i < kPassWaves * 256; i += kPassThreads) wcnt[i] = 0;
        __syncthreads();

        // 1. load ITEMS rows per lane (wave w owns a contiguous 64*ITEMS segment), rank by digit
        Key<KW, NC> keys[ITEMS];
        uint32_t ids[ITEMS];
        uint32_t pk[ITEMS];  // digit | rank << 8, or ~0 for no row
        const uint32_t wbase = lo + w * 64 * ITEMS + lane;
        // every load issued before the first use (past the end: row lo again)
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint32_t i = wbase + it * 64;
            keys[it] = load_key<KW, NC>(skeys_g, i < hi ? i : lo);
        }
        if (sids_g) {  // wave-uniform
#pragma unroll
            for (int it = 0; it < ITEMS; it++) {
                const uint32_t i = wbase + it * 64;
                ids[it] = sids_g[i < hi ? i : lo];
            }
        } else {
#pragma unroll
            for (int it = 0; it < ITEMS; it++) ids[it] = wbase + it * 64;
        }
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            const uint32_t i = wbase + it * 64;
            const bool valid = i < hi;
            uint32_t d = 0;
            if (valid) d = (uint32_t)(hash_key<HP, KW, NC>(keys[it], hc) >> shift) & mask;
            const uint32_t r = valid ? atomicAdd(&wcnt[w * 256 + d], 1u) : 0u;
            pk[it] = valid ? (d | (r << 8)) : ~0u;
        }
        __syncthreads();

        // 2. per digit: wave prefixes, tile count, local offsets
        uint32_t cnt = 0;
        if (tid < bins) {
#pragma unroll
            for (int ww = 0; ww < kPassWaves; ww++) {
                uint32_t c = wcnt[ww * 256 + tid];
                wcnt[ww * 256 + tid] = cnt;
                cnt += c;
            }
        }
        uint32_t total;
        const uint32_t my_loc = block_excl_scan<uint32_t>(cnt, sh_scan, &total);
        if (tid < bins) {
            loc[tid] = my_loc;
            gdo[tid] = gdst[tid] - my_loc;  // output row of staged slot k = gdo[digit] + k
        }
        __syncthreads();

        // 3. stage the key columns in LDS ordered by digit; consecutive slots of
        //    one digit go to consecutive output rows
        uint32_t slot[ITEMS];
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            slot[it] = ~0u;
            if (pk[it] != ~0u) {
                const uint32_t d = pk[it] & 255u;
                slot[it] = loc[d] + wcnt[w * 256 + d] + (pk[it] >> 8);
#pragma unroll
                for (int c = 0; c < NC; c++) skeys[c * T + slot[it]] = keys[it].c[c];
                sdig[slot[it]] = (uint8_t)d;
            }
        }
        __syncthreads();
        // 4. write the keys in slot order
        const uint32_t n = hi - lo;
        for (uint32_t k = tid; k < n; k += kPassThreads) {
            const uint32_t dst = gdo[sdig[k]] + k;
#pragma unroll
            for (int c = 0; c < NC; c++) dcols[c][dst] = skeys[c * T + k];
        }
        __syncthreads();
        // 5. the row ids, through the same staging buffer
#pragma unroll
        for (int it = 0; it < ITEMS; it++)
            if (slot[it] != ~0u) sids[slot[it]] = ids[it];
        __syncthreads();
        for (uint32_t k = tid; k < n; k += kPassThreads) dids[gdo[sdig[k]] + k] = sids[k];
        __syncthreads();
        if (tid < bins) gdst[tid] += cnt;  // next step of a pass-2 chunk
    }
}

// ---------------------------------------------------------------------------
// Build + probe
// ---------------------------------------------------------------------------

struct ProbeTask {
    TView r, s;
    uint32_t *out_r, *out_s;
    unsigned long long *count;
    uint64_t cap;
    unsigned long long out_base;  // first output row of this task
    uint32_t r_lo, r_n, s_lo, s_n;
    uint32_t nslots, shift, one_based;
    uint32_t dup;
    uint32_t done;
    uint32_t wave_count[kProbeWaves];
};

constexpr int kProbeUnroll = 2;
constexpr int kBucketSlots = 4;  // LDS table: buckets of 4 slots (one 16-byte load)
constexpr uint16_t kNoMatch = 0xFFFF;

// One task's parameters for the walks (all wave-uniform).
struct WalkArgs {
    GView s;
    gptr<const uint32_t> r_ids;  // build rows' ids (null: the row index)
    uint32_t r_lo;
    gptr<uint32_t> out_r, out_s;
    uint64_t cap;
    uint32_t s_lo, s_end, bmask, shift, ob1;
    bool dup;
};

__device__ __forceinline__ WalkArgs walk_args(const ProbeTask &t) {
    WalkArgs w;
    w.s = uniform_view(t.s);
    w.r_ids = uniform_global(t.r.rowids);
    w.r_lo = uniform(t.r_lo);
    w.out_r = uniform_global(t.out_r);
    w.out_s = uniform_global(t.out_s);
    w.cap = uniform((uint64_t)t.cap);
    w.s_lo = uniform(t.s_lo);
    w.s_end = uniform(t.s_lo + t.s_n);
    w.bmask = uniform((t.nslots >> 2) - 1);
    w.shift = uniform(t.shift);
    w.ob1 = uniform(t.one_based);
    w.dup = uniform(t.dup) != 0;
    return w;
}

// Row id of build row o of the task's chunk, read from the partitioned build
// side (its rows were just loaded by the build step, so the read is cached).
__device__ __forceinline__ uint32_t build_id(const WalkArgs &wa, uint32_t o) {
    return wa.r_ids ? wa.r_ids[wa.r_lo + o] : wa.r_lo + o;
}

// Writes a wave's matches straight to the join output: the matching lanes of
// one ballot take consecutive rows from `pos`, so every store instruction
// covers one contiguous run of each output column.
struct Emitter {
    unsigned long long pos;
    __device__ __forceinline__ void put(bool m, uint32_t r, uint32_t s, const WalkArgs &wa) {
        const uint64_t mb = __ballot(m);
        if (m) {
            const unsigned long long o = pos + lanes_below(mb);
            if (o < wa.cap) {
                wa.out_r[o] = r;
                wa.out_s[o] = s;
            }
        }
        pos += __builtin_popcountll(mb);
    }
};

// Walks over this wave's rows of the S chunk. The chunk holds up to
// kProbeChunk = kProbeIters x kProbeThreads x kProbeUnroll rows; in iteration
// `it` lane `l` of wave `w` owns rows it*kProbeThreads*U + w*64*U + u*64 + l.
// The table holds one entry per distinct key (equal keys of the build chunk
// are chained behind it, see the build step), so a row matches at most one
// entry. probe_walk finds it, keeps its index in registers and counts the
// row's matches (the chain length); replay_walk then writes them without
// probing again. Probing runs in rounds: each round reads one whole bucket
// for every unfinished row of the lane (U independent LDS chains per lane);
// a round is branch-free (selects, no per-row branches). The S keys of the
// next iteration are loaded before the current one is probed, so one HBM round
// trip overlaps each iteration's probing (where registers allow, see PF).

template <int N> __device__ __forceinline__ uint32_t res_get(const uint32_t (&r)[N], int k) {
    return (r[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
}
template <int N> __device__ __forceinline__ void res_set(uint32_t (&r)[N], int k, uint32_t v) {
    const int sh = (k & 1) * 16;
    r[k >> 1] = (r[k >> 1] & ~(0xFFFFu << sh)) | (v << sh);
}
constexpr int kProbeIters = kProbeChunk / (kProbeThreads * kProbeUnroll);
static_assert(kProbeIters * kProbeThreads * kProbeUnroll == kProbeChunk, "probe chunk shape");

// First S row of this lane in iteration `it` of a walk with U rows per lane.
template <int U> __device__ __forceinline__ uint32_t walk_row(const WalkArgs &wa, int it) {
    return wa.s_lo + it * kProbeThreads * U + (threadIdx.x >> 6) * 64 * U + lane_id();
}
// First S row of the wave in iteration `it` (wave-uniform).
template <int U> __device__ __forceinline__ uint32_t walk_wave_row(const WalkArgs &wa, int it) {
    return wa.s_lo + it * kProbeThreads * U + uniform((uint32_t)(threadIdx.x >> 6)) * 64 * U;
}

// Loads this lane's S keys of iteration `it` (rows past the chunk are left alone).
template <int U, int KW, int NC>
__device__ __forceinline__ void walk_load(const WalkArgs &wa, int it, Key<KW, NC> (&k)[U]) {
    const uint32_t row = walk_row<U>(wa, it);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (row + u * 64 < wa.s_end) k[u] = load_key<KW, NC>(wa.s, row + u * 64);
}

template <int N> __device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <int HP, int KW, int NC>
__device__ __forceinline__ uint32_t probe_walk(const WalkArgs &wa, const HashCfg &hc, uint32_t tag_shift,
                                               const typename KeyT<KW>::T *rk, const uint32_t *slots,
                                               const uint16_t *nxt,
                                               uint32_t (&res)[kProbeIters * kProbeUnroll / 2]) {
    constexpr int CAP = build_cap(KW * NC);
    constexpr int U = kProbeUnroll;
    // Next-iteration loads double the key registers; they are used where the
    // kernel stays within its register budget (single 8-byte key, FNV / XOR).
    constexpr bool PF = KW * NC <= 8 && HP != HP_GENERIC;
    const uint32_t bmask = wa.bmask, shift = wa.shift, s_end = wa.s_end;
    uint32_t total = 0;  // this lane's matches
    Key<KW, NC> nk[U];
    walk_load<U, KW, NC>(wa, 0, nk);
#pragma unroll
    for (int it = 0; it < kProbeIters; it++) {
        if (walk_wave_row<U>(wa, it) >= s_end) break;  // wave-uniform
        const uint32_t row = walk_row<U>(wa, it);
        Key<KW, NC> sk[U];
        uint32_t st[U];   // 16-bit tag << 16 | bucket (buckets < 2^16)
        uint32_t act = 0;  // bit u: row u of this lane still probing
        if (!PF && it > 0) walk_load<U, KW, NC>(wa, it, nk);
#pragma unroll
        for (int u = 0; u < U; u++) {
            sk[u] = nk[u];
            res_set(res, it * U + u, kNoMatch);
            act |= (row + u * 64 < s_end ? 1u : 0u) << u;
        }
        if (PF && it + 1 < kProbeIters) walk_load<U, KW, NC>(wa, it + 1, nk);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t h = hash_key<HP, KW, NC>(sk[u], hc);
            st[u] = ((((uint32_t)(h >> tag_shift) & 0xFFFFu) | 1u) << 16) | ((uint32_t)(h >> shift) & bmask);
        }
        while (__ballot(act != 0)) {
            uint4 E[U];
#pragma unroll
            for (int u = 0; u < U; u++) E[u] = *reinterpret_cast<const uint4 *>(slots + ((st[u] & 0xFFFFu) << 2));
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool live = (act >> u) & 1u;
                const uint32_t tg = st[u] >> 16;
                const uint32_t e4[kBucketSlots] = {E[u].x, E[u].y, E[u].z, E[u].w};
                // the first entry whose tag matches (0: none; tags are never 0, so
                // neither an entry nor a match is 0); a bucket's slots fill in
                // order, so it has a free slot (the row's key cannot be further
                // on) iff its last slot is free
                uint32_t ej = 0;
#pragma unroll
                for (int j = kBucketSlots - 1; j >= 0; j--) ej = (e4[j] >> 16) == tg ? e4[j] : ej;
                const bool emp = e4[kBucketSlots - 1] == 0;
                if (!live) ej = 0;
                uint32_t fo = ej ? (ej & 0xFFFFu) - 1 : 0u;
                bool found = ej != 0;
#pragma unroll
                for (int c = 0; c < NC; c++) found &= rk[c * CAP + fo] == sk[u].c[c];
                if (__ballot(ej != 0 && !found)) {  // a 16-bit tag collision: the later candidates, in slot order
                    uint32_t cand = 0;
#pragma unroll
                    for (int j = 0; j < kBucketSlots; j++) cand |= ((e4[j] >> 16) == tg ? 1u : 0u) << j;
                    cand = ej != 0 && !found ? cand & (cand - 1) : 0u;
                    while (__ballot(cand != 0)) {
                        const bool has = cand != 0;
                        const uint32_t j = __builtin_ctz(cand | (1u << kBucketSlots));
                        const uint32_t ejj = j == 0 ? e4[0] : (j == 1 ? e4[1] : (j == 2 ? e4[2] : e4[3]));
                        cand &= cand - 1;
                        const uint32_t o = has ? (ejj & 0xFFFFu) - 1 : 0u;
                        bool eq = has;
#pragma unroll
                        for (int c = 0; c < NC; c++) eq &= rk[c * CAP + o] == sk[u].c[c];
                        if (eq) {
                            found = true;
                            fo = o;
                            cand = 0;
                        }
                    }
                }
                const uint32_t cur = res_get(res, it * U + u);
                res_set(res, it * U + u, found ? fo : cur);
                uint32_t n = found ? 1u : 0u;
                if (wa.dup) {  // wave-uniform: equal keys of the build chunk are chained behind the entry
                    uint32_t x = found ? nxt[fo] : 0u;
                    while (__ballot(x != 0)) {
                        n += x != 0;
                        x = x != 0 ? nxt[x - 1] : 0u;
                    }
                }
                total += n;
                act &= ~((live && (found || emp) ? 1u : 0u) << u);
                st[u] = (st[u] & 0xFFFF0000u) | ((st[u] + 1) & bmask);
            }
        }
    }
    return wave_sum<64>(total);
}

// Writes the matches probe_walk kept: each matched row's entry, then (when the
// build chunk has equal keys) the rest of its chain. The probe-side and
// build-side row ids of a group of G iterations are loaded before any is
// written, so their round trips overlap.
template <int G>
__device__ __forceinline__ void replay_walk(const WalkArgs &wa, const uint16_t *nxt,
                                            const uint32_t (&res)[kProbeIters * kProbeUnroll / 2], Emitter &em) {
    constexpr int U = kProbeUnroll;
    static_assert(kProbeIters % G == 0, "replay groups");
    const uint32_t ob1 = wa.ob1, s_end = wa.s_end;
#pragma unroll
    for (int g = 0; g < kProbeIters; g += G) {
        if (walk_wave_row<U>(wa, g) >= s_end) break;  // wave-uniform
        // ids of every row, matched or not, with clamped indexes: no per-row
        // branches, and the loads of a group are issued together (rows past
        // the chunk hold stale results: they read row s_lo / build row 0)
        uint32_t sid[G * U], bid[G * U];
#pragma unroll
        for (int it = 0; it < G; it++) {
            const uint32_t row = walk_row<U>(wa, g + it);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = row + u * 64;
                const bool in = i < s_end;
                const uint32_t o = res_get(res, (g + it) * U + u);
                sid[it * U + u] = wa.s.rowids ? wa.s.rowids[in ? i : wa.s_lo] : i;  // wave-uniform test
                bid[it * U + u] = build_id(wa, in && o != kNoMatch ? o : 0u);
            }
        }
#pragma unroll
        for (int it = 0; it < G; it++) {
            if (walk_wave_row<U>(wa, g + it) >= s_end) break;  // wave-uniform
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t o = res_get(res, (g + it) * U + u);
                const bool m = o != kNoMatch;
                em.put(m, bid[it * U + u] + ob1, sid[it * U + u] + ob1, wa);
                if (wa.dup) {  // wave-uniform
                    uint32_t x = m ? nxt[o] : 0u;
                    while (__ballot(x != 0)) {
                        const bool has = x != 0;
                        const uint32_t r = has ? x - 1 : 0u;
                        em.put(has, (has ? build_id(wa, r) : 0u) + ob1, sid[it * U + u] + ob1, wa);
                        x = has ? nxt[r] : 0u;
                    }
                }
            }
        }
    }
}

// Waves per SIMD the probe kernel is compiled for: 8 (64 VGPRs, four blocks
// per CU) for single 8-byte keys with FNV / XOR hashing; wider keys and the
// generic hash path get 4 (128 VGPRs, two blocks per CU) so they do not spill.
template <int HP, int KB> constexpr int probe_waves_per_simd() { return KB <= 8 && HP != HP_GENERIC ? 8 : 4; }

template <int HP, int KW, int NC>
__global__ __launch_bounds__(kProbeThreads, (probe_waves_per_simd<HP, KW * NC>())) void k_probe(PipeArgs a) {
    using KT = typename KeyT<KW>::T;
    constexpr int KB = KW * NC;
    constexpr int CAP = build_cap(KB);
    extern __shared__ __align__(16) unsigned char smem[];
    KT *rk = reinterpret_cast<KT *>(smem);                       // [NC][CAP]
    uint32_t *slots = reinterpret_cast<uint32_t *>(rk + NC * CAP);  // [2 * CAP]
    uint16_t *nxt = reinterpret_cast<uint16_t *>(slots + 2 * CAP);  // [CAP] chain of equal keys (index + 1, 0: end)
    __shared__ ProbeTask t;
    uint32_t res[kProbeIters * kProbeUnroll / 2];  // matched build row per probe row, two per word

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    Emitter em;
    em.pos = 0;
    const HashCfg hc = a.hash;
    const unsigned long long total = a.ctl->task_total;
    const uint32_t tag_shift = hc.bits - 16;

    // Lane 0 of wave 0 holds the ticket of the block's next task, fetched
    // while the current one runs.
    unsigned long long next_task = ~0ull;
    while (true) {
        if (w == 0) {
            unsigned long long task = 0;
            if (lane == 0) task = next_task != ~0ull ? next_task : atomicAdd(&a.ctl->task_ticket, 1ull);
            task = __shfl(task, 0, 64);
            if (task >= total) {
                if (lane == 0) t.done = 1;
            } else {
                uint32_t jn, q;
                if (task < a.taskmap_cap) {  // wave-uniform
                    const unsigned long long e = a.taskmap[task];
                    jn = (uint32_t)(e >> 32);
                    q = (uint32_t)e;
                } else {
                    jn = wave_upper_index(a.task0, a.n_joins, task);
                    const JoinDesc &jd0 = a.joins[jn];
                    q = wave_upper_index(a.ptask + jd0.ptask_off, jd0.nparts, task - a.task0[jn]);
                }
                const JoinDesc &jd = a.joins[jn];
                const unsigned long long *pt = a.ptask + jd.ptask_off;
                const unsigned long long local = task - a.task0[jn];
                if (lane == 0) {
                    const JobDesc &jr = a.jobs[jd.job_r];
                    const JobDesc &js = a.jobs[jd.job_s];
                    const uint32_t *rb = a.base + jr.base_off, *sb = a.base + js.base_off;
                    const uint32_t nr = rb[q + 1] - rb[q], ns = sb[q + 1] - sb[q];
                    const uint32_t s_chunks = (ns + jd.cap_s - 1) / jd.cap_s;
                    const uint32_t k = (uint32_t)(local - pt[q]);
                    const uint32_t rc = k / s_chunks, sc = k - rc * s_chunks;
                    if (jr.b1) {
                        for (int c = 0; c < 4; c++) t.r.cols[c] = jr.fin_cols[c];
                        t.r.rowids = jr.fin_ids;
                    } else {
                        t.r = jr.in;
                    }
                    if (js.b1) {
                        for (int c = 0; c < 4; c++) t.s.cols[c] = js.fin_cols[c];
                        t.s.rowids = js.fin_ids;
                    } else {
                        t.s = js.in;
                    }
                    t.r_lo = rb[q] + rc * jd.cap_r;
                    t.r_n = nr - rc * jd.cap_r < jd.cap_r ? nr - rc * jd.cap_r : jd.cap_r;
                    t.s_lo = sb[q] + sc * jd.cap_s;
                    t.s_n = ns - sc * jd.cap_s < jd.cap_s ? ns - sc * jd.cap_s : jd.cap_s;
                    uint32_t ns2 = 64;
                    // load <= 1/4 where the slot array (2 * CAP) allows, else <= 1/2:
                    // a probe round costs the wave's longest chain, not the mean
                    while (ns2 < 4 * t.r_n && ns2 < 2 * (uint32_t)CAP) ns2 <<= 1;
                    t.nslots = ns2;
                    t.shift = jd.slot_shift;
                    t.one_based = jd.one_based;
                    t.out_r = jd.out_r;
                    t.out_s = jd.out_s;
                    t.cap = jd.cap;
                    t.count = &a.out_count[jn];
                    t.dup = 0;
                    t.done = 0;
                }
            }
        }
        __syncthreads();
        if (t.done) break;
        if (w == 0 && lane == 0) next_task = atomicAdd(&a.ctl->task_ticket, 1ull);

        // build: load the R chunk, then insert (index + 1, 16-bit tag) into buckets;
        // a key already in the table is chained in front of its entry instead
        const uint32_t r_lo = t.r_lo, r_n = t.r_n, nslots = t.nslots, bmask = (nslots >> 2) - 1, shift = t.shift;
        for (uint32_t i = tid; i < nslots; i += kProbeThreads) slots[i] = 0;
        {
            // two R rows per lane in flight at once (a task has r_n >= 1)
            constexpr int RB = 2;
            const GView rv = uniform_view(t.r);
            for (uint32_t i0 = 0; i0 < r_n; i0 += RB * kProbeThreads) {
                Key<KW, NC> rkey[RB];
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    const uint32_t i = i0 + tid + r * kProbeThreads;
                    rkey[r] = load_key<KW, NC>(rv, r_lo + (i < r_n ? i : 0u));
                }
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    const uint32_t i = i0 + tid + r * kProbeThreads;
                    if (i < r_n) {
#pragma unroll
                        for (int c = 0; c < NC; c++) rk[c * CAP + i] = rkey[r].c[c];
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < r_n; i += kProbeThreads) {
            Key<KW, NC> k;
#pragma unroll
            for (int c = 0; c < NC; c++) k.c[c] = rk[c * CAP + i];
            const uint64_t h = hash_key<HP, KW, NC>(k, hc);
            const uint32_t tag = ((uint32_t)(h >> tag_shift) & 0xFFFFu) | 1u;  // never 0: an entry is never 0
            const uint32_t e = (tag << 16) | (i + 1);
            uint32_t b = (uint32_t)(h >> shift) & bmask;
            // Scan slots in order (the bucket, then the next buckets). An empty
            // slot takes the row as a new entry; the entry of an equal key takes
            // it as its new chain head (nxt[i] = old head). nxt[i] is written
            // before the CAS that publishes the row, and never changes after.
            bool placed = false;
            while (!placed) {
                for (uint32_t j = 0; j < (uint32_t)kBucketSlots && !placed; j++) {
                    uint32_t *sl = &slots[b * kBucketSlots + j];
                    nxt[i] = 0;
                    uint32_t old = atomicCAS(sl, 0u, e);
                    placed = old == 0;
                    while (!placed && (old >> 16) == tag) {
                        const uint32_t o = (old & 0xFFFFu) - 1;
                        bool eq = true;
#pragma unroll
                        for (int c = 0; c < NC; c++) eq &= rk[c * CAP + o] == k.c[c];
                        if (!eq) break;  // another key with the same tag: go on scanning
                        t.dup = 1;
                        nxt[i] = (uint16_t)(o + 1);
                        const uint32_t prev = atomicCAS(sl, old, e);
                        placed = prev == old;
                        old = prev;  // another row of this key became the head first: chain behind it
                    }
                }
                b = (b + 1) & bmask;
            }
        }
        __syncthreads();

        // find and count every row's matches (kept in registers), reserve the
        // task's output rows with one atomic, then write them
        const WalkArgs wa = walk_args(t);
        const uint32_t wc = probe_walk<HP, KW, NC>(wa, hc, tag_shift, rk, slots, nxt, res);
        if (lane == 0) t.wave_count[w] = wc;
        __syncthreads();
        if (tid == 0) {
            uint32_t sum = 0;
            for (int k = 0; k < kProbeWaves; k++) sum += t.wave_count[k];
            t.out_base = sum ? atomicAdd(t.count, (unsigned long long)sum) : 0;
        }
        __syncthreads();
        if (wc) {
            em.pos = t.out_base;
            for (uint32_t k = 0; k < w; k++) em.pos += t.wave_count[k];
            // row-id groups of two iterations where registers allow (single 8-byte key, FNV / XOR)
            replay_walk<(KB <= 8 && HP != HP_GENERIC) ? 2 : 1>(wa, nxt, res, em);
        }
        __syncthreads();
    }
}

__global__ void k_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
        offsets[i] = base[i];
        hist[i] = base[i + 1] - base[i];
    }
}

// ---------------------------------------------------------------------------
// Dispatch over (hash path, key width, key columns)
// ---------------------------------------------------------------------------

template <typename F> hipError_t dispatch(const LaunchShape &sh, F &&f) {
#define VJ_CASE(HP, KW, NC) \
    if (sh.hp == HP && sh.kw == KW && sh.nc == NC) return f.template operator()<HP, KW, NC>();
#define VJ_CASES_HP(HP) \
    VJ_CASE(HP, 4, 1) VJ_CASE(HP, 4, 2) VJ_CASE(HP, 4, 3) VJ_CASE(HP, 4, 4) \
    VJ_CASE(HP, 8, 1) VJ_CASE(HP, 8, 2) VJ_CASE(HP, 8, 3) VJ_CASE(HP, 8, 4)
    VJ_CASES_HP(HP_FNV)
    VJ_CASES_HP(HP_XOR_HW)
    VJ_CASES_HP(HP_GENERIC)
#undef VJ_CASES_HP
#undef VJ_CASE
    return hipErrorInvalidValue;
}

size_t pass_lds_bytes(int kw, int nc) {
    const int kb = kw * nc, T = pass_tile(kb);
    return (size_t)T * (kb > 4 ? kb : 4) + (size_t)(kPassWaves * 256 + 768) * 4 + (size_t)T;
}

size_t probe_lds_bytes(int kw, int nc) {
    const int kb = kw * nc, cap = build_cap(kb);
    return (size_t)cap * kb + (size_t)cap * 8 + (size_t)cap * 2;
}

}  // namespace

hipError_t launch_hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s) {
    if (a.total_hist_blocks == 0) return hipSuccess;
    const size_t lds = (size_t)a.max_hist_bins * 4;
    return dispatch(sh, [&]<int HP, int KW, int NC>() {
        hipLaunchKernelGGL((k_hist<HP, KW, NC>), dim3(a.total_hist_blocks), dim3(kHistThreads), lds, s, a);
        return hipGetLastError();
    });
}

hipError_t launch_scan(const PipeArgs &a, hipStream_t s) {
    hipError_t e;
    if (a.n_colscan) {
        hipLaunchKernelGGL(k_colscan, dim3(a.n_colscan), dim3(1024), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bases, dim3(a.n_jobs ? a.n_jobs : 1), dim3(1024), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.n_joins) {
        hipLaunchKernelGGL(k_tasks, dim3(a.n_joins), dim3(1024), 0, s, a);
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

hipError_t launch_reset_probe(const PipeArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_probe, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s) {
    const uint32_t total = pass == 1 ? a.total_p1_tiles : a.total_p2_chunks;
    if (total == 0) return hipSuccess;
    const uint32_t groups = (total + kXcdGroup - 1) / kXcdGroup;  // pass 2: every XCD gets the same number of groups
    const uint32_t grid = pass == 1 ? (total + kXcds - 1) / kXcds * kXcds : (groups + kXcds - 1) / kXcds * kXcds * kXcdGroup;
    const size_t lds = pass_lds_bytes(sh.kw, sh.nc);
    return dispatch(sh, [&]<int HP, int KW, int NC>() {
        hipLaunchKernelGGL((k_pass<HP, KW, NC>), dim3(grid), dim3(kPassThreads), lds, s, a, pass);
        return hipGetLastError();
    });
}

int probe_blocks_per_cu(const LaunchShape &sh) {
    const size_t lds = probe_lds_bytes(sh.kw, sh.nc);
    int blocks = 1;
    (void)dispatch(sh, [&]<int HP, int KW, int NC>() {
        int n = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_probe<HP, KW, NC>, kProbeThreads, lds);
        if (e == hipSuccess && n > 0) blocks = n;
        return e;
    });
    return blocks > 8 ? 8 : blocks;
}

hipError_t launch_probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s) {
    const size_t lds = probe_lds_bytes(sh.kw, sh.nc);
    return dispatch(sh, [&]<int HP, int KW, int NC>() {
        hipLaunchKernelGGL((k_probe<HP, KW, NC>), dim3(blocks), dim3(kProbeThreads), lds, s, a);
        return hipGetLastError();
    });
}

hipError_t launch_emit_hist(const uint32_t *base, uint32_t nb, uint32_t *hist, uint32_t *offsets, hipStream_t s) {
    unsigned blocks = (nb + 255) / 256;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_emit_hist, dim3(blocks), dim3(256), 0, s, base, nb, hist, offsets);
    return hipGetLastError();
}

}  // namespace vj
