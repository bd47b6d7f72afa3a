// The join pipeline's hashing stages, templated on the hash path, the key
// shape and the kernel shapes:
//
//   k_hist   per segment of a job: per-tile pass-1 digit prefixes, the
//            segment's pass-1 and (pass-1, pass-2) digit counts (LDS)
//   k_pass   one radix pass over a tile (pass 1) or a (digit, segment) chunk
//            (pass 2)
//   k_probe  persistent build + probe over (join, partition, R-chunk,
//            S-chunk) tasks from a device queue
//
// Included once per hash path (vj_join_fnv.hip, vj_join_xorhw.hip,
// vj_join_generic.hip), each of which instantiates StageLaunch<HP>, so the
// three paths compile in parallel. The reference's counterparts are
// histogram_kernel / prefix_sum_kernel / element_swap_kernel
// (src/join/join_provider.cu:315-380) and the build / probe / extract
// kernels (:382-608).
#pragma once
#include "vj_kernel_common.hpp"

namespace vj {

namespace {

// ---------------------------------------------------------------------------

// One block per segment. The block sweeps the segment one pass-1 tile at a
// time (all 1024 threads on one tile, coalesced), with one LDS atomic per row
// into the tile's pass-1 digit counters and, for 2-pass jobs, one into the
// segment's (pass-1, pass-2) histogram. Tile counters live in a ring of
// kHistRing tiles; every kHistRing tiles the ring is folded into the running
// in-segment prefix, which is stored per tile (tilepre).
constexpr int kHistRing = kHistThreads / 64;

template <int HP, int KW, int NC, int TV>
__global__ __launch_bounds__(kHistThreads) void k_hist(PipeArgs a) {
    constexpr int KB = KW * NC;
    constexpr int T = pass_tile(KB, TV);
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
    const bool lo32 = s2 + b2 <= 32;  // every digit below bit 32 (joins): the low half of the hash suffices
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
                    const uint64_t h = lo32 ? (uint64_t)hash_key_lo<HP, KW, NC>(key[u], hc) : hash_key<HP, KW, NC>(key[u], hc);
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
    // region layout (region == 0: exact layout): digit d's counter is
    // counts[rc0 + d * rstride], its region starts at row rrow0 + d * rdstride
    uint32_t region, rstride, rrow0, rdstride, trash;
    uint64_t rc0;
};

// Waves per SIMD the pass kernel is compiled for: 8 (64 VGPRs, four blocks per
// CU) for single 8-byte keys with FNV / XOR hashing; wider keys and the
// generic hash path get 6 (80 VGPRs, three blocks) to limit spills; the large
// tile (more rows per lane) 6, or 5 beyond 8-byte keys.
template <int HP, int KB, int TV> constexpr int pass_waves_per_simd() {
    return TV == 2 ? (KB <= 8 ? 6 : 5) : (KB <= 8 && HP != HP_GENERIC ? 8 : 6);
}

// Pass 1: one tile of a job. Pass 2: one (pass-1 digit, segment) chunk, in
// tile-sized steps; a running count per digit carries the destination.
// TV: tile shape (pass_items).
template <int HP, int KW, int NC, int TV>
__global__ __launch_bounds__(kPassThreads, (pass_waves_per_simd<HP, KW * NC, TV>())) void k_pass(PipeArgs a, int pass) {
    using KT = typename KeyT<KW>::T;
    constexpr int KB = KW * NC;
    constexpr int ITEMS = pass_items(KB, TV);
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
        // In the region layout the runs into one final region arrive in
        // atomic order, so pass 2 keeps every chunk of a pass-1 digit d1 on
        // XCD d1 % kXcds instead: a region's partial lines then merge in one
        // L2. Each job's chunk count (2^b1 digits x c chunks) is a multiple of
        // kXcds; XCD x takes job j's digits x, x + kXcds, ... in turn.
        const uint32_t *pfx = pass == 1 ? a.p1_tile0 : a.p2_chunk0;
        const uint32_t total = pass == 1 ? a.total_p1_tiles : a.total_p2_chunks;
        const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
        uint32_t b, valid, j;
        if (pass == 2 && a.regions) {
            valid = k * kXcds < total;
            j = valid ? wave_upper_index(pfx, a.n_jobs, k * kXcds) : 0;
            const JobDesc &jb = a.jobs[j];
            const uint32_t c = jb.nseg * jb.region_tiles, kk = k - pfx[j] / kXcds;
            b = valid ? pfx[j] + (x + kXcds * (kk / c)) * c + kk % c : total;
        } else {
            b = pass == 1 ? x * ((total + kXcds - 1) / kXcds) + k
                          : ((k / kXcdGroup) * kXcds + x) * kXcdGroup + k % kXcdGroup;
            valid = b < total;
            j = valid ? wave_upper_index(pfx, a.n_jobs, b) : 0;
        }
        if (lane == 0) {
            ti.valid = valid;
            ti.region = 0;
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
                    if (job.region1) {  // region (d1, g) of mid, g = this tile's segment
                        const uint32_t g = local / job.seg_tiles;
                        ti.region = job.region1;
                        ti.rc0 = job.m1_off + g;
                        ti.rstride = job.nseg;
                        ti.rrow0 = g * job.region1;
                        ti.rdstride = job.nseg * job.region1;
                        ti.trash = (job.nseg << job.b1) * job.region1;
                    }
                } else if (job.region1) {
                    // chunk = tile kt of pass-1 region (d1, g); digit d2's region is (d1 << b2) | d2 of fin
                    const uint32_t per = job.nseg * job.region_tiles;
                    const uint32_t d1 = local / per, r = local - d1 * per;
                    const uint32_t g = r / job.region_tiles, kt = r - g * job.region_tiles;
                    const uint32_t c1 = a.counts[job.m1_off + d1 * job.nseg + g];
                    const uint32_t filled = c1 < job.region1 ? c1 : job.region1;
                    const uint32_t start = (d1 * job.nseg + g) * job.region1;
                    for (int c = 0; c < 4; c++) ti.src.cols[c] = job.mid_cols[c];
                    ti.src.rowids = job.mid_ids;
                    ti.src.n = job.n;
                    for (int c = 0; c < 4; c++) ti.dst_cols[c] = job.fin_cols[c];
                    ti.dst_ids = job.fin_ids;
                    ti.lo = start + kt * T;
                    ti.hi = start + ((kt + 1) * T < filled ? (kt + 1) * T : filled);
                    ti.valid = kt * T < filled;
                    ti.bins = 1u << job.b2;
                    ti.shift = job.shift0 + job.b1;
                    ti.region = job.region2;
                    ti.rc0 = job.m2_off + (d1 << job.b2);
                    ti.rstride = 1;
                    ti.rrow0 = (d1 << job.b2) * job.region2;
                    ti.rdstride = job.region2;
                    ti.trash = (1u << (job.b1 + job.b2)) * job.region2;
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
    if (tid < ti.bins && !ti.region) {  // first destination row of every digit (exact layout)
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
    const bool lo32 = shift + (uint32_t)__builtin_ctz(bins) <= 32;  // the digit lies in the hash's low half
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
            if (valid) {
                const uint64_t h = lo32 ? (uint64_t)hash_key_lo<HP, KW, NC>(keys[it], hc) : hash_key<HP, KW, NC>(keys[it], hc);
                d = (uint32_t)(h >> shift) & mask;
            }
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
        const uint32_t my_loc = block_excl_scan<uint32_t, kPassWaves>(cnt, sh_scan, &total);
        if (tid < bins) {
            loc[tid] = my_loc;
            uint32_t dst;
            if (ti.region) {  // reserve the run in the digit's region (a run that does not fit goes to the trash tile)
                const uint32_t got = cnt ? atomicAdd(a.counts + ti.rc0 + (uint64_t)tid * ti.rstride, cnt) : 0u;
                dst = (uint64_t)got + cnt <= ti.region ? ti.rrow0 + tid * ti.rdstride + got : ti.trash;
            } else {
                dst = gdst[tid];
            }
            gdo[tid] = dst - my_loc;  // output row of staged slot k = gdo[digit] + k
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
    uint32_t wave_count[kMaxProbeWaves];
    // the join whose fixed fields the record holds (~0: none yet), its
    // sides' partition start / end offsets in base, chunk sizes, first task
    // and per-partition task prefix
    uint32_t jn, rb, re, sb, se, cap_r, cap_s;
    unsigned long long task0;
    const unsigned long long *pt;
};

constexpr int kProbeUnroll = 2;
constexpr int kBucketSlots = 4;  // LDS table: buckets of 4 slots (one 16-byte load)

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
// (IDS: from the table's copy in LDS, `rid`.)
template <bool IDS> __device__ __forceinline__ uint32_t build_id(const WalkArgs &wa, const uint32_t *rid, uint32_t o) {
    if constexpr (IDS) return rid[o];
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
// kProbeChunk = ITERS x TH x kProbeUnroll rows (TH threads per block, see
// ProbeShape); in iteration `it` lane `l` of wave `w` owns rows
// it*TH*U + w*64*U + u*64 + l.
// The table holds one entry per distinct key (equal keys of the build chunk
// are chained behind it, see the build step), so a row matches at most one
// entry. probe_walk finds it, keeps its index in registers and counts the
// row's matches (the chain length); replay_walk then writes them without
// probing again. Probing runs in rounds: each round reads one whole bucket
// for every unfinished row of the lane (U independent LDS chains per lane);
// a round is branch-free (selects, no per-row branches). The S keys of the
// next iteration are loaded before the current one is probed, so one HBM round
// trip overlaps each iteration's probing (where registers allow, see PF).

// A key's place in the LDS table: its 16-bit tag (never 0) << 16 | its bucket,
// from the low 32 hash bits above the partition digits (the digits are the
// same for every row of a partition; digits and buckets end below bit 27).
// The tag is the top of a multiplicative hash of those bits, which depends on
// all of them (the top 16 bits of a 64-bit hash would not do: custom_xor_hw's
// 64-bit hash of 4-byte keys has a zero upper half, so every key would get
// the same tag).
__device__ __forceinline__ uint32_t table_place(uint32_t lo, uint32_t shift, uint32_t bmask) {
    const uint32_t y = lo >> shift;
    return ((((y * 0x9E3779B1u) >> 16) | 1u) << 16) | (y & bmask);
}

template <int N> __device__ __forceinline__ uint32_t res_get(const uint32_t (&r)[N], int k) {
    return (r[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
}
template <int N> __device__ __forceinline__ void res_set(uint32_t (&r)[N], int k, uint32_t v) {
    const int sh = (k & 1) * 16;
    r[k >> 1] = (r[k >> 1] & ~(0xFFFFu << sh)) | (v << sh);
}
// Block shape of the build+probe kernel for table shape PV (probe_threads).
template <int PV> struct ProbeShape {
    static constexpr int TH = probe_threads(PV);
    static constexpr int WAVES = TH / 64;
    static constexpr int ITERS = kProbeChunk / (TH * kProbeUnroll);
    static constexpr int RES = ITERS * kProbeUnroll / 2;  // words of packed 16-bit results per lane
    static_assert(kProbeUnroll % 2 == 0, "an iteration's results fill whole words");
    static_assert(ITERS * TH * kProbeUnroll == kProbeChunk && WAVES <= kMaxProbeWaves, "probe chunk shape");
};

// First S row of this lane in iteration `it` of a walk with U rows per lane.
template <int U, int TH> __device__ __forceinline__ uint32_t walk_row(const WalkArgs &wa, int it) {
    return wa.s_lo + it * TH * U + (threadIdx.x >> 6) * 64 * U + lane_id();
}
// First S row of the wave in iteration `it` (wave-uniform).
template <int U, int TH> __device__ __forceinline__ uint32_t walk_wave_row(const WalkArgs &wa, int it) {
    return wa.s_lo + it * TH * U + uniform((uint32_t)(threadIdx.x >> 6)) * 64 * U;
}

// Loads this lane's S keys of iteration `it` (rows past the chunk are left alone).
template <int U, int TH, int KW, int NC>
__device__ __forceinline__ void walk_load(const WalkArgs &wa, int it, Key<KW, NC> (&k)[U]) {
    const uint32_t row = walk_row<U, TH>(wa, it);
#pragma unroll
    for (int u = 0; u < U; u++)
        if (row + u * 64 < wa.s_end) k[u] = load_key<KW, NC>(wa.s, row + u * 64);
}

template <int N> __device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

template <int HP, int KW, int NC, int PV>
__device__ __forceinline__ uint32_t probe_walk(const WalkArgs &wa, const HashCfg &hc,
                                               const typename KeyT<KW>::T *rk, const uint32_t *slots,
                                               const uint16_t *nxt, uint32_t (&res)[ProbeShape<PV>::RES]) {
    constexpr int CAP = probe_cap(KW * NC, PV);
    constexpr int TH = ProbeShape<PV>::TH, ITERS = ProbeShape<PV>::ITERS;
    constexpr int U = kProbeUnroll;
    // Next-iteration loads double the key registers; they are used where the
    // kernel stays within its register budget (single 8-byte key, or the wide
    // block's 128 VGPRs; FNV / XOR).
    constexpr bool PF = (KW * NC <= 8 || PV == 1) && HP != HP_GENERIC;
    const uint32_t bmask = wa.bmask, shift = wa.shift, s_end = wa.s_end;
    uint32_t total = 0;  // this lane's matches
    Key<KW, NC> nk[U];
    walk_load<U, TH, KW, NC>(wa, 0, nk);
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
        if (walk_wave_row<U, TH>(wa, it) >= s_end) break;  // wave-uniform
        const uint32_t row = walk_row<U, TH>(wa, it);
        Key<KW, NC> sk[U];
        // 16-bit tag << 16 | first bucket (buckets < 2^16); round r reads bucket
        // first + r. 0 once the row is done (or past the chunk): tag 0 matches
        // only empty slots, so such a row finds nothing and stays done.
        uint32_t st[U];
        if (!PF && it > 0) walk_load<U, TH, KW, NC>(wa, it, nk);
#pragma unroll
        for (int u = 0; u < U; u++) sk[u] = nk[u];
        if (PF && it + 1 < ITERS) walk_load<U, TH, KW, NC>(wa, it + 1, nk);
#pragma unroll
        for (int u = 0; u < U; u++)
            st[u] = row + u * 64 < s_end ? table_place(hash_key_lo<HP, KW, NC>(sk[u], hc), shift, bmask) : 0u;
        uint32_t rnd = 0;
        while (true) {
            uint32_t any = 0;
#pragma unroll
            for (int u = 0; u < U; u++) any |= st[u];
            if (!__ballot(any != 0)) break;
            uint4 E[U];
#pragma unroll
            for (int u = 0; u < U; u++) E[u] = *reinterpret_cast<const uint4 *>(slots + (((st[u] + rnd) & bmask) << 2));
#pragma unroll
            for (int u = 0; u < U; u++) {
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
                res[(it * U + u) >> 1] |= (found ? fo + 1 : 0u) << (((it * U + u) & 1) * 16);
                uint32_t n = found ? 1u : 0u;
                if (wa.dup) {  // wave-uniform: equal keys of the build chunk are chained behind the entry
                    uint32_t x = found ? nxt[fo] : 0u;
                    while (__ballot(x != 0)) {
                        n += x != 0;
                        x = x != 0 ? nxt[x - 1] : 0u;
                    }
                }
                total += n;
                st[u] = found || emp ? 0u : st[u];
            }
            rnd++;
        }
    }
    return wave_sum<64>(total);
}

// Writes the matches probe_walk kept: each matched row's entry, then (when the
// build chunk has equal keys) the rest of its chain. The probe-side and
// build-side row ids of a group of G iterations are loaded before any is
// written, so their round trips overlap.
template <int G, int PV, bool IDS>
__device__ __forceinline__ void replay_walk(const WalkArgs &wa, const uint16_t *nxt, const uint32_t *rid,
                                            const uint32_t (&res)[ProbeShape<PV>::RES], Emitter &em) {
    constexpr int U = kProbeUnroll;
    constexpr int TH = ProbeShape<PV>::TH, ITERS = ProbeShape<PV>::ITERS;
    static_assert(ITERS % G == 0, "replay groups");
    const uint32_t ob1 = wa.ob1, s_end = wa.s_end;
#pragma unroll
    for (int g = 0; g < ITERS; g += G) {
        if (walk_wave_row<U, TH>(wa, g) >= s_end) break;  // wave-uniform
        // ids of every row, matched or not, with clamped indexes: no per-row
        // branches, and the loads of a group are issued together (rows past
        // the chunk, and all rows of iterations the walk skipped, hold no
        // match: they read row s_lo / build row 0)
        uint32_t sid[G * U], bid[G * U];
#pragma unroll
        for (int it = 0; it < G; it++) {
            const uint32_t row = walk_row<U, TH>(wa, g + it);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t i = row + u * 64;
                const bool in = i < s_end;
                const uint32_t f = res_get(res, (g + it) * U + u);
                sid[it * U + u] = wa.s.rowids ? wa.s.rowids[in ? i : wa.s_lo] : i;  // wave-uniform test
                bid[it * U + u] = build_id<IDS>(wa, rid, in && f != 0 ? f - 1 : 0u);
            }
        }
#pragma unroll
        for (int it = 0; it < G; it++) {
            if (walk_wave_row<U, TH>(wa, g + it) >= s_end) break;  // wave-uniform
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t f = res_get(res, (g + it) * U + u);
                const bool m = f != 0;
                const uint32_t o = m ? f - 1 : 0u;
                em.put(m, bid[it * U + u] + ob1, sid[it * U + u] + ob1, wa);
                if (wa.dup) {  // wave-uniform
                    uint32_t x = m ? nxt[o] : 0u;
                    while (__ballot(x != 0)) {
                        const bool has = x != 0;
                        const uint32_t r = has ? x - 1 : 0u;
                        em.put(has, (has ? build_id<IDS>(wa, rid, r) : 0u) + ob1, sid[it * U + u] + ob1, wa);
                        x = has ? nxt[r] : 0u;
                    }
                }
            }
        }
    }
}

// Waves per SIMD the probe kernel is compiled for. Standard table: 8 (64
// VGPRs, four blocks per CU) for single 8-byte keys with FNV / XOR hashing;
// wider keys and the generic hash path get 4 (128 VGPRs, two blocks per CU) so
// they do not spill. Wide table: one 16-wave block per CU, 4 waves per SIMD
// (128 VGPRs).
template <int HP, int KB, int PV> constexpr int probe_waves_per_simd() {
    return PV ? 4 : (KB <= 8 && HP != HP_GENERIC ? 8 : 4);
}

// PV: table shape (probe_cap, probe_threads).
template <int HP, int KW, int NC, int PV>
__global__ __launch_bounds__(probe_threads(PV), (probe_waves_per_simd<HP, KW * NC, PV>())) void k_probe(PipeArgs a) {
    using KT = typename KeyT<KW>::T;
    constexpr int KB = KW * NC;
    constexpr int CAP = probe_cap(KB, PV);
    constexpr int TH = ProbeShape<PV>::TH, WAVES = ProbeShape<PV>::WAVES;
    extern __shared__ __align__(16) unsigned char smem[];
    KT *rk = reinterpret_cast<KT *>(smem);                       // [NC][CAP]
    uint32_t *slots = reinterpret_cast<uint32_t *>(rk + NC * CAP);  // [2 * CAP]
    uint16_t *nxt = reinterpret_cast<uint16_t *>(slots + 2 * CAP);  // [CAP] chain of equal keys (index + 1, 0: end)
    constexpr bool IDS = table_keeps_ids(KB);
    uint32_t *rid = reinterpret_cast<uint32_t *>(nxt + CAP);  // [CAP] build rows' ids (IDS only)
    __shared__ ProbeTask t;
    uint32_t res[ProbeShape<PV>::RES];  // per probe row its matched build row + 1 (0: none), two per word

    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    Emitter em;
    em.pos = 0;
    const HashCfg hc = a.hash;
    const unsigned long long total = a.ctl->task_total;

    // Lane 0 of wave 0 holds the ticket of the block's next task, fetched
    // while the current one runs.
    unsigned long long next_task = ~0ull;
    if (tid == 0) t.jn = ~0u;  // read and written by this thread only (lane 0 of wave 0)
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
                if (lane == 0) {
                    if (jn != t.jn) {  // a join's fixed fields, kept while its tasks follow each other
                        const JoinDesc &jd = a.joins[jn];
                        const JobDesc &jr = a.jobs[jd.job_r];
                        const JobDesc &js = a.jobs[jd.job_s];
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
                        t.rb = jr.base_off;
                        t.re = jr.end_off;
                        t.sb = js.base_off;
                        t.se = js.end_off;
                        t.cap_r = jd.cap_r;
                        t.cap_s = jd.cap_s;
                        t.task0 = a.task0[jn];
                        t.pt = a.ptask + jd.ptask_off;
                        t.shift = jd.slot_shift;
                        t.one_based = jd.one_based;
                        t.out_r = jd.out_r;
                        t.out_s = jd.out_s;
                        t.cap = jd.cap;
                        t.count = &a.out_count[jn];
                        t.jn = jn;
                    }
                    // the partition's sides and the task's chunks: one round of loads
                    const uint32_t cap_r = t.cap_r, cap_s = t.cap_s;
                    const uint32_t r0 = a.base[t.rb + q], s0 = a.base[t.sb + q];
                    const uint32_t nr = a.base[t.re + q] - r0, ns = a.base[t.se + q] - s0;
                    const uint32_t s_chunks = (ns + cap_s - 1) / cap_s;
                    const uint32_t k = (uint32_t)(task - t.task0 - t.pt[q]);
                    const uint32_t rc = k / s_chunks, sc = k - rc * s_chunks;
                    t.r_lo = r0 + rc * cap_r;
                    t.r_n = nr - rc * cap_r < cap_r ? nr - rc * cap_r : cap_r;
                    t.s_lo = s0 + sc * cap_s;
                    t.s_n = ns - sc * cap_s < cap_s ? ns - sc * cap_s : cap_s;
                    uint32_t ns2 = 64;
                    // load <= 1/4 where the slot array (2 * CAP) allows, else <= 1/2:
                    // a probe round costs the wave's longest chain, not the mean
                    while (ns2 < 4 * t.r_n && ns2 < 2 * (uint32_t)CAP) ns2 <<= 1;
                    t.nslots = ns2;
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
        for (uint32_t i = tid; i < nslots; i += TH) slots[i] = 0;
        {
            // two R rows per lane in flight at once (a task has r_n >= 1)
            constexpr int RB = 2;
            const GView rv = uniform_view(t.r);
            for (uint32_t i0 = 0; i0 < r_n; i0 += RB * TH) {
                Key<KW, NC> rkey[RB];
                uint32_t rrow[RB];
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    const uint32_t i = i0 + tid + r * TH;
                    const uint32_t row = r_lo + (i < r_n ? i : 0u);
                    rkey[r] = load_key<KW, NC>(rv, row);
                    if constexpr (IDS) rrow[r] = rv.rowids ? rv.rowids[row] : row;  // wave-uniform test
                }
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    const uint32_t i = i0 + tid + r * TH;
                    if (i < r_n) {
#pragma unroll
                        for (int c = 0; c < NC; c++) rk[c * CAP + i] = rkey[r].c[c];
                        if constexpr (IDS) rid[i] = rrow[r];
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < r_n; i += TH) {
            Key<KW, NC> k;
#pragma unroll
            for (int c = 0; c < NC; c++) k.c[c] = rk[c * CAP + i];
            const uint32_t place = table_place(hash_key_lo<HP, KW, NC>(k, hc), shift, bmask);  // tag never 0: an entry is never 0
            const uint32_t tag = place >> 16;
            const uint32_t e = (tag << 16) | (i + 1);
            uint32_t b = place & 0xFFFFu;
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
#pragma unroll
        for (int k = 0; k < ProbeShape<PV>::RES; k++) res[k] = 0;  // results of skipped iterations read as no match
        const uint32_t wc = probe_walk<HP, KW, NC, PV>(wa, hc, rk, slots, nxt, res);
        if (lane == 0) t.wave_count[w] = wc;
        __syncthreads();
        if (tid == 0) {
            uint32_t sum = 0;
            for (int k = 0; k < WAVES; k++) sum += t.wave_count[k];
            t.out_base = sum ? atomicAdd(t.count, (unsigned long long)sum) : 0;
        }
        __syncthreads();
        if (wc) {
            em.pos = t.out_base;
            for (uint32_t k = 0; k < w; k++) em.pos += t.wave_count[k];
            // row-id groups of four iterations for single keys of up to 8 bytes (FNV / XOR) and where the
            // build rows' ids come from LDS (keys wider than 8 bytes; eight in the wide block)
            replay_walk<(KB <= 8 && HP != HP_GENERIC) ? 4 : (IDS ? (PV ? 8 : 4) : 1), PV, IDS>(wa, nxt, rid, res, em);
        }
        __syncthreads();
    }
}

// Dispatch over the key shape (key width, key columns).
template <typename F> hipError_t dispatch_keys(const LaunchShape &sh, F &&f) {
#define VJ_CASE(KW, NC) \
    if (sh.kw == KW && sh.nc == NC) return f.template operator()<KW, NC>();
    VJ_CASE(4, 1) VJ_CASE(4, 2) VJ_CASE(4, 3) VJ_CASE(4, 4)
    VJ_CASE(8, 1) VJ_CASE(8, 2) VJ_CASE(8, 3) VJ_CASE(8, 4)
#undef VJ_CASE
    return hipErrorInvalidValue;
}

size_t pass_lds_bytes(int kw, int nc, int tile) {
    const int kb = kw * nc, T = pass_tile(kb, tile);
    return (size_t)T * (kb > 4 ? kb : 4) + (size_t)(kPassWaves * 256 + 768) * 4 + (size_t)T;
}

size_t probe_lds_bytes(int kw, int nc, int table) {
    const int kb = kw * nc;
    return (size_t)probe_cap(kb, table) * table_row_bytes(kb);
}

}  // namespace

template <int HP> hipError_t StageLaunch<HP>::hist(const PipeArgs &a, const LaunchShape &sh, hipStream_t s) {
    const size_t lds = (size_t)a.max_hist_bins * 4;
    return dispatch_keys(sh, [&]<int KW, int NC>() {
        if (sh.tile == 1)
            hipLaunchKernelGGL((k_hist<HP, KW, NC, 1>), dim3(a.total_hist_blocks), dim3(kHistThreads), lds, s, a);
        else if (sh.tile == 2)
            hipLaunchKernelGGL((k_hist<HP, KW, NC, 2>), dim3(a.total_hist_blocks), dim3(kHistThreads), lds, s, a);
        else
            hipLaunchKernelGGL((k_hist<HP, KW, NC, 0>), dim3(a.total_hist_blocks), dim3(kHistThreads), lds, s, a);
        return hipGetLastError();
    });
}

template <int HP> hipError_t StageLaunch<HP>::pass(const PipeArgs &a, const LaunchShape &sh, int pass, hipStream_t s) {
    const uint32_t total = pass == 1 ? a.total_p1_tiles : a.total_p2_chunks;
    const uint32_t groups = (total + kXcdGroup - 1) / kXcdGroup;  // pass 2: every XCD gets the same number of groups
    const uint32_t grid = pass == 1 ? (total + kXcds - 1) / kXcds * kXcds : (groups + kXcds - 1) / kXcds * kXcds * kXcdGroup;
    const size_t lds = pass_lds_bytes(sh.kw, sh.nc, sh.tile);
    return dispatch_keys(sh, [&]<int KW, int NC>() {
        if (sh.tile == 1)
            hipLaunchKernelGGL((k_pass<HP, KW, NC, 1>), dim3(grid), dim3(kPassThreads), lds, s, a, pass);
        else if (sh.tile == 2)
            hipLaunchKernelGGL((k_pass<HP, KW, NC, 2>), dim3(grid), dim3(kPassThreads), lds, s, a, pass);
        else
            hipLaunchKernelGGL((k_pass<HP, KW, NC, 0>), dim3(grid), dim3(kPassThreads), lds, s, a, pass);
        return hipGetLastError();
    });
}

template <int HP> int StageLaunch<HP>::probe_blocks_per_cu(const LaunchShape &sh) {
    const size_t lds = probe_lds_bytes(sh.kw, sh.nc, sh.table);
    int blocks = 1;
    (void)dispatch_keys(sh, [&]<int KW, int NC>() {
        int n = 0;
        const hipError_t e =
            sh.table ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_probe<HP, KW, NC, 1>, probe_threads(1), lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_probe<HP, KW, NC, 0>, probe_threads(0), lds);
        if (e == hipSuccess && n > 0) blocks = n;
        return e;
    });
    return blocks > 8 ? 8 : blocks;
}

template <int HP> hipError_t StageLaunch<HP>::probe(const PipeArgs &a, const LaunchShape &sh, int blocks, hipStream_t s) {
    const size_t lds = probe_lds_bytes(sh.kw, sh.nc, sh.table);
    return dispatch_keys(sh, [&]<int KW, int NC>() {
        if (sh.table)
            hipLaunchKernelGGL((k_probe<HP, KW, NC, 1>), dim3(blocks), dim3(probe_threads(1)), lds, s, a);
        else
            hipLaunchKernelGGL((k_probe<HP, KW, NC, 0>), dim3(blocks), dim3(probe_threads(0)), lds, s, a);
        return hipGetLastError();
    });
}

}  // namespace vj
