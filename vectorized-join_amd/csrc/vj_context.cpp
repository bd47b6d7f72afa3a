// vj_ctx: device arenas, the join planner, pipeline launch, and the join /
// partition / batch entry points of the C ABI.
//
// Reference counterparts: JoinProvider (src/join/join_provider.cuh:222-500)
// with join_gpu's stage order (:286-421), DeviceConfig's memory admission
// (:124-220), partition_gpu (src/join/join_provider.cu:610-661) and the
// bucket-pair probing loop (join_provider.cuh:792-911), here a single
// device-side work queue instead of host threads.
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <cmath>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "vj_host.hpp"
#include "vj_internal.hpp"
#include "vj_pipeline.hpp"

using namespace vj;

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kNoAllocStageBytes = 64 << 10;  // descriptors + counts of one join (a few KiB)

inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

enum Ev { EV_START, EV_HIST, EV_SCAN, EV_P1, EV_P2, EV_PROBE, EV_COUNT };

}  // namespace

struct vj_ctx {
    int device = 0;
    int cus = 256;
    uint64_t budget = 0;
    bool profiling = false;
    std::string last_error;
    vj_join_summary summary{};
    DevBuf tuples, counts, base, ptask, meta;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t ev[EV_COUNT] = {};
    // Captured pipeline (vj_join_batch with use_graph): the graph is reused
    // while the launch signature (every kernel argument, grid and copy) is
    // unchanged; descriptors are re-read from the pinned buffer at each replay.
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_sig;
    std::vector<uint64_t> last_sig;  // signature of the last eager pipeline (automatic capture)
    bool graph_failed = false;       // capture failed once: stay eager
    uint64_t graph_replays = 0, graph_captures = 0;
    // Autotuned plans (vj_tune): size class -> tuned_word(passes, bits1, bits2, tile, table, layout)
    std::unordered_map<uint64_t, uint64_t> tuned;
    // size classes whose tuned region layout overflowed: the exact layout from then on
    std::unordered_set<uint64_t> region_overflowed;
    // vj_join_into_noalloc: this device's fixed scratch (looked up at creation)
    // and a pinned staging block for its descriptors and counts (allocated at
    // creation, so the entry point itself allocates nothing)
    void *na_scratch = nullptr;
    void *na_stage = nullptr;
    std::string na_error;
};

namespace {

// --- memory ----------------------------------------------------------------

uint64_t held_bytes(const vj_ctx *c) {
    return c->tuples.bytes + c->counts.bytes + c->base.bytes + c->ptask.bytes + c->meta.bytes;
}

void free_buf(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

void free_all(vj_ctx *c) {
    free_buf(c->tuples);
    free_buf(c->counts);
    free_buf(c->base);
    free_buf(c->ptask);
    free_buf(c->meta);
}

// Grows `b` to at least `bytes`; with `slack` an arena that must grow takes
// 1/8 more, so slowly growing shapes do not reallocate every time (a
// reallocation frees the old buffer, which waits for the device).
Err ensure(DevBuf &b, size_t bytes, bool slack) {
    if (b.bytes >= bytes) return Err::ok();
    free_buf(b);
    size_t want = align_up(slack ? bytes + bytes / 8 : bytes);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return Err::make(VJ_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    b.bytes = want;
    return Err::ok();
}

Err ensure_pinned(vj_ctx *c, size_t bytes) {
    if (c->pinned_bytes >= bytes) return Err::ok();
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_bytes = 0;
    size_t want = align_up(bytes * 2);
    VJ_HIP_TRY(hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
    c->pinned_bytes = want;
    return Err::ok();
}

// --- planning ----------------------------------------------------------------


struct Plan {
    std::vector<JobDesc> jobs;
    std::vector<JoinDesc> joins;
    std::vector<ColScanDesc> colscan;
    std::vector<uint32_t> hist_block0, p1_tile0, p2_chunk0;
    std::vector<size_t> fin_off, mid_off;  // byte offsets in the tuples arena (SIZE_MAX: none)
    std::vector<uint64_t> fin_rows, mid_rows;  // rows of each buffer (the region layout leaves gaps)
    bool regions = false;                  // region layout (every job has two passes)
    size_t tuples_bytes = 0;
    uint64_t counts_words = 0, base_words = 0, ptask_words = 0;
    uint64_t taskmap_cap = 0;  // entries of the task map (after the task prefixes in the ptask arena)
    uint32_t max_nparts = 0;
    uint32_t max_hist_bins = 1, total_hist_blocks = 0, total_p1_tiles = 0, total_p2_chunks = 0;
    bool any_part = false, any_p2 = false;
    // histogram blocks wanted over the whole pipeline (two per CU) and the rows
    // they share; a 2-pass job gets about its share of the blocks
    uint64_t hist_rows_total = 0;
    uint32_t hist_blocks_target = 0;
    LaunchShape shape{};
    HashCfg hash{};
    int kb = 0;
    std::vector<std::pair<uint32_t, vj_built *>> keeps;  // (job, table to fill)
    uint64_t keep_bytes = 0;
    // meta layout
    size_t off_jobs = 0, off_joins = 0, off_cs = 0, off_hb0 = 0, off_pt0 = 0, off_pc0 = 0, upload_bytes = 0;
    size_t off_jtasks = 0, off_task0 = 0, off_count = 0, off_ctl = 0, meta_bytes = 0;
    // Bytes downloaded after a launch: the match counts and the control words
    // behind them (the queue's task total).
    size_t down_bytes() const { return off_ctl + sizeof(Ctl) - off_count; }
};

int ceil_log2(uint64_t v) {
    int b = 0;
    while ((1ull << b) < v) b++;
    return b;
}

// Radix bits that bring nr build rows to half a build chunk of `cap` rows
// per partition.
int need_bits(uint64_t nr, int cap) {
    const uint64_t target = (uint64_t)cap / 2;
    return nr > target ? ceil_log2((nr + target - 1) / target) : 0;
}

// Radix passes and bits for one join (the reference fixes one radix width
// per level, join_provider.cuh:296-300, and recurses until a partition fits,
// :569-571; here the partition target is half an LDS build chunk of the
// pipeline's table shape).
Err plan_bits(uint64_t nr, int kb, int table, const vj_join_opts *o, int *P, int *b1, int *b2) {
    const int need = need_bits(nr, probe_cap(kb, table));
    int passes = o->passes;
    if (passes < -1 || passes > 2) return Err::make(VJ_ERR_INVALID_ARGUMENT, "passes must be -1, 0, 1 or 2");
    if (passes == -1) passes = need == 0 ? 0 : (need <= kMaxPassBits ? 1 : 2);
    if (o->bits1 < 0 || o->bits2 < 0 || o->bits1 > kMaxPassBits || o->bits2 > kMaxPassBits)
        return Err::make(VJ_ERR_UNSUPPORTED, "radix bits per pass must be 1..8");
    *P = passes;
    if (passes == 0) {
        *b1 = *b2 = 0;
    } else if (passes == 1) {
        *b1 = o->bits1 ? o->bits1 : std::max(1, std::min(need, kMaxPassBits));
        *b2 = 0;
    } else {
        int total = std::max(2, std::min(need, kMaxHistBits));
        int x1 = o->bits1 ? o->bits1 : std::min(kMaxPassBits, (total + 1) / 2);
        int x2 = o->bits2 ? o->bits2 : std::max(1, std::min(kMaxPassBits, total - x1));
        *b1 = x1;
        *b2 = x2;
    }
    if (*b1 + *b2 > kMaxHistBits) return Err::make(VJ_ERR_UNSUPPORTED, "bits1 + bits2 must be <= 14");
    return Err::ok();
}

Err check_pair(const vj_table *r, const vj_table *s) {
    VJ_TRY(check_table(r));
    VJ_TRY(check_table(s));
    // join_gpu asserts equal column counts (join_provider.cuh:291)
    if (r->n_cols != s->n_cols || r->key_width != s->key_width || (r->key_signed != 0) != (s->key_signed != 0))
        return Err::make(VJ_ERR_INVALID_ARGUMENT, "build and probe tables must have the same key columns");
    return Err::ok();
}


// Bytes of a resident table of n rows.
size_t built_bytes(uint64_t n, int nc, int kw, uint32_t b1, uint32_t b2) {
    return align_up((size_t)kw * n) * nc + align_up((size_t)4 * n) + align_up((size_t)4 * ((1u << (b1 + b2)) + 1));
}

Err alloc_built(vj_built *b) {
    b->bytes = built_bytes(b->n, b->n_cols, b->key_width, b->b1, b->b2);
    if (hipMalloc(&b->mem, b->bytes) != hipSuccess) {
        (void)hipGetLastError();
        b->mem = nullptr;
        return Err::make(VJ_ERR_OUT_OF_MEMORY, "hipMalloc of a resident table failed");
    }
    char *p = static_cast<char *>(b->mem);
    const size_t col = align_up((size_t)b->key_width * b->n);
    for (int c = 0; c < 4; c++) b->cols[c] = c < b->n_cols ? p + c * col : nullptr;
    b->ids = reinterpret_cast<uint32_t *>(p + b->n_cols * col);
    b->base = reinterpret_cast<uint32_t *>(p + b->n_cols * col + align_up((size_t)4 * b->n));
    return Err::ok();
}

void free_built(vj_built *b) {
    if (b->mem) (void)hipFree(b->mem);
    b->mem = nullptr;
}

// A resident build side as an already partitioned job: no hist / pass work.
uint32_t add_resident_job(Plan &pl, const vj_built *b) {
    JobDesc jd;
    memset(&jd, 0, sizeof(jd));
    jd.in.n = b->n;
    jd.n = (uint32_t)b->n;
    jd.b1 = b->b1;
    jd.b2 = b->b2;
    for (int c = 0; c < 4; c++) jd.fin_cols[c] = b->cols[c];
    jd.fin_ids = b->ids;
    jd.res_base = b->base;
    jd.base_off = (uint32_t)pl.base_words;
    jd.end_off = jd.base_off + 1;
    pl.base_words += (1u << (b->b1 + b->b2)) + 1;
    pl.any_part = true;
    pl.hist_block0.push_back(pl.total_hist_blocks);
    pl.p1_tile0.push_back(pl.total_p1_tiles);
    pl.p2_chunk0.push_back(pl.total_p2_chunks);
    pl.fin_off.push_back(SIZE_MAX);
    pl.mid_off.push_back(SIZE_MAX);
    pl.fin_rows.push_back(0);
    pl.mid_rows.push_back(0);
    pl.jobs.push_back(jd);
    return (uint32_t)pl.jobs.size() - 1;
}

// Region layout: pass-1 tiles per segment (each of a segment's 2^b1 region
// counters takes one device atomic per tile; a counter serialises its
// atomics, so segments keep each one to about a thousand).
constexpr uint32_t kRegionSegTiles = 1024;

// Rows of a region whose expected fill is `mean` rows: the mean plus 32
// standard deviations of a uniform (Poisson) spread plus 256, in whole
// 64-row groups. Keys repeated a few times widen the spread by about the
// square root of the repeat count (FK probe sides: 8 repeats, ~11 standard
// deviations left); heavier skew overflows and falls back to the exact layout.
uint64_t region_rows(uint64_t mean) {
    const uint64_t r = mean + (uint64_t)(32.0 * std::sqrt((double)mean)) + 256;
    return (r + 63) & ~63ull;
}

// Region sizes of a two-pass job of n rows (false: its buffers would pass
// VJ_MAX_ROWS rows, the bound every 32-bit row index of the kernels keeps).
struct RegionShape {
    uint32_t nseg, seg_tiles, region1, region2, region_tiles;
    uint64_t mid_rows, fin_rows;
};
bool region_shape(uint64_t n, int b1, int b2, uint32_t T, RegionShape *r) {
    const uint64_t tiles = (n + T - 1) / T;
    const uint64_t nseg = std::max<uint64_t>(1, std::min<uint64_t>((tiles + kRegionSegTiles - 1) / kRegionSegTiles,
                                                                   kMaxSegments));
    const uint64_t seg_tiles = std::max<uint64_t>(1, (tiles + nseg - 1) / nseg);
    const uint64_t seg_rows = std::min<uint64_t>(n, seg_tiles * T);
    const uint64_t r1 = region_rows((seg_rows + (1ull << b1) - 1) >> b1);
    const uint64_t r2 = region_rows((n + (1ull << (b1 + b2)) - 1) >> (b1 + b2));
    r->nseg = (uint32_t)((tiles + seg_tiles - 1) / seg_tiles);
    if (r->nseg == 0) r->nseg = 1;
    r->seg_tiles = (uint32_t)seg_tiles;
    r->mid_rows = ((uint64_t)r->nseg << b1) * r1 + T;
    r->fin_rows = (1ull << (b1 + b2)) * r2 + T;
    r->region1 = (uint32_t)r1;
    r->region2 = (uint32_t)r2;
    r->region_tiles = (uint32_t)((r1 + T - 1) / T);
    return r->mid_rows <= VJ_MAX_ROWS && r->fin_rows <= VJ_MAX_ROWS;
}

void add_job(Plan &pl, const vj_table *t, int P, int b1, int b2, uint32_t shift0, bool fin_external = false) {
    JobDesc jd;
    memset(&jd, 0, sizeof(jd));
    jd.in = to_view(t);
    jd.n = (uint32_t)t->n_rows;
    jd.b1 = P >= 1 ? b1 : 0;
    jd.b2 = P >= 2 ? b2 : 0;
    jd.shift0 = shift0;
    const uint32_t T = pass_tile(pl.kb, pl.shape.tile);
    const uint32_t n1 = 1u << jd.b1, nb = 1u << (jd.b1 + jd.b2);
    jd.base_off = (uint32_t)pl.base_words;
    jd.end_off = jd.base_off + 1;
    pl.base_words += (jd.b1 ? nb : 1) + 1;
    if (pl.regions && jd.b2) {  // region layout: the partition ends after the starts
        jd.end_off = (uint32_t)pl.base_words;
        pl.base_words += nb;
    }
    const uint32_t job_index = (uint32_t)pl.jobs.size();
    uint32_t chunks = 0;
    uint64_t mid_rows = jd.n, fin_rows = jd.n;
    if (pl.regions && jd.b2) {
        // region layout: no histogram rows, no column scans; counters m1
        // [2^b1][nseg] and m2 [2^(b1+b2)] (zeroed by k_region_reset)
        RegionShape rs;
        (void)region_shape(jd.n, (int)jd.b1, (int)jd.b2, T, &rs);  // checked by the planner
        pl.any_part = true;
        pl.any_p2 = true;
        jd.p1_tiles = (uint32_t)(((uint64_t)jd.n + T - 1) / T);
        jd.nseg = rs.nseg;
        jd.seg_tiles = rs.seg_tiles;
        jd.hsplit = 1;
        jd.region1 = rs.region1;
        jd.region2 = rs.region2;
        jd.region_tiles = jd.n ? rs.region_tiles : 0;  // an empty job has no pass-2 chunks
        jd.m1_off = pl.counts_words;
        pl.counts_words += (uint64_t)jd.nseg * n1;
        jd.m2_off = pl.counts_words;
        pl.counts_words += nb;
        chunks = jd.nseg * n1 * jd.region_tiles;
        mid_rows = rs.mid_rows;
        fin_rows = rs.fin_rows;
    } else if (jd.b1 && jd.n) {
        pl.any_part = true;
        jd.p1_tiles = (uint32_t)(((uint64_t)jd.n + T - 1) / T);
        // segments: pass-2 chunks (a segment's rows of one pass-1 partition)
        // averaging 15/16 of a tile, so that chunk sizes (spread around the
        // mean) rarely need a second pass-2 step; for one pass, enough
        // segments to spread the count
        const uint64_t chunk = (uint64_t)T * 15 / 16;
        uint64_t want = jd.b2 ? ((uint64_t)jd.n + (uint64_t)n1 * chunk - 1) / ((uint64_t)n1 * chunk)
                              : ((uint64_t)jd.p1_tiles + 15) / 16;
        want = std::max<uint64_t>(1, std::min<uint64_t>({want, (uint64_t)kMaxSegments, jd.p1_tiles}));
        jd.seg_tiles = (uint32_t)((jd.p1_tiles + want - 1) / want);
        jd.nseg = (jd.p1_tiles + jd.seg_tiles - 1) / jd.seg_tiles;
        // 2-pass jobs with fewer segments than their share of the histogram
        // blocks split each segment into hsplit histogram rows
        uint64_t hs = 1;
        if (jd.b2 && pl.hist_rows_total && pl.hist_blocks_target) {
            const uint64_t share = (uint64_t)pl.hist_blocks_target * jd.n / pl.hist_rows_total;
            hs = std::max<uint64_t>(1, std::min<uint64_t>({share / jd.nseg, (uint64_t)kMaxSegments / jd.nseg,
                                                           (uint64_t)jd.seg_tiles}));
        }
        jd.hsplit = (uint32_t)hs;
        jd.hrow_tiles = (jd.seg_tiles + jd.hsplit - 1) / jd.hsplit;
        jd.nrow = jd.nseg * jd.hsplit;
        jd.tilepre_off = pl.counts_words;
        pl.counts_words += (uint64_t)jd.p1_tiles * n1;
        jd.m1_off = pl.counts_words;
        pl.counts_words += (uint64_t)(jd.nrow + 1) * n1;
        for (uint32_t c0 = 0; c0 < n1; c0 += kColScanCols) pl.colscan.push_back({job_index, 1u, c0, 0u});
        if (jd.b2) {
            pl.any_p2 = true;
            jd.m2_off = pl.counts_words;
            pl.counts_words += (uint64_t)(jd.nrow + 1) * nb;
            for (uint32_t c0 = 0; c0 < nb; c0 += kColScanCols) pl.colscan.push_back({job_index, 2u, c0, 0u});
            chunks = n1 * jd.nseg;
        }
        pl.max_hist_bins = std::max(pl.max_hist_bins, (jd.b2 ? nb : 0) + n1 * (1 + kHistThreads / 64));
    }
    pl.hist_block0.push_back(pl.total_hist_blocks);
    pl.total_hist_blocks += jd.nrow;
    pl.p1_tile0.push_back(pl.total_p1_tiles);
    pl.total_p1_tiles += jd.p1_tiles;
    pl.p2_chunk0.push_back(pl.total_p2_chunks);
    pl.total_p2_chunks += chunks;
    // tuple buffers: fin (and mid for two passes), SoA key columns + row ids
    size_t fin = SIZE_MAX, mid = SIZE_MAX;
    auto bytes = [&](uint64_t rows) {
        return align_up((size_t)t->key_width * rows) * t->n_cols + align_up((size_t)4 * rows);
    };
    if (jd.b1) {
        if (!fin_external) {
            fin = pl.tuples_bytes;
            pl.tuples_bytes += bytes(fin_rows);
        }
        if (jd.b2) {
            mid = pl.tuples_bytes;
            pl.tuples_bytes += bytes(mid_rows);
        }
    }
    pl.fin_off.push_back(fin);
    pl.mid_off.push_back(mid);
    pl.fin_rows.push_back(fin_rows);
    pl.mid_rows.push_back(mid_rows);
    pl.jobs.push_back(jd);
}

void finish_layout(Plan &pl) {
    const size_t nj = pl.jobs.size(), nn = pl.joins.size();
    pl.hist_block0.push_back(pl.total_hist_blocks);
    pl.p1_tile0.push_back(pl.total_p1_tiles);
    pl.p2_chunk0.push_back(pl.total_p2_chunks);
    size_t o = 0;
    pl.off_jobs = o;
    o = align_up(o + nj * sizeof(JobDesc));
    pl.off_joins = o;
    o = align_up(o + nn * sizeof(JoinDesc));
    pl.off_cs = o;
    o = align_up(o + pl.colscan.size() * sizeof(ColScanDesc));
    pl.off_hb0 = o;
    o = align_up(o + (nj + 1) * 4);
    pl.off_pt0 = o;
    o = align_up(o + (nj + 1) * 4);
    pl.off_pc0 = o;
    o = align_up(o + (nj + 1) * 4);
    pl.upload_bytes = o;
    pl.off_jtasks = o;
    o = align_up(o + nn * 8);
    pl.off_task0 = o;
    o = align_up(o + (nn + 1) * 8);
    pl.off_count = o;
    o = align_up(o + nn * 8);
    pl.off_ctl = o;
    o = align_up(o + sizeof(Ctl));
    pl.meta_bytes = o;
}

// Size class of an autotuned plan: build rows (rounded up to a power of two),
// key bytes and hash path.
uint64_t tune_key(uint64_t nr, int kb, int hp) {
    return (uint64_t)ceil_log2(std::max<uint64_t>(nr, 1)) | (uint64_t)kb << 8 | (uint64_t)hp << 16;
}

// Tuned plan word: passes | bits1 << 8 | bits2 << 16 | tile << 24 | table << 28 | layout << 32
// (tile, table: 0-based shapes; layout: VJ_LAYOUT_*).
uint64_t tuned_word(int P, int b1, int b2, int tile, int table, int layout) {
    return (uint64_t)P | (uint64_t)b1 << 8 | (uint64_t)b2 << 16 | (uint64_t)tile << 24 | (uint64_t)table << 28 |
           (uint64_t)layout << 32;
}

Err check_shapes(const vj_join_opts *o) {
    if (o->tile < 0 || o->tile > VJ_TILE_LARGE || o->table < 0 || o->table > VJ_TABLE_WIDE)
        return Err::make(VJ_ERR_INVALID_ARGUMENT,
                         "tile must be 0 (planner), 1 (standard), 2 (small) or 3 (large); table 0, 1 or 2");
    if (o->layout < 0 || o->layout > VJ_LAYOUT_REGIONS)
        return Err::make(VJ_ERR_INVALID_ARGUMENT, "layout must be 0 (planner), 1 (exact) or 2 (regions)");
    return Err::ok();
}

// Size class of a pipeline: its largest build side.
uint64_t largest_build(const std::vector<JoinSpec> &specs) {
    uint64_t nr = 0;
    for (const JoinSpec &js : specs) nr = std::max<uint64_t>(nr, js.res ? js.res->n : js.r->n_rows);
    return nr;
}

// Kernel shapes of a pipeline (one for all of its joins): the caller's, else
// the autotuned ones of the size class of the batch's largest build, else
// the wide table when a standard one would need more than kMaxHistBits bits
// of partitioning to bring the largest build down to half a chunk per
// partition; a tile left open (-1) becomes the large tile in the region
// layout and the standard one otherwise (build_join_plan).
Err choose_shapes(Plan &pl, const std::vector<JoinSpec> &specs, const vj_join_opts *o, const vj_ctx *c) {
    VJ_TRY(check_shapes(o));
    const uint64_t nr = largest_build(specs);
    int tile = o->tile - 1, table = o->table - 1;  // -1: choose
    if (c && o->passes == -1 && !o->bits1 && !o->bits2 && !c->tuned.empty()) {
        auto it = c->tuned.find(tune_key(nr, pl.kb, pl.shape.hp));
        if (it != c->tuned.end()) {
            if (tile < 0) tile = (int)((it->second >> 24) & 0xF);
            if (table < 0) table = (int)((it->second >> 28) & 0xF);
        }
    }
    if (table < 0) table = need_bits(nr, build_cap(pl.kb)) > kMaxHistBits ? 1 : 0;
    pl.shape.tile = tile;
    pl.shape.table = table;
    return Err::ok();
}

bool same_hash(const vj_hash_opts &a, const vj_hash_opts &b) {
    return a.kind == b.kind && a.hash_bits == b.hash_bits && a.chunk_bits == b.chunk_bits && a.word_mode == b.word_mode;
}

// The partition layout of a pipeline: the caller's, else the autotuned one of
// its size class (unless that class has overflowed a region before), else
// exact; regions only where every join has two passes and no resident or
// kept build side, and every region's rows fit 32-bit indexes.
bool choose_regions(const Plan &pl, const std::vector<JoinSpec> &specs, const std::vector<std::array<int, 3>> &bits,
                    const vj_join_opts *o, const vj_ctx *c, bool allow) {
    int layout = o->layout;
    if (!layout && c && o->passes == -1 && !o->bits1 && !o->bits2 && !c->tuned.empty()) {
        const uint64_t key = tune_key(largest_build(specs), pl.kb, pl.shape.hp);
        auto it = c->tuned.find(key);
        if (it != c->tuned.end() && !c->region_overflowed.count(key)) layout = (int)((it->second >> 32) & 0xF);
    }
    if (layout != VJ_LAYOUT_REGIONS || !allow) return false;
    const uint32_t T = pass_tile(pl.kb, pl.shape.tile < 0 ? 2 : pl.shape.tile);
    for (size_t i = 0; i < specs.size(); i++) {
        const JoinSpec &js = specs[i];
        // pass 2 deals a job's digits to the XCDs: 2^b1 must be a multiple of kXcds
        if (js.res || js.keep || bits[i][0] != 2 || (1u << bits[i][1]) % kXcds) return false;
        RegionShape rs;
        if (!region_shape(js.r->n_rows, bits[i][1], bits[i][2], T, &rs) ||
            !region_shape(js.s->n_rows, bits[i][1], bits[i][2], T, &rs))
            return false;
    }
    return true;
}

Err build_join_plan(Plan &pl, const std::vector<JoinSpec> &specs, const vj_join_opts *o, const vj_ctx *c = nullptr,
                    bool allow_regions = true) {
    const vj_table *t0 = specs[0].s;
    VJ_TRY(check_table(t0));
    VJ_TRY(make_hash_cfg(&o->hash, t0, &pl.hash));
    pl.kb = t0->key_width * t0->n_cols;
    pl.shape.hp = hash_path(pl.hash);
    pl.shape.kw = t0->key_width;
    pl.shape.nc = t0->n_cols;
    for (const JoinSpec &js : specs)
        if (!js.res) VJ_TRY(check_table(js.r));
    VJ_TRY(choose_shapes(pl, specs, o, c));
    for (const JoinSpec &js : specs) pl.hist_rows_total += (js.res ? 0 : js.r->n_rows) + js.s->n_rows;
    pl.hist_blocks_target = 2u * (uint32_t)(c ? c->cus : 256);  // sizing without a context: an MI355X
    // radix plan of every join: (passes, bits1, bits2)
    std::vector<std::array<int, 3>> bits(specs.size());
    for (size_t i = 0; i < specs.size(); i++) {
        const JoinSpec &js = specs[i];
        int &P = bits[i][0], &b1 = bits[i][1], &b2 = bits[i][2];
        if (js.res) {
            VJ_TRY(check_table(js.s));
            const vj_built *b = js.res;
            if (b->n_cols != js.s->n_cols || b->key_width != js.s->key_width ||
                (b->key_signed != 0) != (js.s->key_signed != 0))
                return Err::make(VJ_ERR_INVALID_ARGUMENT, "probe table and resident table have different key columns");
            if (!same_hash(b->hash, o->hash))
                return Err::make(VJ_ERR_INVALID_ARGUMENT, "the resident table was partitioned with another hash");
            P = b->b2 ? 2 : 1;
            b1 = (int)b->b1;
            b2 = (int)b->b2;
        } else {
            VJ_TRY(check_pair(js.r, js.s));
            VJ_TRY(plan_bits(js.r->n_rows, pl.kb, pl.shape.table, o, &P, &b1, &b2));
            if (c && o->passes == -1 && !o->bits1 && !o->bits2 && !c->tuned.empty()) {  // autotuned plan
                auto it = c->tuned.find(tune_key(js.r->n_rows, pl.kb, pl.shape.hp));
                if (it != c->tuned.end()) {
                    P = (int)(it->second & 0xFF);
                    b1 = (int)((it->second >> 8) & 0xFF);
                    b2 = (int)((it->second >> 16) & 0xFF);
                }
            }
            if (js.keep && P == 0) {  // a kept table is always partitioned
                P = 1;
                b1 = 1;
                b2 = 0;
            }
        }
    }
    pl.regions = choose_regions(pl, specs, bits, o, c, allow_regions);
    if (pl.shape.tile < 0) pl.shape.tile = pl.regions ? 2 : 0;
    for (size_t i = 0; i < specs.size(); i++) {
        const JoinSpec &js = specs[i];
        const int P = bits[i][0], b1 = bits[i][1], b2 = bits[i][2];
        JoinDesc jd;
        memset(&jd, 0, sizeof(jd));
        if (js.res) {
            jd.job_r = add_resident_job(pl, js.res);
        } else {
            jd.job_r = (uint32_t)pl.jobs.size();
            add_job(pl, js.r, P, b1, b2, 0, js.keep != nullptr);
            if (js.keep) {
                vj_built *k = js.keep;
                k->n = js.r->n_rows;
                k->n_cols = js.r->n_cols;
                k->key_width = js.r->key_width;
                k->key_signed = js.r->key_signed != 0;
                k->hash = o->hash;
                k->b1 = (uint32_t)b1;
                k->b2 = (uint32_t)(P >= 2 ? b2 : 0);
                pl.keeps.push_back({jd.job_r, k});
                pl.keep_bytes += built_bytes(k->n, k->n_cols, k->key_width, k->b1, k->b2);
            }
        }
        if (js.s->n_cols != t0->n_cols || js.s->key_width != t0->key_width ||
            (js.s->key_signed != 0) != (t0->key_signed != 0))
            return Err::make(VJ_ERR_INVALID_ARGUMENT, "all joins of a batch must share the key columns' shape");
        jd.job_s = (uint32_t)pl.jobs.size();
        add_job(pl, js.s, P, b1, b2, 0);
        jd.out_r = js.out_r;
        jd.out_s = js.out_s;
        jd.cap = js.cap;
        jd.cap_r = probe_cap(pl.kb, pl.shape.table);
        jd.cap_s = kProbeChunk;
        jd.nparts = 1u << (P >= 1 ? b1 : 0) << (P >= 2 ? b2 : 0);
        jd.ptask_off = (uint32_t)pl.ptask_words;
        pl.ptask_words += jd.nparts + 1;
        jd.part_off = (uint32_t)pl.ptask_words;  // the task-prefix blocks' partials
        pl.ptask_words += 4ull * ((jd.nparts + kTaskBlock - 1) / kTaskBlock);
        pl.taskmap_cap += 2ull * jd.nparts;
        pl.max_nparts = std::max(pl.max_nparts, jd.nparts);
        jd.slot_shift = (P >= 1 ? b1 : 0) + (P >= 2 ? b2 : 0);
        jd.one_based = o->one_based ? 1 : 0;
        pl.joins.push_back(jd);
    }
    finish_layout(pl);
    return Err::ok();
}

// Arena sizes of a plan, as allocate() requests them (ensure() rounds each up
// to kAlign; under a budget it adds nothing else).
struct ArenaSizes {
    uint64_t tuples, counts, base, ptask, meta;
    uint64_t sum() const { return tuples + counts + base + ptask + meta; }
};
ArenaSizes arena_sizes(const Plan &pl) {
    return {align_up(std::max<size_t>(pl.tuples_bytes, kAlign)), align_up(std::max<size_t>(pl.counts_words * 4, kAlign)),
            align_up(std::max<size_t>(pl.base_words * 4, kAlign)),
            align_up(std::max<size_t>((pl.ptask_words + pl.taskmap_cap) * 8, kAlign)), align_up(pl.meta_bytes)};
}

uint64_t plan_scratch_bytes(const Plan &pl) { return arena_sizes(pl).sum(); }

// Admission (DeviceConfig::is_out_of_memory, join_provider.cuh:148-150): the
// scratch of this pipeline plus `extra` output bytes must fit the budget.
// Under a budget the arenas are allocated at exactly these sizes, so the
// context then holds at most budget - extra bytes.
Err admit(vj_ctx *c, const Plan &pl, uint64_t extra) {
    if (!c->budget) return Err::ok();
    const ArenaSizes a = arena_sizes(pl);
    const uint64_t need = a.sum() + extra;
    if (need > c->budget)
        return Err::make(VJ_ERR_BUDGET, "Not enough memory for swap tables and hashes: need " + std::to_string(need) +
                                            " bytes, budget " + std::to_string(c->budget));
    // drop larger-than-needed arenas when keeping them would break the budget
    const uint64_t keep = std::max<uint64_t>(c->tuples.bytes, a.tuples) + std::max<uint64_t>(c->counts.bytes, a.counts) +
                          std::max<uint64_t>(c->base.bytes, a.base) + std::max<uint64_t>(c->ptask.bytes, a.ptask) +
                          std::max<uint64_t>(c->meta.bytes, a.meta);
    if (keep + extra > c->budget) free_all(c);
    return Err::ok();
}

// Points every partitioned job's mid / fin tuple columns into the tuple
// arena at `tb` (offsets from the plan).
void place_tuples(Plan &pl, char *tb) {
    for (size_t j = 0; j < pl.jobs.size(); j++) {
        JobDesc &jd = pl.jobs[j];
        if (pl.fin_off[j] != SIZE_MAX) {
            const size_t col = align_up((size_t)pl.shape.kw * pl.fin_rows[j]);
            char *p = tb + pl.fin_off[j];
            for (int cc = 0; cc < 4; cc++) jd.fin_cols[cc] = cc < pl.shape.nc ? p + cc * col : nullptr;
            jd.fin_ids = reinterpret_cast<uint32_t *>(p + pl.shape.nc * col);
        }
        if (pl.mid_off[j] != SIZE_MAX) {
            const size_t col = align_up((size_t)pl.shape.kw * pl.mid_rows[j]);
            char *p = tb + pl.mid_off[j];
            for (int cc = 0; cc < 4; cc++) jd.mid_cols[cc] = cc < pl.shape.nc ? p + cc * col : nullptr;
            jd.mid_ids = reinterpret_cast<uint32_t *>(p + pl.shape.nc * col);
        }
    }
}

// Bytes the arenas hold once those that must grow have grown with 1/8 slack.
uint64_t grown_with_slack(const vj_ctx *c, const ArenaSizes &a) {
    auto g = [](const DevBuf &b, uint64_t need) { return b.bytes >= need ? (uint64_t)b.bytes : align_up(need + need / 8); };
    return g(c->tuples, a.tuples) + g(c->counts, a.counts) + g(c->base, a.base) + g(c->ptask, a.ptask) +
           g(c->meta, a.meta);
}

// Without a budget the arenas that must grow take slack. Under a budget
// (after admit()) they grow to exactly the plan, or, with `budget_slack`, with
// slack where the budget leaves room for it beside `extra` bytes (a serving
// lane, whose batches change size all the time; a direct context keeps the
// room for a regrown output).
Err allocate(vj_ctx *c, Plan &pl, bool budget_slack = false, uint64_t extra = 0) {
    const bool slack =
        c->budget == 0 || (budget_slack && grown_with_slack(c, arena_sizes(pl)) + extra <= c->budget);
    VJ_TRY(ensure(c->tuples, std::max<size_t>(pl.tuples_bytes, kAlign), slack));
    VJ_TRY(ensure(c->counts, std::max<size_t>(pl.counts_words * 4, kAlign), slack));
    VJ_TRY(ensure(c->base, std::max<size_t>(pl.base_words * 4, kAlign), slack));
    VJ_TRY(ensure(c->ptask, std::max<size_t>((pl.ptask_words + pl.taskmap_cap) * 8, kAlign), slack));
    VJ_TRY(ensure(c->meta, pl.meta_bytes, slack));
    VJ_TRY(ensure_pinned(c, pl.meta_bytes));
    place_tuples(pl, static_cast<char *>(c->tuples.p));
    // kept build sides: their own allocation receives the last pass and the bases
    for (auto &kp : pl.keeps) {
        vj_built *b = kp.second;
        if (!b->mem) {
            b->device = c->device;
            VJ_TRY(alloc_built(b));
        }
        JobDesc &jd = pl.jobs[kp.first];
        for (int cc = 0; cc < 4; cc++) jd.fin_cols[cc] = b->cols[cc];
        jd.fin_ids = b->ids;
        jd.keep_base = b->base;
    }
    return Err::ok();
}

// Device and staging memory of one pipeline launch: the context's arenas, or
// the fixed scratch of the no-allocation entry point.
struct Arenas {
    char *counts, *base, *ptask, *meta;  // device
    char *pinned;                        // host staging of the descriptors and the match counts
};

Arenas ctx_arenas(vj_ctx *c) {
    return {static_cast<char *>(c->counts.p), static_cast<char *>(c->base.p), static_cast<char *>(c->ptask.p),
            static_cast<char *>(c->meta.p), static_cast<char *>(c->pinned)};
}

PipeArgs make_args(const Arenas &ar, const Plan &pl) {
    char *m = ar.meta;
    PipeArgs a;
    memset(&a, 0, sizeof(a));
    a.jobs = reinterpret_cast<const JobDesc *>(m + pl.off_jobs);
    a.joins = reinterpret_cast<const JoinDesc *>(m + pl.off_joins);
    a.colscan = reinterpret_cast<const ColScanDesc *>(m + pl.off_cs);
    a.hist_block0 = reinterpret_cast<const uint32_t *>(m + pl.off_hb0);
    a.p1_tile0 = reinterpret_cast<const uint32_t *>(m + pl.off_pt0);
    a.p2_chunk0 = reinterpret_cast<const uint32_t *>(m + pl.off_pc0);
    a.join_tasks = reinterpret_cast<unsigned long long *>(m + pl.off_jtasks);
    a.task0 = reinterpret_cast<unsigned long long *>(m + pl.off_task0);
    a.out_count = reinterpret_cast<unsigned long long *>(m + pl.off_count);
    a.ctl = reinterpret_cast<Ctl *>(m + pl.off_ctl);
    a.counts = reinterpret_cast<uint32_t *>(ar.counts);
    a.base = reinterpret_cast<uint32_t *>(ar.base);
    a.ptask = reinterpret_cast<unsigned long long *>(ar.ptask);
    a.taskmap = a.ptask + pl.ptask_words;
    a.taskmap_cap = pl.taskmap_cap;
    a.max_nparts = pl.max_nparts;
    a.n_jobs = (uint32_t)pl.jobs.size();
    a.n_joins = (uint32_t)pl.joins.size();
    a.n_colscan = (uint32_t)pl.colscan.size();
    a.total_hist_blocks = pl.total_hist_blocks;
    a.total_p1_tiles = pl.total_p1_tiles;
    a.total_p2_chunks = pl.total_p2_chunks;
    a.max_hist_bins = pl.max_hist_bins;
    a.pass_tile = pass_tile(pl.kb, pl.shape.tile);
    a.regions = pl.regions ? 1u : 0u;
    a.hash = pl.hash;
    return a;
}

// Writes the descriptors into the pinned staging buffer.
void stage_descriptors(char *h, const Plan &pl) {
    memcpy(h + pl.off_jobs, pl.jobs.data(), pl.jobs.size() * sizeof(JobDesc));
    if (!pl.joins.empty()) memcpy(h + pl.off_joins, pl.joins.data(), pl.joins.size() * sizeof(JoinDesc));
    if (!pl.colscan.empty()) memcpy(h + pl.off_cs, pl.colscan.data(), pl.colscan.size() * sizeof(ColScanDesc));
    memcpy(h + pl.off_hb0, pl.hist_block0.data(), pl.hist_block0.size() * 4);
    memcpy(h + pl.off_pt0, pl.p1_tile0.data(), pl.p1_tile0.size() * 4);
    memcpy(h + pl.off_pc0, pl.p2_chunk0.data(), pl.p2_chunk0.size() * 4);
}

Err upload(const Arenas &ar, const Plan &pl, hipStream_t s) {
    stage_descriptors(ar.pinned, pl);
    VJ_HIP_TRY(hipMemcpyAsync(ar.meta, ar.pinned, pl.upload_bytes, hipMemcpyHostToDevice, s));
    return Err::ok();
}

#define VJ_LAUNCH(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return Err::make(VJ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

// Launches every stage on `s` (no host synchronisation). `events`: record
// the start / end events (and the per-stage ones when profiling); off while
// a graph is being captured.
Err launch_stages(vj_ctx *c, const Plan &pl, const PipeArgs &a, hipStream_t s, bool with_probe, bool events = true) {
    const bool prof = c->profiling && events;
    if (events) VJ_HIP_TRY(hipEventRecord(c->ev[EV_START], s));
    if (pl.regions) {
        // region layout: counters zeroed, both passes reserve runs, then the task prefix
        VJ_LAUNCH(launch_region_reset(a, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_HIST], s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_SCAN], s));
        VJ_LAUNCH(launch_pass(a, pl.shape, 1, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_P1], s));
        VJ_LAUNCH(launch_pass(a, pl.shape, 2, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_P2], s));
        VJ_LAUNCH(launch_tasks(a, s));
    } else {
        if (pl.any_part) VJ_LAUNCH(launch_hist(a, pl.shape, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_HIST], s));
        VJ_LAUNCH(launch_scan(a, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_SCAN], s));
        if (pl.any_part) VJ_LAUNCH(launch_pass(a, pl.shape, 1, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_P1], s));
        if (pl.any_p2) VJ_LAUNCH(launch_pass(a, pl.shape, 2, s));
        if (prof) VJ_HIP_TRY(hipEventRecord(c->ev[EV_P2], s));
    }
    if (with_probe && !pl.joins.empty())
        VJ_LAUNCH(launch_probe(a, pl.shape, c->cus * probe_blocks_per_cu(pl.shape), s));
    if (events) VJ_HIP_TRY(hipEventRecord(c->ev[EV_PROBE], s));
    return Err::ok();
}

// Everything a captured pipeline bakes in: the kernel arguments (arena
// pointers, totals, hash), which stages run, the probe grid, and the copies.
std::vector<uint64_t> graph_signature(const vj_ctx *c, const Plan &pl, const PipeArgs &a) {
    std::vector<uint64_t> sig((sizeof(PipeArgs) + 7) / 8 + 10, 0);  // zeroed: PipeArgs padding is zero too
    memcpy(sig.data(), &a, sizeof(PipeArgs));
    uint64_t *x = sig.data() + (sizeof(PipeArgs) + 7) / 8;
    x[0] = (uint64_t)pl.any_part | (uint64_t)pl.any_p2 << 1 | (uint64_t)pl.regions << 2;
    x[1] = (uint64_t)pl.shape.hp | (uint64_t)pl.shape.kw << 8 | (uint64_t)pl.shape.nc << 16 |
           (uint64_t)pl.shape.tile << 24 | (uint64_t)pl.shape.table << 32;
    x[2] = (uint64_t)c->cus * (uint64_t)probe_blocks_per_cu(pl.shape);
    x[3] = pl.upload_bytes;
    x[4] = pl.off_count;
    x[5] = pl.joins.size();
    x[6] = reinterpret_cast<uintptr_t>(c->pinned);
    x[7] = reinterpret_cast<uintptr_t>(c->meta.p);
    x[8] = pl.colscan.size();
    x[9] = (uint64_t)c->device;
    return sig;
}

void drop_graph(vj_ctx *c) {
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_sig.clear();
}

// Captures upload -> all stages -> count download on the context's capture
// stream (capture needs a stream other than the null stream) and instantiates
// it. The upload and the download are kernels that read / write the pinned
// staging buffer (no copy nodes: a replay stays on the compute queue). Events
// are recorded around the graph launch, not inside the graph (event nodes
// inside a replayed graph gave stage times that did not add up).
Err capture_pipeline(vj_ctx *c, const Plan &pl, const PipeArgs &a) {
    drop_graph(c);
    if (!c->capture_stream) VJ_HIP_TRY(hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking));
    hipStream_t cs = c->capture_stream;
    char *h = static_cast<char *>(c->pinned);
    VJ_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    hipError_t he = launch_stage_copy(c->meta.p, h, pl.upload_bytes, cs);
    Err e = he == hipSuccess ? launch_stages(c, pl, a, cs, true, false)
                             : Err::make(VJ_ERR_HIP, std::string("capture upload: ") + hipGetErrorString(he));
    if (!e.failed()) {
        he = launch_stage_copy(h + pl.off_count, static_cast<char *>(c->meta.p) + pl.off_count, pl.down_bytes(), cs);
        if (he != hipSuccess) e = Err::make(VJ_ERR_HIP, std::string("capture download: ") + hipGetErrorString(he));
    }
    hipGraph_t g = nullptr;
    he = hipStreamEndCapture(cs, &g);
    if (!e.failed() && he != hipSuccess) e = Err::make(VJ_ERR_HIP, std::string("end capture: ") + hipGetErrorString(he));
    if (!e.failed()) {
        he = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        if (he != hipSuccess) {
            c->graph = nullptr;
            e = Err::make(VJ_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(he));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    if (e.failed()) (void)hipGetLastError();
    c->graph_captures++;
    return e;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

// Algorithmic bytes of each stage of a one-join plan: the histogram reads the
// keys of every side it partitions; pass 1 reads keys (and explicit row ids)
// and writes keys + row ids; pass 2 reads and writes keys + row ids; build +
// probe reads every tuple once and writes 8 bytes per pair. Resident build
// sides were partitioned earlier and only count in build + probe.
void stage_bytes(const Plan &pl, uint64_t matches, vj_join_summary &s) {
    const uint64_t kb = (uint64_t)pl.kb;
    for (const JobDesc &jd : pl.jobs) {
        const uint64_t n = jd.n, ids_in = jd.in.rowids ? 4 : 0;
        const bool resident = jd.res_base != nullptr;
        if (jd.b1 && !resident) {
            if (!pl.regions) s.bytes_histogram += n * kb;
            s.bytes_pass1 += n * (kb + ids_in) + n * (kb + 4);
        }
        if (jd.b2 && !resident) s.bytes_pass2 += 2 * n * (kb + 4);
        s.bytes_build_probe += jd.b1 || resident ? n * (kb + 4) : n * (kb + ids_in);
    }
    s.bytes_build_probe += matches * 8;
    s.bytes_total = s.bytes_histogram + s.bytes_pass1 + s.bytes_pass2 + s.bytes_build_probe;
}

double gbps(uint64_t bytes, double ms) { return ms > 0 ? (double)bytes / (ms * 1e-3) / 1e9 : 0.0; }

void fill_summary(vj_ctx *c, const Plan &pl, const JoinSpec &js, uint64_t matches, uint64_t tasks, int retries,
                  bool replay, uint64_t scratch_bytes) {
    vj_join_summary &s = c->summary;
    memset(&s, 0, sizeof(s));
    s.build_rows = js.res ? js.res->n : js.r->n_rows;
    s.probe_rows = js.s->n_rows;
    s.matches = matches;
    const JobDesc &jr = pl.jobs[0];
    s.passes = jr.b2 ? 2 : (jr.b1 ? 1 : 0);
    s.bits1 = jr.b1;
    s.bits2 = jr.b2;
    s.retries = retries;
    s.partitions = pl.joins[0].nparts;
    s.tasks = tasks;
    s.ms_total = elapsed(c->ev[EV_START], c->ev[EV_PROBE]);
    if (c->profiling) {
        s.ms_histogram = elapsed(c->ev[EV_START], c->ev[EV_HIST]);
        s.ms_pass1 = elapsed(c->ev[EV_SCAN], c->ev[EV_P1]);
        s.ms_pass2 = elapsed(c->ev[EV_P1], c->ev[EV_P2]);
        s.ms_build_probe = elapsed(c->ev[EV_P2], c->ev[EV_PROBE]);
    }
    s.scratch_bytes = scratch_bytes;
    s.graph_replay = replay ? 1 : 0;
    s.tile_rows = pass_tile(pl.kb, pl.shape.tile);
    s.probe_threads = probe_threads(pl.shape.table);
    s.table_rows = probe_cap(pl.kb, pl.shape.table);
    s.layout = pl.regions ? VJ_LAYOUT_REGIONS : VJ_LAYOUT_EXACT;
    stage_bytes(pl, matches, s);
    s.gbps_histogram = gbps(s.bytes_histogram, s.ms_histogram);
    s.gbps_pass1 = gbps(s.bytes_pass1, s.ms_pass1);
    s.gbps_pass2 = gbps(s.bytes_pass2, s.ms_pass2);
    s.gbps_build_probe = gbps(s.bytes_build_probe, s.ms_build_probe);
    s.gbps_total = gbps(s.bytes_total, s.ms_total);
    s.probe_tuples_per_s = s.ms_total > 0 ? (double)s.probe_rows / (s.ms_total * 1e-3) : 0.0;
}

vj_status fail(vj_ctx *c, const Err &e) {
    if (c) c->last_error = e.msg;
    return e.code;
}

Err set_device(vj_ctx *c) {
    VJ_HIP_TRY(hipSetDevice(c->device));
    return Err::ok();
}

// Runs the pipeline for `specs`; counts[i] receives join i's match count.
// When `grow` is set, joins whose output overflowed get a fresh buffer of the
// exact size (hipMalloc, owned by the caller) and only build+probe is rerun.
// Graph use of a pipeline: never, always (vj_join_batch with use_graph), or
// automatically when the same launch signature comes twice in a row.
enum GraphMode { GRAPH_OFF, GRAPH_ON, GRAPH_AUTO };

Err run_joins(vj_ctx *c, std::vector<JoinSpec> &specs, const vj_join_opts *o, hipStream_t s,
              std::vector<uint64_t> &counts, bool grow, uint64_t extra_bytes, int *retries_out,
              GraphMode gmode = GRAPH_OFF) {
    Plan pl;
    VJ_TRY(build_join_plan(pl, specs, o, c));
    VJ_TRY(admit(c, pl, extra_bytes + pl.keep_bytes));
    VJ_TRY(allocate(c, pl));
    PipeArgs a = make_args(ctx_arenas(c), pl);
    char *h = static_cast<char *>(c->pinned);
    unsigned long long *hc = reinterpret_cast<unsigned long long *>(h + pl.off_count);
    bool replay = false;
    // automatic replay only without per-stage timing (it needs events between the stages)
    if (gmode == GRAPH_AUTO && c->profiling) gmode = GRAPH_OFF;
    if (gmode != GRAPH_OFF && !c->graph_failed) {
        std::vector<uint64_t> sig = graph_signature(c, pl, a);
        if (c->graph && sig == c->graph_sig) {
            replay = true;
        } else if (gmode == GRAPH_ON || sig == c->last_sig) {
            // the previous use of the pinned buffer has completed (every call ends synchronised)
            stage_descriptors(static_cast<char *>(c->pinned), pl);
            const Err ce = capture_pipeline(c, pl, a);
            if (ce.failed()) {
                if (gmode == GRAPH_ON) return ce;
                c->graph_failed = true;  // automatic use only: stay eager from now on
            } else {
                c->graph_sig = std::move(sig);
                replay = true;
            }
        } else {
            c->last_sig = std::move(sig);
        }
    }
    if (replay) {
        stage_descriptors(static_cast<char *>(c->pinned), pl);  // the graph's upload node reads them at launch
        VJ_HIP_TRY(hipEventRecord(c->ev[EV_START], s));
        VJ_HIP_TRY(hipGraphLaunch(c->graph, s));
        VJ_HIP_TRY(hipEventRecord(c->ev[EV_PROBE], s));
        c->graph_replays++;
    } else {
        VJ_TRY(upload(ctx_arenas(c), pl, s));
        VJ_TRY(launch_stages(c, pl, a, s, true));
        VJ_HIP_TRY(hipMemcpyAsync(hc, static_cast<char *>(c->meta.p) + pl.off_count, pl.down_bytes(),
                                  hipMemcpyDeviceToHost, s));
    }
    VJ_HIP_TRY(hipStreamSynchronize(s));
    if (pl.regions && reinterpret_cast<const Ctl *>(h + pl.off_ctl)->overflow) {
        // a region overflowed (the keys do not spread evenly): the same joins
        // with the exact layout; an autotuned size class keeps to it from now on
        if (!o->layout) c->region_overflowed.insert(tune_key(largest_build(specs), pl.kb, pl.shape.hp));
        vj_join_opts ox = *o;
        ox.layout = VJ_LAYOUT_EXACT;
        VJ_TRY(run_joins(c, specs, &ox, s, counts, grow, extra_bytes, retries_out, gmode));
        if (specs.size() == 1) c->summary.layout_fallback = 1;  // the summary describes single joins
        return Err::ok();
    }
    counts.assign(hc, hc + pl.joins.size());
    int retries = 0;
    if (grow) {
        // the regrown outputs are admitted like the first ones: the context's
        // arenas plus every output it allocates must fit the budget
        uint64_t out_bytes = 0;
        for (size_t i = 0; i < specs.size(); i++) out_bytes += std::max<uint64_t>(counts[i], specs[i].cap) * 8;
        if (c->budget && held_bytes(c) + out_bytes > c->budget)
            return Err::make(VJ_ERR_BUDGET, "the join has " + std::to_string(counts[0]) +
                                                " pairs; its output does not fit the HBM budget");
        bool again = false;
        for (size_t i = 0; i < specs.size(); i++) {
            if (counts[i] > specs[i].cap) {
                if (specs[i].out_r) (void)hipFree(specs[i].out_r);
                if (specs[i].out_s) (void)hipFree(specs[i].out_s);
                specs[i].out_r = specs[i].out_s = nullptr;
                if (hipMalloc(&specs[i].out_r, counts[i] * 4) != hipSuccess ||
                    hipMalloc(&specs[i].out_s, counts[i] * 4) != hipSuccess) {
                    (void)hipGetLastError();
                    return Err::make(VJ_ERR_OUT_OF_MEMORY, "output allocation failed");
                }
                specs[i].cap = counts[i];
                pl.joins[i].out_r = specs[i].out_r;
                pl.joins[i].out_s = specs[i].out_s;
                pl.joins[i].cap = specs[i].cap;
                again = true;
            }
        }
        if (again) {
            retries = 1;
            memcpy(h + pl.off_joins, pl.joins.data(), pl.joins.size() * sizeof(JoinDesc));
            VJ_HIP_TRY(hipMemcpyAsync(static_cast<char *>(c->meta.p) + pl.off_joins, h + pl.off_joins,
                                      pl.joins.size() * sizeof(JoinDesc), hipMemcpyHostToDevice, s));
            VJ_LAUNCH(launch_reset_probe(a, s));
            VJ_LAUNCH(launch_probe(a, pl.shape, c->cus * probe_blocks_per_cu(pl.shape), s));
            VJ_HIP_TRY(hipEventRecord(c->ev[EV_PROBE], s));
            VJ_HIP_TRY(hipMemcpyAsync(hc, static_cast<char *>(c->meta.p) + pl.off_count, pl.down_bytes(),
                                      hipMemcpyDeviceToHost, s));
            VJ_HIP_TRY(hipStreamSynchronize(s));
            counts.assign(hc, hc + pl.joins.size());
        }
    }
    if (retries_out) *retries_out = retries;
    if (specs.size() == 1)
        fill_summary(c, pl, specs[0], counts[0], reinterpret_cast<const Ctl *>(h + pl.off_ctl)->task_total, retries,
                     replay, held_bytes(c));
    return Err::ok();
}

// vj_join's body: a callee-allocated output of the hinted size (or the probe
// rows), regrown when the join has more pairs. `held_extra` is device memory
// the caller already holds for this join (counted against the budget).
Err join_callee_alloc(vj_ctx *c, const vj_table *build, const vj_table *probe, const vj_join_opts *opts, hipStream_t s,
                      uint64_t held_extra, vj_pairs *out) {
    const uint64_t cap = opts->out_capacity_hint ? opts->out_capacity_hint : std::max<uint64_t>(probe->n_rows, 1);
    if (c->budget && cap * 8 + held_extra > c->budget)  // the output alone breaks the budget: nothing is allocated
        return Err::make(VJ_ERR_BUDGET, "an output of " + std::to_string(cap) + " pairs exceeds the HBM budget");
    std::vector<JoinSpec> specs{{build, probe, nullptr, nullptr, cap}};
    if (hipMalloc(&specs[0].out_r, cap * 4) != hipSuccess || hipMalloc(&specs[0].out_s, cap * 4) != hipSuccess) {
        (void)hipGetLastError();
        if (specs[0].out_r) (void)hipFree(specs[0].out_r);
        return Err::make(VJ_ERR_OUT_OF_MEMORY, "output allocation failed");
    }
    std::vector<uint64_t> counts;
    const Err e = run_joins(c, specs, opts, s, counts, true, cap * 8 + held_extra, nullptr);
    if (e.failed()) {
        if (specs[0].out_r) (void)hipFree(specs[0].out_r);
        if (specs[0].out_s) (void)hipFree(specs[0].out_s);
        return e;
    }
    out->count = counts[0];
    out->build_rowids = specs[0].out_r;
    out->probe_rowids = specs[0].out_s;
    return Err::ok();
}

}  // namespace

namespace vj {

Err pipeline_bytes(const std::vector<JoinSpec> &specs, const vj_join_opts *o, uint64_t *bytes) {
    Plan pl;
    VJ_TRY(build_join_plan(pl, specs, o, nullptr, false));
    // a lane launches under a budget: its arenas hold at least these sizes,
    // and growth slack only where the budget leaves room for it
    *bytes = plan_scratch_bytes(pl) + pl.keep_bytes;
    return Err::ok();
}

Err pipeline_launch(vj_ctx *c, std::vector<JoinSpec> &specs, const vj_join_opts *o, hipStream_t s, uint64_t budget,
                    size_t *count_off) {
    VJ_TRY(set_device(c));
    Plan pl;
    VJ_TRY(build_join_plan(pl, specs, o, c, false));  // served batches complete asynchronously: exact layout
    c->budget = budget;
    VJ_TRY(admit(c, pl, pl.keep_bytes));
    VJ_TRY(allocate(c, pl, true, pl.keep_bytes));
    VJ_TRY(upload(ctx_arenas(c), pl, s));
    PipeArgs a = make_args(ctx_arenas(c), pl);
    VJ_TRY(launch_stages(c, pl, a, s, true));
    char *h = static_cast<char *>(c->pinned);
    VJ_HIP_TRY(hipMemcpyAsync(h + pl.off_count, static_cast<char *>(c->meta.p) + pl.off_count, pl.down_bytes(),
                              hipMemcpyDeviceToHost, s));
    *count_off = pl.off_count;
    return Err::ok();
}

void pipeline_counts(const vj_ctx *c, size_t count_off, size_t n, std::vector<uint64_t> &counts) {
    const unsigned long long *hc =
        reinterpret_cast<const unsigned long long *>(static_cast<const char *>(c->pinned) + count_off);
    counts.assign(hc, hc + n);
}

uint64_t ctx_bytes(const vj_ctx *c) { return held_bytes(c); }
void ctx_release(vj_ctx *c) {
    drop_graph(c);
    free_all(c);
}
void built_release(vj_built *b) { free_built(b); }

}  // namespace vj

extern "C" {

vj_status vj_ctx_create(int device, vj_ctx **out) {
    if (!out) return VJ_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return VJ_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return VJ_ERR_INVALID_ARGUMENT;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return VJ_ERR_NO_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return VJ_ERR_NO_DEVICE;  // built for gfx950 only
    if (hipSetDevice(device) != hipSuccess) return VJ_ERR_HIP;
    vj_ctx *c = new vj_ctx();
    c->device = device;
    c->cus = device_cu_count(device);
    for (int i = 0; i < EV_COUNT; i++) {
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            for (int k = 0; k < i; k++) (void)hipEventDestroy(c->ev[k]);
            delete c;
            return VJ_ERR_HIP;
        }
    }
    // set-up of the no-allocation entry point (its failure only disables that entry point)
    const Err ne = noalloc_scratch(device, &c->na_scratch);
    if (ne.failed()) {
        c->na_error = ne.msg;
    } else if (hipHostMalloc(&c->na_stage, kNoAllocStageBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->na_stage = nullptr;
        c->na_error = "hipHostMalloc of the staging block failed";
    }
    *out = c;
    return VJ_OK;
}

vj_status vj_ctx_destroy(vj_ctx *c) {
    if (!c) return VJ_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    drop_graph(c);
    if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
    free_all(c);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->na_stage) (void)hipHostFree(c->na_stage);
    for (int i = 0; i < EV_COUNT; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    delete c;
    return VJ_OK;
}

vj_status vj_ctx_set_budget(vj_ctx *c, uint64_t bytes) {
    if (!c) return VJ_ERR_INVALID_ARGUMENT;
    c->budget = bytes;
    return VJ_OK;
}

vj_status vj_ctx_memory(vj_ctx *c, uint64_t *bytes) {
    if (!c || !bytes) return VJ_ERR_INVALID_ARGUMENT;
    *bytes = held_bytes(c);
    return VJ_OK;
}

vj_status vj_ctx_set_profiling(vj_ctx *c, int enabled) {
    if (!c) return VJ_ERR_INVALID_ARGUMENT;
    c->profiling = enabled != 0;
    return VJ_OK;
}

const char *vj_last_error(const vj_ctx *c) { return c ? c->last_error.c_str() : ""; }

void vj_join_opts_default(vj_join_opts *o) {
    if (!o) return;
    vj_hash_opts_default(&o->hash);
    o->passes = -1;
    o->bits1 = 0;
    o->bits2 = 0;
    o->one_based = 0;
    o->out_capacity_hint = 0;
    o->tile = 0;
    o->table = 0;
    o->layout = 0;
    o->reserved = 0;
}

vj_status vj_join_into(vj_ctx *c, const vj_table *build, const vj_table *probe, const vj_join_opts *opts,
                       uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count, void *stream) {
    if (!c || !count) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    *count = 0;
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    if (capacity && (!out_build || !out_probe)) return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "output is NULL"));
    Err e = check_pair(build, probe);
    if (e.failed()) return fail(c, e);
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    std::vector<JoinSpec> specs{{build, probe, out_build, out_probe, capacity}};
    std::vector<uint64_t> counts;
    e = run_joins(c, specs, opts, (hipStream_t)stream, counts, false, 0, nullptr, GRAPH_AUTO);
    if (e.failed()) return fail(c, e);
    *count = counts[0];
    if (counts[0] > capacity)
        return fail(c, Err::make(VJ_ERR_CAPACITY, "output holds " + std::to_string(capacity) + " pairs, join has " +
                                                      std::to_string(counts[0])));
    return VJ_OK;
}

vj_status vj_join_into_noalloc(vj_ctx *c, const vj_table *build, const vj_table *probe, const vj_join_opts *opts,
                               uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count,
                               void *stream) {
    if (!c || !count) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    *count = 0;
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    if (capacity && (!out_build || !out_probe)) return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "output is NULL"));
    Err e = check_pair(build, probe);
    if (e.failed()) return fail(c, e);
    if (!c->na_scratch || !c->na_stage)
        return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "no-allocation scratch unavailable: " + c->na_error));
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    hipStream_t s = (hipStream_t)stream;
    std::vector<JoinSpec> specs{{build, probe, out_build, out_probe, capacity}};
    Plan pl;
    e = build_join_plan(pl, specs, opts, c, false);  // exact layout: its scratch is bounded by the join's rows
    if (e.failed()) return fail(c, e);
    // the arenas, laid out back to back in the fixed scratch
    const ArenaSizes az = arena_sizes(pl);
    if (az.sum() > kNoAllocScratchBytes)
        return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "this join needs " + std::to_string(az.sum()) +
                                                         " bytes of scratch; the no-allocation entry point has " +
                                                         std::to_string(kNoAllocScratchBytes) +
                                                         " (vj_join_into has no such limit)"));
    if (pl.meta_bytes > kNoAllocStageBytes)
        return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "this join's descriptors exceed the no-allocation staging block"));
    char *base = static_cast<char *>(c->na_scratch);
    Arenas ar;
    ar.counts = base + az.tuples;
    ar.base = ar.counts + az.counts;
    ar.ptask = ar.base + az.base;
    ar.meta = ar.ptask + az.ptask;
    ar.pinned = static_cast<char *>(c->na_stage);
    place_tuples(pl, base);
    const PipeArgs a = make_args(ar, pl);
    e = upload(ar, pl, s);
    if (!e.failed()) e = launch_stages(c, pl, a, s, true);
    if (!e.failed()) {
        hipError_t he = hipMemcpyAsync(ar.pinned + pl.off_count, ar.meta + pl.off_count, pl.down_bytes(),
                                       hipMemcpyDeviceToHost, s);
        if (he == hipSuccess) he = hipStreamSynchronize(s);
        if (he != hipSuccess) e = Err::make(VJ_ERR_HIP, std::string("no-allocation join: ") + hipGetErrorString(he));
    }
    if (e.failed()) return fail(c, e);
    const uint64_t m = *reinterpret_cast<const unsigned long long *>(ar.pinned + pl.off_count);
    fill_summary(c, pl, specs[0], m, reinterpret_cast<const Ctl *>(ar.pinned + pl.off_ctl)->task_total, 0, false,
                 az.sum());
    *count = m;
    if (m > capacity)
        return fail(c, Err::make(VJ_ERR_CAPACITY, "output holds " + std::to_string(capacity) + " pairs, join has " +
                                                      std::to_string(m)));
    return VJ_OK;
}

vj_status vj_join(vj_ctx *c, const vj_table *build, const vj_table *probe, const vj_join_opts *opts, vj_pairs *out,
                  void *stream) {
    if (!c || !out) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    memset(out, 0, sizeof(*out));
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    Err e = check_pair(build, probe);
    if (e.failed()) return fail(c, e);
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    e = join_callee_alloc(c, build, probe, opts, (hipStream_t)stream, 0, out);
    return e.failed() ? fail(c, e) : VJ_OK;
}

vj_status vj_join_rows(vj_ctx *c, const vj_rows *r, const vj_rows *s, const vj_join_opts *opts, vj_rows *rs,
                       void *stream) {
    if (!c || !r || !s || !rs) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    memset(rs, 0, sizeof(*rs));
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    // join_gpu asserts equal column counts (join_provider.cuh:291)
    if (r->n_cols != s->n_cols || r->n_cols < 1 || r->n_cols > VJ_MAX_KEY_COLS)
        return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "both tables need the same number (1..4) of key columns"));
    for (const vj_rows *t : {r, s}) {
        if (t->n_rows > VJ_MAX_ROWS)
            return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "tables must have at most 2^32 - 2^16 rows"));
        if (t->n_rows && (!t->column_values || !t->primary_keys))
            return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "column_values / primary_keys is NULL"));
    }
    Err e = set_device(c);
    if (e.failed()) return fail(c, e);
    hipStream_t st = (hipStream_t)stream;
    const int nc = r->n_cols;
    // Key columns for the pipeline: one column of a row-major table is already
    // contiguous; wider rows are transposed into temporary columns.
    uint64_t *tmp = nullptr;
    const uint64_t tmp_words = nc > 1 ? (r->n_rows + s->n_rows) * (uint64_t)nc : 0;
    if (tmp_words) {
        if (c->budget && held_bytes(c) + tmp_words * 8 > c->budget)
            return fail(c, Err::make(VJ_ERR_BUDGET, "the transposed key columns exceed the HBM budget"));
        if (hipMalloc(&tmp, tmp_words * 8) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, Err::make(VJ_ERR_OUT_OF_MEMORY, "hipMalloc of the transposed key columns failed"));
        }
        e = launch_rows_to_cols(r->column_values, r->n_rows, nc, tmp, st, c->cus);
        if (!e.failed()) e = launch_rows_to_cols(s->column_values, s->n_rows, nc, tmp + r->n_rows * nc, st, c->cus);
    }
    vj_table rt{}, stt{};
    for (int k = 0; k < 2; k++) {
        const vj_rows *src = k ? s : r;
        vj_table &t = k ? stt : rt;
        t.n_rows = src->n_rows;
        t.n_cols = nc;
        t.key_width = 8;
        uint64_t *base = nc > 1 ? tmp + (k ? r->n_rows * nc : 0) : src->column_values;
        for (int cc = 0; cc < nc; cc++) t.cols[cc] = nc > 1 ? base + (uint64_t)cc * src->n_rows : base;
    }
    vj_pairs pairs{};
    if (!e.failed()) e = join_callee_alloc(c, &rt, &stt, opts, st, tmp_words * 8, &pairs);
    const uint64_t m = pairs.count;
    if (!e.failed() && m) {  // rs: (r pk, s pk) per pair, row-major, and primary keys 1..M
        if (c->budget && held_bytes(c) + tmp_words * 8 + m * 8 + m * 24 > c->budget) {
            e = Err::make(VJ_ERR_BUDGET, "the joined table exceeds the HBM budget");
        } else if (hipMalloc(&rs->column_values, m * 16) != hipSuccess ||
                   hipMalloc(&rs->primary_keys, m * 8) != hipSuccess) {
            (void)hipGetLastError();
            e = Err::make(VJ_ERR_OUT_OF_MEMORY, "hipMalloc of the joined table failed");
        } else {
            e = launch_materialize(pairs.build_rowids, pairs.probe_rowids, m, r->primary_keys, s->primary_keys,
                                   rs->column_values, rs->primary_keys, st, c->cus);
        }
    }
    if (!e.failed()) {
        const hipError_t he = hipStreamSynchronize(st);
        if (he != hipSuccess) e = Err::make(VJ_ERR_HIP, std::string("vj_join_rows: ") + hipGetErrorString(he));
    } else {
        (void)hipStreamSynchronize(st);  // nothing queued may still use what is freed below
    }
    (void)vj_pairs_free(&pairs);
    if (tmp) (void)hipFree(tmp);
    if (e.failed()) {
        (void)vj_rows_free(rs);
        return fail(c, e);
    }
    rs->n_rows = m;
    rs->n_cols = 2;
    return VJ_OK;
}

vj_status vj_rows_free(vj_rows *t) {
    if (!t) return VJ_ERR_INVALID_ARGUMENT;
    if (t->column_values) (void)hipFree(t->column_values);
    if (t->primary_keys) (void)hipFree(t->primary_keys);
    t->column_values = t->primary_keys = nullptr;
    t->n_rows = 0;
    return VJ_OK;
}

vj_status vj_pairs_free(vj_pairs *p) {
    if (!p) return VJ_ERR_INVALID_ARGUMENT;
    if (p->build_rowids) (void)hipFree(p->build_rowids);
    if (p->probe_rowids) (void)hipFree(p->probe_rowids);
    p->build_rowids = p->probe_rowids = nullptr;
    p->count = 0;
    return VJ_OK;
}

vj_status vj_join_summary_get(vj_ctx *c, vj_join_summary *s) {
    if (!c || !s) return VJ_ERR_INVALID_ARGUMENT;
    *s = c->summary;
    return VJ_OK;
}

vj_status vj_partition(vj_ctx *c, const vj_table *table, const vj_hash_opts *hash, int32_t shift, int32_t bits,
                       const vj_table *out, uint32_t *hist, uint32_t *offsets, void *stream) {
    if (!c) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    Err e = check_table(table);
    if (e.failed()) return fail(c, e);
    e = check_table(out);
    if (e.failed()) return fail(c, e);
    if (bits < 1 || bits > kMaxPassBits) return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "bits must be 1..8"));
    if (shift < 0 || shift > 63) return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "shift must be 0..63"));
    if (out->n_cols != table->n_cols || out->key_width != table->key_width || out->n_rows < table->n_rows ||
        !out->rowids || !hist || !offsets)
        return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "output table must match the input and carry row ids"));
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    hipStream_t s = (hipStream_t)stream;
    Plan pl;
    e = make_hash_cfg(hash, table, &pl.hash);
    if (e.failed()) return fail(c, e);
    pl.kb = table->key_width * table->n_cols;
    pl.shape.hp = hash_path(pl.hash);
    pl.shape.kw = table->key_width;
    pl.shape.nc = table->n_cols;
    add_job(pl, table, 1, bits, 0, (uint32_t)shift);
    finish_layout(pl);
    pl.tuples_bytes = 0;  // writes into the caller's table
    pl.fin_off[0] = SIZE_MAX;
    e = admit(c, pl, 0);
    if (!e.failed()) e = allocate(c, pl);
    if (e.failed()) return fail(c, e);
    JobDesc &jd = pl.jobs[0];
    for (int cc = 0; cc < 4; cc++) jd.fin_cols[cc] = cc < out->n_cols ? const_cast<void *>(out->cols[cc]) : nullptr;
    jd.fin_ids = const_cast<uint32_t *>(out->rowids);
    e = upload(ctx_arenas(c), pl, s);
    if (e.failed()) return fail(c, e);
    PipeArgs a = make_args(ctx_arenas(c), pl);
    e = launch_stages(c, pl, a, s, false);
    if (e.failed()) return fail(c, e);
    hipError_t he = launch_emit_hist(a.base + jd.base_off, 1u << bits, hist, offsets, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) return fail(c, Err::make(VJ_ERR_HIP, hipGetErrorString(he)));
    return VJ_OK;
}

vj_status vj_join_batch(vj_ctx *c, int32_t n_joins, const vj_table *builds, const vj_table *probes,
                        const vj_join_opts *opts, uint32_t *const *out_build, uint32_t *const *out_probe,
                        const uint64_t *caps, uint64_t *counts, int32_t use_graph, void *stream) {
    if (!c || n_joins <= 0 || !builds || !probes || !out_build || !out_probe || !caps || !counts)
        return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    Err e = set_device(c);
    if (e.failed()) return fail(c, e);
    std::vector<JoinSpec> specs((size_t)n_joins);
    for (int i = 0; i < n_joins; i++) {
        specs[i] = {&builds[i], &probes[i], out_build[i], out_probe[i], caps[i]};
        if (caps[i] && (!out_build[i] || !out_probe[i]))
            return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "output is NULL"));
    }
    std::vector<uint64_t> cnt;
    e = run_joins(c, specs, opts, (hipStream_t)stream, cnt, false, 0, nullptr, use_graph ? GRAPH_ON : GRAPH_OFF);
    if (e.failed()) return fail(c, e);
    bool over = false;
    for (int i = 0; i < n_joins; i++) {
        counts[i] = cnt[i];
        over |= cnt[i] > caps[i];
    }
    if (over) return fail(c, Err::make(VJ_ERR_CAPACITY, "an output buffer of the batch is too small"));
    return VJ_OK;
}

vj_status vj_build(vj_ctx *c, const vj_table *build, const vj_join_opts *opts, vj_built **out, void *stream) {
    if (!c || !out) return VJ_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    c->last_error.clear();
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    Err e = check_table(build);
    if (e.failed()) return fail(c, e);
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    hipStream_t s = (hipStream_t)stream;
    vj_built *b = new vj_built();
    Plan pl;
    e = make_hash_cfg(&opts->hash, build, &pl.hash);
    int P = 0, b1 = 0, b2 = 0;
    if (!e.failed()) {
        pl.kb = build->key_width * build->n_cols;
        pl.shape.hp = hash_path(pl.hash);
        pl.shape.kw = build->key_width;
        pl.shape.nc = build->n_cols;
        e = check_shapes(opts);
        if (!e.failed()) {
            pl.shape.tile = opts->tile ? opts->tile - 1 : 0;
            pl.shape.table = opts->table ? opts->table - 1 : (need_bits(build->n_rows, build_cap(pl.kb)) > kMaxHistBits ? 1 : 0);
            e = plan_bits(build->n_rows, pl.kb, pl.shape.table, opts, &P, &b1, &b2);
        }
    }
    if (!e.failed()) {
        if (P == 0) {
            P = 1;
            b1 = 1;
            b2 = 0;
        }
        add_job(pl, build, P, b1, b2, 0, true);
        b->n = build->n_rows;
        b->n_cols = build->n_cols;
        b->key_width = build->key_width;
        b->key_signed = build->key_signed != 0;
        b->hash = opts->hash;
        b->b1 = (uint32_t)b1;
        b->b2 = (uint32_t)(P >= 2 ? b2 : 0);
        b->device = c->device;
        pl.keeps.push_back({0u, b});
        pl.keep_bytes = built_bytes(b->n, b->n_cols, b->key_width, b->b1, b->b2);
        finish_layout(pl);
        e = admit(c, pl, pl.keep_bytes);
    }
    if (!e.failed()) e = allocate(c, pl);
    if (!e.failed()) e = upload(ctx_arenas(c), pl, s);
    if (!e.failed()) {
        PipeArgs a = make_args(ctx_arenas(c), pl);
        e = launch_stages(c, pl, a, s, false);
    }
    if (!e.failed()) {
        hipError_t he = hipStreamSynchronize(s);
        if (he != hipSuccess) e = Err::make(VJ_ERR_HIP, hipGetErrorString(he));
    }
    if (e.failed()) {
        free_built(b);
        delete b;
        return fail(c, e);
    }
    *out = b;
    return VJ_OK;
}

vj_status vj_probe_into(vj_ctx *c, const vj_built *built, const vj_table *probe, const vj_join_opts *opts,
                        uint32_t *out_build, uint32_t *out_probe, uint64_t capacity, uint64_t *count, void *stream) {
    if (!c || !built || !count) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    *count = 0;
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    if (capacity && (!out_build || !out_probe)) return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "output is NULL"));
    if (built->device != c->device) return fail(c, Err::make(VJ_ERR_INVALID_ARGUMENT, "table is on another device"));
    Err e = set_device(c);
    if (e.failed()) return fail(c, e);
    std::vector<JoinSpec> specs{{nullptr, probe, out_build, out_probe, capacity}};
    specs[0].res = built;
    std::vector<uint64_t> counts;
    e = run_joins(c, specs, opts, (hipStream_t)stream, counts, false, 0, nullptr);
    if (e.failed()) return fail(c, e);
    *count = counts[0];
    if (counts[0] > capacity)
        return fail(c, Err::make(VJ_ERR_CAPACITY, "output holds " + std::to_string(capacity) + " pairs, join has " +
                                                      std::to_string(counts[0])));
    return VJ_OK;
}

vj_status vj_built_bytes(const vj_built *b, uint64_t *bytes) {
    if (!b || !bytes) return VJ_ERR_INVALID_ARGUMENT;
    *bytes = b->bytes;
    return VJ_OK;
}

vj_status vj_built_free(vj_built *b) {
    if (!b) return VJ_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(b->device);
    free_built(b);
    delete b;
    return VJ_OK;
}

vj_status vj_tune(vj_ctx *c, const vj_table *build, const vj_table *probe, const vj_join_opts *opts, int32_t reps,
                  vj_join_opts *best, double *best_ms, void *stream) {
    if (!c || !best || reps < 1) return VJ_ERR_INVALID_ARGUMENT;
    c->last_error.clear();
    vj_join_opts d;
    vj_join_opts_default(&d);
    if (!opts) opts = &d;
    Err e = check_pair(build, probe);
    if (e.failed()) return fail(c, e);
    e = set_device(c);
    if (e.failed()) return fail(c, e);
    hipStream_t s = (hipStream_t)stream;
    HashCfg hc;
    e = make_hash_cfg(&opts->hash, build, &hc);
    if (e.failed()) return fail(c, e);
    e = check_shapes(opts);
    if (e.failed()) return fail(c, e);
    const int kb = build->key_width * build->n_cols, hp = hash_path(hc);
    const uint64_t nr = build->n_rows;
    // Candidates: for each table shape, plans around the planner's choice for
    // that table's chunk (every pass count, bits -1..+1, both splits of two
    // passes; two-pass plans in both partition layouts), with the standard
    // tile; then the small tile with the fastest.
    struct Cand {
        int P, b1, b2, tile, table, layout;
    };
    std::vector<Cand> cands;
    auto add = [&](Cand x) {
        for (auto &y : cands)
            if (y.P == x.P && y.b1 == x.b1 && y.b2 == x.b2 && y.tile == x.tile && y.table == x.table &&
                y.layout == x.layout)
                return;
        if (opts->passes >= 0 && x.P != opts->passes) return;
        if (opts->layout && x.layout != opts->layout) return;
        cands.push_back(x);
    };
    // the caller's tile, else the standard tile (exact layout) or the large one (region layout)
    const int tile0 = opts->tile ? opts->tile - 1 : 0, tile_r = opts->tile ? opts->tile - 1 : 2;
    for (int table = 0; table <= 1; table++) {
        if (opts->table && table != opts->table - 1) continue;
        const int cap = probe_cap(kb, table);
        const int need = need_bits(nr, cap);
        if (nr <= (uint64_t)cap) add({0, 0, 0, tile0, table, VJ_LAYOUT_EXACT});
        for (int b = std::max(1, need - 1); b <= std::min(kMaxPassBits, need + 1); b++)
            add({1, b, 0, tile0, table, VJ_LAYOUT_EXACT});
        for (int tot = std::max(2, need - 1); tot <= std::min(kMaxHistBits, need + 1); tot++)
            for (int b1 : {(tot + 1) / 2, tot / 2}) {
                const int b2 = tot - b1;
                if (b1 >= 1 && b1 <= kMaxPassBits && b2 >= 1 && b2 <= kMaxPassBits) {
                    add({2, b1, b2, tile0, table, VJ_LAYOUT_EXACT});
                    if ((1u << b1) % kXcds == 0) add({2, b1, b2, tile_r, table, VJ_LAYOUT_REGIONS});
                }
            }
    }
    if (cands.empty()) return fail(c, Err::make(VJ_ERR_UNSUPPORTED, "no candidate plan with that pass count"));
    // output buffers sized by a counting run (the count does not depend on the plan)
    std::vector<JoinSpec> specs{{build, probe, nullptr, nullptr, 0}};
    std::vector<uint64_t> counts;
    c->tuned.erase(tune_key(nr, kb, hp));
    c->region_overflowed.erase(tune_key(nr, kb, hp));
    auto plan_opts = [&](const Cand &x) {
        vj_join_opts o2 = *opts;
        o2.passes = x.P;
        o2.bits1 = x.b1;
        o2.bits2 = x.b2;
        o2.tile = x.tile + 1;
        o2.table = x.table + 1;
        o2.layout = x.layout;
        return o2;
    };
    vj_join_opts oc = plan_opts(cands[0]);  // counting run: any valid plan
    e = run_joins(c, specs, &oc, s, counts, false, 0, nullptr);
    if (e.failed()) return fail(c, e);
    const uint64_t cap = std::max<uint64_t>(counts[0], 1);
    uint32_t *ob = nullptr, *op = nullptr;
    if (hipMalloc(&ob, cap * 4) != hipSuccess || hipMalloc(&op, cap * 4) != hipSuccess) {
        (void)hipGetLastError();
        if (ob) (void)hipFree(ob);
        return fail(c, Err::make(VJ_ERR_OUT_OF_MEMORY, "tuning output allocation failed"));
    }
    specs[0].out_r = ob;
    specs[0].out_s = op;
    specs[0].cap = cap;
    auto time_plan = [&](const Cand &x, double *med) {
        const vj_join_opts o2 = plan_opts(x);
        std::vector<double> ts;
        for (int r = 0; r <= reps; r++) {  // the first run warms the arenas
            Err er = run_joins(c, specs, &o2, s, counts, false, 0, nullptr);
            if (er.failed()) return er;
            if (r) ts.push_back(c->summary.ms_total);
        }
        std::sort(ts.begin(), ts.end());
        *med = ts[ts.size() / 2];
        return Err::ok();
    };
    double best_t = 1e30;
    Cand best_c = cands[0];
    for (const Cand &x : cands) {
        double t = 0;
        e = time_plan(x, &t);
        if (e.failed()) break;
        if (t < best_t) {
            best_t = t;
            best_c = x;
        }
    }
    if (!e.failed() && !opts->tile && best_c.P > 0) {  // the other pass tiles, for the fastest plan
        const Cand top = best_c;
        for (int tile = 0; tile < kTileShapes && !e.failed(); tile++) {
            if (tile == top.tile) continue;
            Cand x = top;
            x.tile = tile;
            double t = 0;
            e = time_plan(x, &t);
            if (!e.failed() && t < best_t) {
                best_t = t;
                best_c = x;
            }
        }
    }
    (void)hipFree(ob);
    (void)hipFree(op);
    if (e.failed()) return fail(c, e);
    c->tuned[tune_key(nr, kb, hp)] =
        tuned_word(best_c.P, best_c.b1, best_c.b2, best_c.tile, best_c.table, best_c.layout);
    *best = plan_opts(best_c);
    if (best_ms) *best_ms = best_t;
    return VJ_OK;
}

vj_status vj_tune_clear(vj_ctx *c) {
    if (!c) return VJ_ERR_INVALID_ARGUMENT;
    c->tuned.clear();
    c->region_overflowed.clear();
    return VJ_OK;
}

}  // extern "C"
