// Standalone hash kernel: vj_hash (replaces hash_func, src/hash/hash.cu:507-557).
//
// The reference walks a row-major buffer with one thread per row, so a wave
// reads 64 rows x row-width bytes with a stride; here key columns are SoA and
// each of the NC column loads of a wave is one contiguous 64 x KW-byte read.
// Every thread hashes UNROLL rows per grid-stride step to keep several loads
// in flight. The kernel is HBM-bound: bytes = n * (NC * KW + hash_bits / 8).
#include "vj_internal.hpp"

namespace vj {

namespace {

constexpr int kHashThreads = 256;
constexpr int kHashUnroll = 4;

template <int HP, int KW, int NC, typename OutT>
__global__ __launch_bounds__(kHashThreads) void k_hash(TView t, HashCfg h, OutT *out_) {
    // key columns and output as global-address-space pointers (global_load /
    // global_store, counted on vmcnt only), copied out of the view once
    GKeys<KW, NC> in;
#pragma unroll
    for (int c = 0; c < NC; c++) in.c[c] = to_global(static_cast<const typename KeyT<KW>::T *>(t.cols[c]));
    const gptr<OutT> out = to_global(out_);
    const uint64_t n = t.n;
    const uint64_t stride = (uint64_t)gridDim.x * kHashThreads * kHashUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * kHashThreads * kHashUnroll + threadIdx.x; base < n; base += stride) {
        Key<KW, NC> k[kHashUnroll];
#pragma unroll
        for (int u = 0; u < kHashUnroll; u++) {  // all loads in flight at once (past the end: row 0 again)
            const uint64_t i = base + (uint64_t)u * kHashThreads;
            k[u] = load_key<KW, NC>(in, i < n ? i : 0);
        }
#pragma unroll
        for (int u = 0; u < kHashUnroll; u++) {
            const uint64_t i = base + (uint64_t)u * kHashThreads;
            if (i < n) out[i] = (OutT)hash_key<HP, KW, NC>(k[u], h);
        }
    }
}

template <int HP, int KW, int NC>
Err launch_typed(const TView &v, const HashCfg &h, void *out, hipStream_t s, int cus) {
    uint64_t per_block = (uint64_t)kHashThreads * kHashUnroll;
    uint64_t want = (v.n + per_block - 1) / per_block;
    uint64_t cap = (uint64_t)cus * 16;
    unsigned blocks = (unsigned)(want < cap ? want : cap);
    if (blocks == 0) return Err::ok();
    if (h.bits == 32)
        hipLaunchKernelGGL((k_hash<HP, KW, NC, uint32_t>), dim3(blocks), dim3(kHashThreads), 0, s, v, h,
                           (uint32_t *)out);
    else
        hipLaunchKernelGGL((k_hash<HP, KW, NC, uint64_t>), dim3(blocks), dim3(kHashThreads), 0, s, v, h,
                           (uint64_t *)out);
    VJ_HIP_TRY(hipGetLastError());
    return Err::ok();
}

template <int HP, int KW>
Err launch_nc(int nc, const TView &v, const HashCfg &h, void *out, hipStream_t s, int cus) {
    switch (nc) {
    case 1: return launch_typed<HP, KW, 1>(v, h, out, s, cus);
    case 2: return launch_typed<HP, KW, 2>(v, h, out, s, cus);
    case 3: return launch_typed<HP, KW, 3>(v, h, out, s, cus);
    case 4: return launch_typed<HP, KW, 4>(v, h, out, s, cus);
    }
    return Err::make(VJ_ERR_INVALID_ARGUMENT, "n_cols must be 1..4");
}

template <int HP>
Err launch_kw(const vj_table *t, const TView &v, const HashCfg &h, void *out, hipStream_t s, int cus) {
    if (t->key_width == 8) return launch_nc<HP, 8>(t->n_cols, v, h, out, s, cus);
    return launch_nc<HP, 4>(t->n_cols, v, h, out, s, cus);
}

}  // namespace

Err launch_hash(const vj_table *t, const HashCfg &h, void *out, hipStream_t s, int cus) {
    TView v = to_view(t);
    switch (hash_path(h)) {
    case HP_FNV: return launch_kw<HP_FNV>(t, v, h, out, s, cus);
    case HP_XOR_HW: return launch_kw<HP_XOR_HW>(t, v, h, out, s, cus);
    default: return launch_kw<HP_GENERIC>(t, v, h, out, s, cus);
    }
}

}  // namespace vj
