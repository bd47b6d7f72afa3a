// Device helpers shared by the join pipeline's kernels: wave-uniform values
// moved into scalar registers, global-address-space views of the tables,
// wave and block scans, and searches over prefix arrays.
#pragma once
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

// Exclusive scan over the block; `sh` holds >= blockDim/64 values. Every wave
// adds up the wave totals below it itself (no serial step); NW: waves per
// block when known at compile time.
template <typename T, int NW = 0> __device__ T block_excl_scan(T v, T *sh, T *total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t nw = NW ? (uint32_t)NW : (blockDim.x + 63) >> 6;
    const T inc = wave_incl_scan(v);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T carry = 0, tot = 0;
    for (uint32_t k = 0; k < nw; k++) {
        const T x = sh[k];
        carry += k < w ? x : T(0);
        tot += x;
    }
    __syncthreads();  // the caller may reuse sh at once
    *total = tot;
    return carry + inc - v;
}

}  // namespace

}  // namespace vj
