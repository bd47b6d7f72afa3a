// Kernels of the reference-layout join (vj_join_rows, vj_rows.cpp).
//
// The reference's db_table (src/base/types.hpp:128-266) keeps every key
// column as a 64-bit column_t, row-major (column_values[row * cols + c]), and
// a 64-bit primary key per row. Its join result (merge_joined_tables,
// src/join/join_provider.cuh:948-996, filled by copy_probe_results_kernel,
// src/join/join_provider.cu:593-608, and generate_primary_key_kernel, :307-313)
// is a db_table of two columns: (r.primary_keys[r], s.primary_keys[s]) per
// matching pair, with primary keys 1..M.
//
//   k_rows_to_cols  row-major n x C u64 keys -> C SoA columns (C = 2..4), so the
//                   pipeline reads each key column contiguously
//   k_materialize   (build row, probe row) index pairs -> the reference's rs
//                   rows (one 16-byte store per pair) and primary keys 1..M
// Both are HBM-bound: the transpose moves 2 x 8C bytes per row; the
// materialization reads 8 bytes of indexes and two 8-byte primary keys and
// writes 24 bytes per pair.
#include <algorithm>

#include "vj_internal.hpp"

namespace vj {

namespace {

constexpr int kRowsThreads = 256;
constexpr int kRowsUnroll = 4;

template <int C>
__global__ __launch_bounds__(kRowsThreads) void k_rows_to_cols(const uint64_t *__restrict__ rows_, uint64_t n,
                                                               uint64_t *__restrict__ cols_) {
    const gptr<const uint64_t> rows = to_global(rows_);
    const gptr<uint64_t> cols = to_global(cols_);
    const uint64_t stride = (uint64_t)gridDim.x * kRowsThreads * kRowsUnroll;
    for (uint64_t b = (uint64_t)blockIdx.x * kRowsThreads * kRowsUnroll + threadIdx.x; b < n; b += stride) {
        uint64_t v[kRowsUnroll][C];
#pragma unroll
        for (int u = 0; u < kRowsUnroll; u++) {
            const uint64_t i = b + (uint64_t)u * kRowsThreads;
            const uint64_t r = i < n ? i : 0;  // past the end: row 0 again (not stored)
#pragma unroll
            for (int c = 0; c < C; c++) v[u][c] = rows[r * C + c];
        }
#pragma unroll
        for (int u = 0; u < kRowsUnroll; u++) {
            const uint64_t i = b + (uint64_t)u * kRowsThreads;
            if (i < n) {
#pragma unroll
                for (int c = 0; c < C; c++) cols[(uint64_t)c * n + i] = v[u][c];
            }
        }
    }
}

__global__ __launch_bounds__(kRowsThreads) void k_materialize(const uint32_t *__restrict__ ri_, const uint32_t *__restrict__ si_,
                                                              uint64_t m, const uint64_t *__restrict__ r_pk_,
                                                              const uint64_t *__restrict__ s_pk_, uint64_t *__restrict__ rs_cols_,
                                                              uint64_t *__restrict__ rs_pk_) {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    const gptr<const uint32_t> ri = to_global(ri_), si = to_global(si_);
    const gptr<const uint64_t> r_pk = to_global(r_pk_), s_pk = to_global(s_pk_);
    const gptr<u64x2> rs_cols = to_global(reinterpret_cast<u64x2 *>(rs_cols_));
    const gptr<uint64_t> rs_pk = to_global(rs_pk_);
    const uint64_t stride = (uint64_t)gridDim.x * kRowsThreads * kRowsUnroll;
    for (uint64_t b = (uint64_t)blockIdx.x * kRowsThreads * kRowsUnroll + threadIdx.x; b < m; b += stride) {
        uint32_t r[kRowsUnroll], s[kRowsUnroll];
#pragma unroll
        for (int u = 0; u < kRowsUnroll; u++) {
            const uint64_t i = b + (uint64_t)u * kRowsThreads;
            r[u] = ri[i < m ? i : 0];
            s[u] = si[i < m ? i : 0];
        }
        uint64_t pr[kRowsUnroll], ps[kRowsUnroll];
#pragma unroll
        for (int u = 0; u < kRowsUnroll; u++) {
            pr[u] = r_pk[r[u]];
            ps[u] = s_pk[s[u]];
        }
#pragma unroll
        for (int u = 0; u < kRowsUnroll; u++) {
            const uint64_t i = b + (uint64_t)u * kRowsThreads;
            if (i < m) {
                rs_cols[i] = u64x2{pr[u], ps[u]};
                rs_pk[i] = i + 1;
            }
        }
    }
}

unsigned grid_for(uint64_t n, int cus) {
    const uint64_t per = (uint64_t)kRowsThreads * kRowsUnroll;
    const uint64_t want = (n + per - 1) / per, cap = (uint64_t)cus * 8;
    return (unsigned)std::max<uint64_t>(1, std::min(want, cap));
}

}  // namespace

Err launch_rows_to_cols(const uint64_t *rows, uint64_t n, int ncols, uint64_t *cols, hipStream_t s, int cus) {
    if (n == 0) return Err::ok();
    const unsigned g = grid_for(n, cus);
    switch (ncols) {
    case 2: hipLaunchKernelGGL(k_rows_to_cols<2>, dim3(g), dim3(kRowsThreads), 0, s, rows, n, cols); break;
    case 3: hipLaunchKernelGGL(k_rows_to_cols<3>, dim3(g), dim3(kRowsThreads), 0, s, rows, n, cols); break;
    case 4: hipLaunchKernelGGL(k_rows_to_cols<4>, dim3(g), dim3(kRowsThreads), 0, s, rows, n, cols); break;
    default: return Err::make(VJ_ERR_INVALID_ARGUMENT, "row-major transpose takes 2..4 columns");
    }
    VJ_HIP_TRY(hipGetLastError());
    return Err::ok();
}

Err launch_materialize(const uint32_t *ri, const uint32_t *si, uint64_t m, const uint64_t *r_pk, const uint64_t *s_pk,
                       uint64_t *rs_cols, uint64_t *rs_pk, hipStream_t s, int cus) {
    if (m == 0) return Err::ok();
    hipLaunchKernelGGL(k_materialize, dim3(grid_for(m, cus)), dim3(kRowsThreads), 0, s, ri, si, m, r_pk, s_pk, rs_cols,
                       rs_pk);
    VJ_HIP_TRY(hipGetLastError());
    return Err::ok();
}

}  // namespace vj
