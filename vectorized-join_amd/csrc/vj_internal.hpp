// Host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "vj.h"
#include "vj_device.hpp"

namespace vj {

// Status with a message, converted to vj_status + vj_last_error at the ABI.
struct Err {
    vj_status code = VJ_OK;
    std::string msg;
    static Err ok() { return Err(); }
    static Err make(vj_status c, std::string m) {
        Err e;
        e.code = c;
        e.msg = std::move(m);
        return e;
    }
    bool failed() const { return code != VJ_OK; }
};

#define VJ_HIP_TRY(expr)                                                                              \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return ::vj::Err::make(VJ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define VJ_TRY(expr)                       \
    do {                                   \
        ::vj::Err e_ = (expr);             \
        if (e_.failed()) return e_;        \
    } while (0)

// Validation shared by hashing and joining.
Err check_table(const vj_table *t, bool need_rowids_absent_ok = true);
Err make_hash_cfg(const vj_hash_opts *o, const vj_table *t, HashCfg *cfg);
TView to_view(const vj_table *t);

// vj_hash.hip
Err launch_hash(const vj_table *t, const HashCfg &cfg, void *out, hipStream_t s, int num_cus);

// vj_rows.hip: the reference's row-major table layout
Err launch_rows_to_cols(const uint64_t *rows, uint64_t n, int ncols, uint64_t *cols, hipStream_t s, int cus);
Err launch_materialize(const uint32_t *ri, const uint32_t *si, uint64_t m, const uint64_t *r_pk, const uint64_t *s_pk,
                       uint64_t *rs_cols, uint64_t *rs_pk, hipStream_t s, int cus);

int device_cu_count(int device);

// vj_noalloc.hip: the fixed device scratch of vj_join_into_noalloc.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // per device, held until the process ends
static_assert(kNoAllocScratchBytes == VJ_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");
constexpr int kNoAllocMaxDevices = 64;
// Device address of `device`'s copy of the scratch array; the first call for a
// device (made with that device current) looks it up, later calls return it.
Err noalloc_scratch(int device, void **ptr);

}  // namespace vj
