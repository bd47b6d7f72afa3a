// Fixed device scratch of the no-allocation join entry point
// (vj_join_into_noalloc, vj_context.cpp).
//
// The pipeline's scratch (partitioned tuple copies, histogram counts,
// partition bases, task prefixes, descriptors) normally lives in arenas a
// context grows with hipMalloc. Callers that cannot allocate use this
// file-scope array instead: every GPU that loads the library holds one copy
// of kNoAllocScratchBytes until the process ends. Host code never passes the
// array's name to a kernel or a copy (on the host it is only a stand-in for
// the symbol): the device address of this device's copy comes from
// hipGetSymbolAddress, looked up once per device with that device current
// (at context creation) and cached here.
//
// Every call on one GPU shares that GPU's copy and leaves its contents
// behind; the pipeline writes every scratch word it reads (counts, bases,
// task prefixes, queue ticket and output counters are all written by earlier
// stages of the same launch sequence), so nothing is cleared between calls.
// Calls that overlap on one GPU race on it.
#include <mutex>
#include <string>

#include "vj_internal.hpp"

namespace vj {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_mu;
void *g_ptr[kNoAllocMaxDevices] = {};

}  // namespace

Err noalloc_scratch(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= kNoAllocMaxDevices)
        return Err::make(VJ_ERR_UNSUPPORTED, "no-allocation scratch supports devices 0.." +
                                                 std::to_string(kNoAllocMaxDevices - 1));
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        VJ_HIP_TRY(hipGetDevice(&cur));
        if (cur != device) return Err::make(VJ_ERR_INTERNAL, "scratch lookup needs its device current");
        void *p = nullptr;
        VJ_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return Err::make(VJ_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return Err::ok();
}

}  // namespace vj
