// colo_noalloc.hip -- colo_scan_slots_noalloc: the slot scan of
// colo_scan_slots for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The scan's only workspace -- the per-tile sums of its
// three-phase scan -- lives in a file-scope device array instead of the
// context's growable scratch.  Each GPU holds its own copy of the array
// for the life of the process; calls on one GPU share it, so calls that
// overlap on one GPU race.  The tile sums are written before they are read
// in every call, so nothing needs clearing between calls.
#include <mutex>

#include "colo_internal.h"

namespace {
constexpr uint64_t kScratchWords = 132096;  // >= scan_ws_words(COLO_SCAN_NOALLOC_MAX_N) = 131140
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint64_t scratch[kScratchWords];

namespace {
std::mutex g_mu;
uint64_t *g_scratch[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device
// with the device current.  Host code uses only this address: the array's
// name in host code is a host-side stand-in.
int scratch_address(colo_ctx *c, uint64_t **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return colo::set_error(c, COLO_EHIP, "colo_scan_slots_noalloc: no current device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return colo::set_error(c, COLO_EHIP, std::string("colo_scan_slots_noalloc: hipGetSymbolAddress: ") +
                                               hipGetErrorString(e));
        g_scratch[dev] = (uint64_t *)p;
    }
    *out = g_scratch[dev];
    return COLO_OK;
}
}  // namespace

extern "C" int colo_scan_slots_noalloc(colo_ctx *c, const uint32_t *d_slots, uint64_t n, uint64_t *d_offsets,
                                       uint64_t *total, void *stream) {
    if (!c) return COLO_EARG;
    if (!d_offsets || (n && !d_slots)) return colo::set_error(c, COLO_EARG, "colo_scan_slots_noalloc: null buffer");
    if (n > COLO_SCAN_NOALLOC_MAX_N || colo::scan_ws_words(n) > kScratchWords)
        return colo::set_error(c, COLO_EARG, "colo_scan_slots_noalloc: n = " + std::to_string(n) +
                                           " slots exceeds its static workspace (at most 2^28 slots); "
                                           "use colo_scan_slots");
    colo_ctx *ctx_for_err = c;
    uint64_t *ws = nullptr;
    int rc = scratch_address(c, &ws);
    if (rc) return rc;
    hipStream_t s = colo::pick_stream(c, stream);
    if ((rc = colo::scan_slots_ws(d_slots, n, d_offsets, ws, s))) return colo::set_error(c, rc, "colo_scan_slots_noalloc: launch");
    if (total) {
        COLO_HIP(hipMemcpyAsync(c->h_pinned, d_offsets + n, 8, hipMemcpyDeviceToHost, s));
        COLO_HIP(hipStreamSynchronize(s));
        *total = c->h_pinned[0];
    }
    return COLO_OK;
}
