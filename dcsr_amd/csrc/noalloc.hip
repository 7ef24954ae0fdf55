// noalloc.hip -- the allocation-free update entry point (dcsr_update_noalloc).
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The update's
// scratch (sort keys, group tables, records, row plans) comes from one fixed
// device array per GPU, `scratch`, sized for NOALLOC_MAX_TRIPLES triples;
// nothing is allocated or freed, the pool is never grown or compacted.
//
// The array's host-side name is not a device address: its address on the
// current device is taken once per device with hipGetSymbolAddress (the first
// call on that device) and only that pointer is handed to kernels.  Every
// call on one GPU shares that GPU's array and leaves its contents behind:
// calls that overlap on one GPU (two streams, two threads) race, so callers
// serialize them.  The only words that must start at zero (the update's
// finish counters) are cleared by a fill kernel at the start of each call.
#include <mutex>

#include "dcsr_internal.h"

namespace {

constexpr int64_t NOALLOC_MAX_TRIPLES = int64_t(1) << 16;
// covers every shape at NOALLOC_MAX_TRIPLES (about 6.9 MB are used: 100
// bytes per triple plus the row-group tables); checked per call
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[NOALLOC_MAX_DEVICES];

// The current device's copy of `scratch` (looked up once per device).
dcsr_status scratch_on_current_device(uint8_t **out)
{
    int dev = 0;
    DCSR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return DCSR_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        DCSR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return DCSR_ERR_HIP;
        g_scratch_ptr[dev] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[dev];
    return DCSR_OK;
}

}  // namespace

extern "C" dcsr_status dcsr_update_noalloc(dcsr_handle h, const int32_t *rows, const int32_t *cols,
                                           const float *vals, int64_t n, uint32_t flags, dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals))) return DCSR_ERR_INVALID;
    if (n > NOALLOC_MAX_TRIPLES) return DCSR_ERR_TOO_LARGE;
    if (update_scratch_bytes(h, n) > NOALLOC_SCRATCH_BYTES) return DCSR_ERR_TOO_LARGE;
    uint8_t *base = nullptr;
    DCSR_TRY(scratch_on_current_device(&base));
    return update_fixed_scratch(h, rows, cols, vals, n, flags, (hipStream_t)stream, base, NOALLOC_SCRATCH_BYTES);
}

extern "C" int64_t dcsr_update_noalloc_max_triples(void) { return NOALLOC_MAX_TRIPLES; }
