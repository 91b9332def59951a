// csx_noalloc.hip -- csx_compact_noalloc: csx_compact_config without a
// workspace argument, for callers that cannot allocate device memory (code
// under graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The scratch is a file-scope device array.  In host code `scratch` names a
// host-side stand-in, not a device address: its address on each GPU comes
// from hipGetSymbolAddress, called once per device with that device current
// on the first call there, and only that pointer is ever handed to kernels.
// Every pipeline initialises the parts of its workspace that must start at
// zero (look-back ticket, error word and status words) with a kernel of its
// own on every call, so the contents left by an earlier call do not matter.
// Calls that overlap on one GPU share its array and race: the caller
// serialises them.
#include <mutex>

#include "csx_internal.hpp"

namespace {

static_assert(CSX_NOALLOC_SCRATCH_BYTES % 256 == 0, "whole 256-byte blocks");
static __device__ __align__(256) uint8_t scratch[CSX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // per device: the array's device address, null until set up

// The current device's scratch address (set up on first use on that device).
csx_status device_scratch(void** p)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return csx::set_error(CSX_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return csx::set_error(CSX_ERR_UNSUPPORTED, "device %d is beyond the %d this build tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void* q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return csx::set_error(CSX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        if (!q) return csx::set_error(CSX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d returned no address", dev);
        g_scratch[dev] = q;
    }
    *p = g_scratch[dev];
    return CSX_OK;
}

} // namespace

extern "C" csx_status csx_compact_noalloc(const void* d_in, const uint8_t* d_mask, void* d_out, uint64_t n, uint32_t elem_size, int bit_order,
                                          const csx_config* cfg, uint64_t* d_count, uint64_t* h_count, hipStream_t stream)
{
    if (!cfg) return csx::set_error(CSX_ERR_INVALID, "null configuration");
    // the size check first: an oversized request fails before any device work
    size_t need = 0;
    csx_status s = csx_workspace_size(n, elem_size, cfg->chunk_len, &need);
    if (s != CSX_OK) return s;
    if (need > CSX_NOALLOC_SCRATCH_BYTES)
        return csx::set_error(CSX_ERR_UNSUPPORTED, "%llu elements of %u bytes at chunk %u need %zu bytes of scratch; the static scratch holds %llu",
                              (unsigned long long)n, elem_size, cfg->chunk_len, need, (unsigned long long)CSX_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    if ((s = device_scratch(&ws)) != CSX_OK) return s;
    return csx::compact_in_workspace(d_in, d_mask, d_out, n, elem_size, bit_order, cfg, d_count, h_count, ws, CSX_NOALLOC_SCRATCH_BYTES, stream);
}
