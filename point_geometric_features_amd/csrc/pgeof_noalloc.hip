// Static device scratch of the no-allocation entry point
// (pgeof_compute_features_device_noalloc, include/pgeof_hip.h).
//
// The array below is a file-scope device symbol: each device the library runs
// on holds its own copy, at its own address, from the first use until the
// process ends.  In host code its name is only a host-side stand-in; the
// device address is looked up with hipGetSymbolAddress once per device (with
// that device current) and that pointer is what the launches receive.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pgeof_hip.h"
#include "pgeof_internal.hpp"

namespace pgeof_amd
{

namespace
{

// 60 MiB: the spatial row order of PGEOF_NOALLOC_MAX_ROWS rows
// (12 bytes per row + 2^10 bucket counts per 4096 rows + scan partials:
// 52.0 MiB)
constexpr size_t kScratchBytes = size_t(60) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void** ptr, size_t* bytes)
{
    static_assert(kScratchBytes % 256 == 0, "scratch size");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (csr_order_bytes(PGEOF_NOALLOC_MAX_ROWS) > kScratchBytes) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> l(g_mu);
    if (!g_ptr[dev])
    {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *ptr = g_ptr[dev];
    *bytes = kScratchBytes;
    return hipSuccess;
}

}  // namespace pgeof_amd
