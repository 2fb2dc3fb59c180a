// etraj_noalloc.hip -- the device scratch of etraj_integrate_and_bin_noalloc.
//
// The entry point's device buffers (the caller's columns, segment tables, the
// field table, re-integration flags and the per-electron binning records) are
// slices of one file-scope device array instead of hipMalloc'ed memory, so a
// caller that cannot allocate (a tight serving loop, a harness that forbids
// device allocation) can call it. Every GPU the library runs on holds its own
// copy of the array until the process ends. The array's name is a host-side
// stand-in, not a device address: host code uses only the address that
// hipGetSymbolAddress returns for the current device, kept per device here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "etraj_engine.hpp"

namespace etraj {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

// The device address of `device`'s copy of the scratch array. The lookup runs
// once per device, with that device current (the caller has set it).
hipError_t noalloc_scratch(int device, void** out) {
    if (!out) return hipErrorInvalidValue;
    *out = nullptr;
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace etraj
