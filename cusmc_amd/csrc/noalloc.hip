// noalloc.hip — the fixed device scratch of smc_run_noalloc.
//
// smc_run sizes its per-run buffers to each call (hipMalloc on growth).
// Callers that cannot allocate (code under stream capture, serving loops,
// harnesses that forbid hipMalloc) use smc_run_noalloc instead, whose
// intermediate buffers live in this file-scope array: one copy per GPU, held
// until the process ends.  Host code must not pass `scratch` itself to a
// kernel or a copy (in host code the name is a host-side stand-in); it uses
// the device address hipGetSymbolAddress returns, looked up once per device
// with that device current, and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

namespace {

__device__ __align__(256) uint8_t scratch[SMC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

namespace smc {

hipError_t noalloc_scratch(int device, void **ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

}  // namespace smc
