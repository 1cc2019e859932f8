// capi_internal.hpp — state and helpers shared by the C ABI's sources.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "compile.hpp"
#include "kernels.hpp"
#include "smc_capi.h"

namespace smc::capi {

extern thread_local std::string g_last_error;

inline int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const HipError &e) {
        return fail(SMC_E_DEVICE, e.what());
    } catch (const std::bad_alloc &) {
        return fail(SMC_E_ALLOC, "out of host memory");
    } catch (const std::exception &e) {
        std::string m = e.what();
        if (m.rfind("limit:", 0) == 0) return fail(SMC_E_LIMIT, m);
        return fail(SMC_E_INVALID, m);
    }
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Per-run device outputs of one instance, grown on demand.
struct RunBuffers {
    size_t cap = 0;
    uint8_t *flag = nullptr;
    double *value = nullptr;
    uint32_t *steps = nullptr, *trans = nullptr;
    uint64_t *digest = nullptr;
    double *time = nullptr;

    void release() {
        for (void *p : {(void *)flag, (void *)value, (void *)steps, (void *)trans, (void *)digest, (void *)time})
            if (p) hipFree(p);
        flag = nullptr; value = nullptr; steps = nullptr; trans = nullptr; digest = nullptr; time = nullptr;
        cap = 0;
    }
    void ensure(size_t n) {
        if (n <= cap) return;
        release();
        size_t c = std::max<size_t>(n, 1024);
        hip_check(hipMalloc(&flag, c), "hipMalloc flags");
        hip_check(hipMalloc(&value, c * 8), "hipMalloc values");
        hip_check(hipMalloc(&steps, c * 4), "hipMalloc steps");
        hip_check(hipMalloc(&trans, c * 4), "hipMalloc transitions");
        hip_check(hipMalloc(&digest, c * 8), "hipMalloc digests");
        hip_check(hipMalloc(&time, c * 8), "hipMalloc times");
        cap = c;
    }
};

}  // namespace smc::capi

struct smc_model {
    int device = 0;
    hipStream_t stream = nullptr;
    smc::ModelImage img;
    DevModel dm{};
    void *arena = nullptr;
    uint32_t table_bytes = 0;  // arena prefix a run kernel may stage in LDS (the hot tables)
    uint32_t small_off = 0;    // the hot tables after the location records: offset and bytes
    uint32_t small_bytes = 0;
    int cu_count = 0;
    smc::capi::RunBuffers runs;
    InstDesc h_inst{};
    InstDesc *d_inst = nullptr;
    unsigned long long *d_counter = nullptr;
    ReducePartial *d_partials = nullptr, *d_acc = nullptr;
    ReducePartial h_acc{};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    smc_launch tuned{};
    bool has_tuned = false;

    ~smc_model() {
        if (device >= 0) hipSetDevice(device);
        runs.release();
        for (void *p : {(void *)d_inst, (void *)d_counter, (void *)d_partials, (void *)d_acc, arena})
            if (p) hipFree(p);
        for (auto &e : ev) if (e) hipEventDestroy(e);
        if (stream) hipStreamDestroy(stream);
    }
};

