// noalloc.hpp — layout of the fixed device scratch of smc_run_noalloc.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kernels.hpp"

// Runs per smc_run_noalloc call, at most.
#define SMC_NOALLOC_MAX_RUNS (1u << 20)

namespace smc {

// Byte offsets into the scratch: the instance record, the work counter, the
// reduction's totals and per-block partials, then the per-run outputs the
// reduction reads (values f64, steps u32, transitions u32, flags u8).
constexpr size_t kNoallocInst = 0;
constexpr size_t kNoallocCounter = 1024;
constexpr size_t kNoallocAcc = 1280;
constexpr size_t kNoallocPartials = 1536;
constexpr size_t kNoallocPartialsMax = 1024;
constexpr size_t kNoallocRuns = kNoallocPartials + kNoallocPartialsMax * sizeof(ReducePartial);
constexpr size_t kNoallocValue = kNoallocRuns;
constexpr size_t kNoallocSteps = kNoallocValue + 8ull * SMC_NOALLOC_MAX_RUNS;
constexpr size_t kNoallocTrans = kNoallocSteps + 4ull * SMC_NOALLOC_MAX_RUNS;
constexpr size_t kNoallocFlag = kNoallocTrans + 4ull * SMC_NOALLOC_MAX_RUNS;
static_assert(sizeof(InstDesc) <= kNoallocCounter - kNoallocInst, "instance record");
static_assert(sizeof(ReducePartial) <= kNoallocPartials - kNoallocAcc, "totals");
static_assert(kNoallocRuns % 256 == 0, "alignment");

// Device address of this device's scratch (looked up once per device).
hipError_t noalloc_scratch(int device, void **ptr);

}  // namespace smc

#define SMC_NOALLOC_SCRATCH_BYTES (smc::kNoallocFlag + 1ull * SMC_NOALLOC_MAX_RUNS)
