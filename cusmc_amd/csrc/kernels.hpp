// kernels.hpp — launch interface of the run kernel and the reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "devmodel.h"

struct SimParams {
    uint64_t seed;
    uint64_t run_begin;
    uint64_t n_runs;
    double bound;
    double threshold;
    double qinit;               // initial max/min of the query variable
    int32_t reach;              // 1: location query
    int32_t qvar;               // query variable or -1
    int32_t is_max;
    int32_t lanes;              // S, lanes per run (power of two <= 64)
    int32_t var_stride;         // doubles of LDS per segment
    uint32_t max_steps;         // runs are cut off after this many race rounds
    uint8_t *flag;              // [n_runs] required
    double *value;              // [n_runs] required
    uint32_t *steps;            // [n_runs] required
    uint32_t *trans;            // [n_runs] required
    uint64_t *digest;           // [n_runs] optional (trace variant)
    double *time;               // [n_runs] optional
    unsigned long long *counter;  // runs handed out so far (zeroed by the caller)
};

struct ReducePartial {
    double sum, sumsq, vmin, vmax;
    unsigned long long hits, steps, trans;
    unsigned long long truncated;
};

namespace smc {

hipError_t launch_sim(const DevModel &m, const SimParams &p, int slots, bool trace, int waves,
                      size_t lds_bytes, hipStream_t stream);
hipError_t sim_occupancy(int slots, bool trace, size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_reduce(const uint8_t *flag, const double *value, const uint32_t *steps,
                         const uint32_t *trans, uint64_t n, uint32_t max_steps, int blocks,
                         ReducePartial *out, hipStream_t s);

}  // namespace smc
