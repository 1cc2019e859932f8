// kernels.hpp — launch interface of the run kernel and the reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "devmodel.h"

// Query, run range and outputs of one instance.
struct SimParams {
    uint64_t seed;
    uint64_t run_begin;         // run index of local run 0 in round 0
    uint64_t n_runs;            // runs per round
    double bound;
    double threshold;
    double qinit;               // initial max/min of the query variable
    int32_t reach;              // 1: location query
    int32_t qvar;               // query variable or -1
    int32_t is_max;
    uint32_t max_steps;         // runs are cut off after this many race rounds
    uint8_t *flag;              // [n_runs] required
    double *value;              // [n_runs] required
    uint32_t *steps;            // [n_runs] required
    uint32_t *trans;            // [n_runs] required
    uint64_t *digest;           // [n_runs] optional (trace variant)
    double *time;               // [n_runs] optional
};

// One (model, query, runs) instance of a launch.  Instances of one launch
// share the kernel's slots per lane C and may differ in lanes per run
// (S = next_pow2(components) / C, at least 1).
struct InstDesc {
    DevModel m;
    SimParams p;
    uint64_t chunk_first;       // first global chunk of this instance
    int32_t lanes;              // S, lanes per run (power of two <= 64)
    uint32_t chunk_runs;        // runs per chunk of this instance
};

struct LaunchParams {
    const InstDesc *inst;       // sorted by chunk_first
    int32_t n_inst;
    int32_t var_stride;         // doubles of LDS per segment (>= every n_vars)
    uint64_t total_chunks;
    unsigned long long *counter;  // chunks handed out (cleared by launch_prep)
    const uint64_t *round;      // optional: runs of round r start at run_begin + r * n_runs
    const SMC_GLOBAL uint8_t *tables;  // tables staged in LDS (one instance): first staged byte
    int32_t table_mode;         // 0 none, 1 every hot table, 2 all but the location records
    uint32_t table_bytes;       // bytes staged (multiple of 16), 0 without staging
    uint32_t wave_lds;          // LDS bytes of one wave after the tables (multiple of 16)
};

// Waves per workgroup of the run kernel, at most.
#define SMC_MAX_GROUP_WAVES 4

struct ReducePartial {
    double sum, sumsq, vmin, vmax;
    unsigned long long hits, steps, trans;
    unsigned long long truncated;
};

namespace smc {

// One build of the run kernel (sim_variant.hip).
struct SimVariant {
    hipError_t (*launch)(const LaunchParams &lp, bool trace, int tab, int waves, int group_waves, size_t lds,
                         hipStream_t s);
    hipError_t (*occupancy)(bool trace, int tab, int group_waves, size_t lds, int *blocks_per_cu);
};
SimVariant sim_variant_c1_o0();
SimVariant sim_variant_c2_o0();
SimVariant sim_variant_c4_o0();
SimVariant sim_variant_c8_o0();
SimVariant sim_variant_c16_o0();
SimVariant sim_variant_c1_o3();
SimVariant sim_variant_c2_o3();
SimVariant sim_variant_c4_o3();
SimVariant sim_variant_c8_o3();
SimVariant sim_variant_c16_o3();

// occ: 0 registers unrestricted, 3 at most 168 (three waves per SIMD); tab:
// table placement (lp.table_mode; 0 global memory); waves: total, a multiple of
// group_waves (waves per workgroup, <= SMC_MAX_GROUP_WAVES)
hipError_t launch_sim(const LaunchParams &lp, int slots, bool trace, int occ, int waves, int group_waves,
                      size_t lds_bytes, hipStream_t stream);
hipError_t sim_occupancy(int slots, bool trace, int occ, int tab, int group_waves, size_t lds_bytes,
                         int *blocks_per_cu);
hipError_t launch_reduce(const InstDesc *inst, int n_inst, uint32_t blocks_per_inst, ReducePartial *partials,
                         ReducePartial *acc, int accumulate, hipStream_t s);
hipError_t launch_prep(unsigned long long *counter, uint64_t *round, int advance, hipStream_t s);

}  // namespace smc
