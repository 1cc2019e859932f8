// sim_kernel.hip — the result reduction and the run kernel's launchers.
//
// The run kernel itself (sim_body.h) is compiled once per slot count and
// register policy (sim_variant.hip); this file dispatches to those builds.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "device_common.h"
#include "kernels.hpp"

namespace {

// ------------------------------------------------------------ reduction

// Per-block partial sums over a contiguous slice of one instance's runs
// (grid.y = instance), fixed tree order.
__global__ __launch_bounds__(256) void reduce_kernel(const InstDesc *inst, uint32_t blocks_per_inst,
                                                     ReducePartial *partials) {
    __shared__ double s_sum[256], s_sq[256], s_min[256], s_max[256];
    __shared__ unsigned long long s_hits[256], s_steps[256], s_trans[256], s_trunc[256];
    const SimParams p = inst[blockIdx.y].p;
    const uint64_t n = p.n_runs;
    const uint64_t per_block = (n + blocks_per_inst - 1) / blocks_per_inst;
    const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
    const uint64_t b1 = b0 + per_block < n ? b0 + per_block : n;
    double sum = 0.0, sq = 0.0, mn = __builtin_huge_val(), mx = -__builtin_huge_val();
    unsigned long long hits = 0, st = 0, tr = 0, tc = 0;
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += 256) {
        const double v = p.value[i];
        sum += v;
        sq = fma(v, v, sq);
        mn = fmin(mn, v);
        mx = fmax(mx, v);
        hits += p.flag[i];
        st += p.steps[i];
        tr += p.trans[i];
        tc += p.steps[i] >= p.max_steps;
    }
    s_sum[threadIdx.x] = sum; s_sq[threadIdx.x] = sq; s_min[threadIdx.x] = mn; s_max[threadIdx.x] = mx;
    s_hits[threadIdx.x] = hits; s_steps[threadIdx.x] = st; s_trans[threadIdx.x] = tr;
    s_trunc[threadIdx.x] = tc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            s_sq[threadIdx.x] += s_sq[threadIdx.x + o];
            s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + o]);
            s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + o]);
            s_hits[threadIdx.x] += s_hits[threadIdx.x + o];
            s_steps[threadIdx.x] += s_steps[threadIdx.x + o];
            s_trans[threadIdx.x] += s_trans[threadIdx.x + o];
            s_trunc[threadIdx.x] += s_trunc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ReducePartial r;
        r.sum = s_sum[0]; r.sumsq = s_sq[0]; r.vmin = s_min[0]; r.vmax = s_max[0];
        r.hits = s_hits[0]; r.steps = s_steps[0]; r.trans = s_trans[0]; r.truncated = s_trunc[0];
        partials[(size_t)blockIdx.y * blocks_per_inst + blockIdx.x] = r;
    }
}

// One thread per instance: fold its partials in order into the running
// totals (a fixed order, so totals do not depend on timing).
__global__ void finalize_kernel(const ReducePartial *partials, int n_inst, uint32_t blocks_per_inst,
                                ReducePartial *acc, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    ReducePartial t;
    if (accumulate) t = acc[i];
    else { t.sum = 0; t.sumsq = 0; t.vmin = __builtin_huge_val(); t.vmax = -__builtin_huge_val();
           t.hits = 0; t.steps = 0; t.trans = 0; t.truncated = 0; }
    for (uint32_t b = 0; b < blocks_per_inst; ++b) {
        const ReducePartial r = partials[(size_t)i * blocks_per_inst + b];
        t.sum += r.sum; t.sumsq += r.sumsq;
        t.vmin = fmin(t.vmin, r.vmin); t.vmax = fmax(t.vmax, r.vmax);
        t.hits += r.hits; t.steps += r.steps; t.trans += r.trans; t.truncated += r.truncated;
    }
    acc[i] = t;
}

// Start of a launch or of a captured round: clear the work counter and
// optionally advance the round (a kernel, not a memset, so it can be
// recorded in a graph and replayed).
__global__ void prep_kernel(unsigned long long *counter, uint64_t *round, int advance) {
    if (threadIdx.x == 0) {
        *counter = 0;
        if (round && advance) *round += 1;
    }
}

}  // namespace

// ------------------------------------------------------------ launchers

namespace smc {

static SimVariant variant(int slots, int occ) {
    if (occ != 0 && occ != 3) return SimVariant{};
    const bool o3 = occ == 3;
    switch (slots) {
    case 1: return o3 ? sim_variant_c1_o3() : sim_variant_c1_o0();
    case 2: return o3 ? sim_variant_c2_o3() : sim_variant_c2_o0();
    case 4: return o3 ? sim_variant_c4_o3() : sim_variant_c4_o0();
    case 8: return o3 ? sim_variant_c8_o3() : sim_variant_c8_o0();
    case 16: return o3 ? sim_variant_c16_o3() : sim_variant_c16_o0();
    default: return SimVariant{};
    }
}

hipError_t launch_sim(const LaunchParams &lp, int slots, bool trace, int occ, int waves, int group_waves,
                      size_t lds_bytes, hipStream_t stream) {
    const SimVariant v = variant(slots, occ);
    if (!v.launch || group_waves < 1 || group_waves > SMC_MAX_GROUP_WAVES || waves % group_waves) return hipErrorInvalidValue;
    return v.launch(lp, trace, lp.table_bytes > 0 ? lp.table_mode : 0, waves, group_waves, lds_bytes, stream);
}

hipError_t sim_occupancy(int slots, bool trace, int occ, int tab, int group_waves, size_t lds_bytes,
                         int *blocks_per_cu) {
    const SimVariant v = variant(slots, occ);
    if (!v.occupancy) return hipErrorInvalidValue;
    return v.occupancy(trace, tab, group_waves, lds_bytes, blocks_per_cu);
}

hipError_t launch_reduce(const InstDesc *inst, int n_inst, uint32_t blocks_per_inst, ReducePartial *partials,
                         ReducePartial *acc, int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3(blocks_per_inst, n_inst), dim3(256), 0, s, inst, blocks_per_inst, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(finalize_kernel, dim3((n_inst + 63) / 64), dim3(64), 0, s, partials, n_inst, blocks_per_inst,
                       acc, accumulate);
    return hipGetLastError();
}

hipError_t launch_prep(unsigned long long *counter, uint64_t *round, int advance, hipStream_t s) {
    hipLaunchKernelGGL(prep_kernel, dim3(1), dim3(64), 0, s, counter, round, advance);
    return hipGetLastError();
}

}  // namespace smc
