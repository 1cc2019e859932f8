// sim_variant.hip — one build of the run kernel (sim_body.h): SMC_C
// component slots per lane, register policy SMC_OCC (0 unrestricted, 3 at
// most 168 registers: three waves per SIMD): with digests (tables in global
// memory), and without, with the tables in global memory, staged in LDS, or
// staged in LDS but for the location records.
// The build compiles this file once per pair, in parallel.
#include "sim_body.h"

#if !defined(SMC_C) || !defined(SMC_OCC)
#error "SMC_C and SMC_OCC must be defined"
#endif

#define SMC_JOIN2(a, b, c, d) a##b##c##d
#define SMC_JOIN(a, b, c, d) SMC_JOIN2(a, b, c, d)

namespace {

#if SMC_OCC == 3
#define SMC_KERNEL sim_kernel_o3
#else
#define SMC_KERNEL sim_kernel
#endif

// Digest runs (parity checks) read the tables from global memory only.
const void *kernel_fn(bool trace, int tab) {
    if (trace) return (const void *)SMC_KERNEL<SMC_C, true, 0>;
    return tab == 1 ? (const void *)SMC_KERNEL<SMC_C, false, 1>
         : tab == 2 ? (const void *)SMC_KERNEL<SMC_C, false, 2> : (const void *)SMC_KERNEL<SMC_C, false, 0>;
}

hipError_t launch(const LaunchParams &lp, bool trace, int tab, int waves, int group_waves, size_t lds,
                  hipStream_t s) {
    const dim3 grid(waves / group_waves), block(64 * group_waves);
    if (trace) {
        if (tab) return hipErrorInvalidValue;
        hipLaunchKernelGGL((SMC_KERNEL<SMC_C, true, 0>), grid, block, lds, s, lp);
    } else {
        if (tab == 1) hipLaunchKernelGGL((SMC_KERNEL<SMC_C, false, 1>), grid, block, lds, s, lp);
        else if (tab == 2) hipLaunchKernelGGL((SMC_KERNEL<SMC_C, false, 2>), grid, block, lds, s, lp);
        else hipLaunchKernelGGL((SMC_KERNEL<SMC_C, false, 0>), grid, block, lds, s, lp);
    }
    return hipGetLastError();
}

hipError_t occupancy(bool trace, int tab, int group_waves, size_t lds, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kernel_fn(trace, tab), 64 * group_waves, lds);
}

}  // namespace

namespace smc {
SimVariant SMC_JOIN(sim_variant_c, SMC_C, _o, SMC_OCC)() { return SimVariant{&launch, &occupancy}; }
}  // namespace smc
