// sim_variant.hip — one build of the run kernel (sim_body.h): SMC_C
// component slots per lane, register policy SMC_OCC (0 unrestricted, 2 at
// most 256 registers: two waves per SIMD).  The build compiles this file
// once per pair, in parallel.
#include "sim_body.h"

#if !defined(SMC_C) || !defined(SMC_OCC)
#error "SMC_C and SMC_OCC must be defined"
#endif

#define SMC_JOIN2(a, b, c, d) a##b##c##d
#define SMC_JOIN(a, b, c, d) SMC_JOIN2(a, b, c, d)

namespace {

template <bool TRACE>
const void *kernel_fn() {
#if SMC_OCC == 2
    return (const void *)sim_kernel_o2<SMC_C, TRACE>;
#else
    return (const void *)sim_kernel<SMC_C, TRACE>;
#endif
}

hipError_t launch(const LaunchParams &lp, bool trace, int waves, size_t lds, hipStream_t s) {
#if SMC_OCC == 2
    if (trace) hipLaunchKernelGGL((sim_kernel_o2<SMC_C, true>), dim3(waves), dim3(64), lds, s, lp);
    else hipLaunchKernelGGL((sim_kernel_o2<SMC_C, false>), dim3(waves), dim3(64), lds, s, lp);
#else
    if (trace) hipLaunchKernelGGL((sim_kernel<SMC_C, true>), dim3(waves), dim3(64), lds, s, lp);
    else hipLaunchKernelGGL((sim_kernel<SMC_C, false>), dim3(waves), dim3(64), lds, s, lp);
#endif
    return hipGetLastError();
}

hipError_t occupancy(bool trace, size_t lds, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, trace ? kernel_fn<true>() : kernel_fn<false>(),
                                                        64, lds);
}

}  // namespace

namespace smc {
SimVariant SMC_JOIN(sim_variant_c, SMC_C, _o, SMC_OCC)() { return SimVariant{&launch, &occupancy}; }
}  // namespace smc
