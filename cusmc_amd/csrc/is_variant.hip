// is_variant.hip — one build of the tilted advance kernel (is_advance.h):
// SMC_C component slots per lane (registers unrestricted, SMC_OCC 0), with
// the tables in global memory or staged in LDS.  The build compiles this
// file once per slot count.
#include "is_advance.h"

#if !defined(SMC_C) || !defined(SMC_OCC) || SMC_OCC != 0
#error "SMC_C must be defined and SMC_OCC be 0"
#endif

#define SMC_JOIN2(a, b, c, d) a##b##c##d
#define SMC_JOIN(a, b, c, d) SMC_JOIN2(a, b, c, d)

namespace {

hipError_t launch(const IsParams &p, int grid, int group_waves, size_t lds, hipStream_t s) {
    const dim3 g(grid), b(64 * group_waves);
    // builds: staged tables (sampling), global tables (sampling), global
    // tables with recording
    if (p.rec_pos && p.table_bytes) return hipErrorInvalidValue;
    if (p.rec_pos) hipLaunchKernelGGL((is_advance_kernel<SMC_C, false, true>), g, b, lds, s, p);
    else if (p.table_bytes) hipLaunchKernelGGL((is_advance_kernel<SMC_C, true, false>), g, b, lds, s, p);
    else hipLaunchKernelGGL((is_advance_kernel<SMC_C, false, false>), g, b, lds, s, p);
    return hipGetLastError();
}

}  // namespace

namespace smc {
IsAdvanceLaunch SMC_JOIN(is_advance_c, SMC_C, _o, SMC_OCC)() { return &launch; }
}  // namespace smc
