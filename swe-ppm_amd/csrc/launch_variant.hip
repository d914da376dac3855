// launch_variant.hip -- solver-independent kernels that depend on the math
// variant (the flavour-U apply pass and its dt rule).  Compiled once per
// variant with -DSWE_VARIANT=<0|1> and the matching flags.
#include <hip/hip_runtime.h>

#include "kernels_u.hpp"
#include "swe_internal.h"

#ifndef SWE_VARIANT
#define SWE_VARIANT 0
#endif

namespace swe {

template <int V>
int launch_apply_u_variant(StateView v, SumsView sums, BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed,
                           int fuseBc, float* dtLog, int logIdx, hipStream_t stream) {
    // one cell per thread: 7 loads and 3 stores in flight per lane, no
    // dependent loop, so the pass streams at memory speed
    constexpr int XS = 1;
    const int nxs = (v.nx + XS - 1) / XS;
    dim3 grid((v.ny + 63) / 64, (nxs + 3) / 4);
    hipLaunchKernelGGL((apply_u_kernel<XS, V>), grid, dim3(256), 0, stream, v, sums, bi, ctl, sl, dtFixed, fuseBc,
                       dtLog, logIdx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "apply_u_kernel launch");
    return 0;
}

template int launch_apply_u_variant<SWE_VARIANT>(StateView, SumsView, BoundaryInfo, DeviceCtl*, Slots, float, int,
                                                 float*, int, hipStream_t);

}  // namespace swe
