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

// A block's copy layer with its pending update applied on the fly, as
// apply_u_kernel would apply it (same dt / dx, dt / dy and update_cell),
// into another block's ghost line.  src, sums and dst point at the lines'
// first cells; one thread per cell.
template <int V>
__global__ __launch_bounds__(256) void copy_line_pending_kernel(LinePending L) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const float dtdx = L.dt / L.dx;
    const float dtdy = L.dt / L.dy;
    const size_t c = (size_t)i * L.sstride;
    const Cell q = update_cell(L.sh[c], L.shu[c], L.shv[c], dtdx, dtdy, L.sums.sx[c], L.sums.sy[c], L.sums.su[c],
                               L.sums.sv[c]);
    const size_t d = (size_t)i * L.dstride;
    L.dh[d] = q.h;
    L.dhu[d] = q.hu;
    L.dhv[d] = q.hv;
}

template <int V>
int launch_copy_line_pending_variant(const LinePending& L, hipStream_t stream) {
    if (L.n <= 0) return 0;
    hipLaunchKernelGGL((copy_line_pending_kernel<V>), dim3((L.n + 255) / 256), dim3(256), 0, stream, L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "copy_line_pending_kernel launch");
    return 0;
}

template int launch_copy_line_pending_variant<SWE_VARIANT>(const LinePending&, hipStream_t);

}  // namespace swe
