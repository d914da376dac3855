// launch_solver.hip -- per-solver, per-math-variant kernel instantiations.
// Compiled once per (solver, variant) with -DSWE_SOLVER_ID=<id> and
// -DSWE_VARIANT=<0|1>, so that the big device solvers build in parallel
// translation units and the two arithmetic variants get their own compile
// flags:
//   variant 0 (IEEE): no contraction, correctly rounded division and square
//                     root -- bit-compatible with an IEEE CPU evaluation;
//   variant 1 (fast): FMA contraction, hardware reciprocal / square root.
// The variant is also a template parameter of every kernel, so the two
// instantiations never share a symbol.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels_fused.hpp"
#include "kernels_s.hpp"
#include "kernels_u.hpp"
#include "swe_internal.h"

#ifndef SWE_SOLVER_ID
#error "SWE_SOLVER_ID must be defined"
#endif
#ifndef SWE_VARIANT
#define SWE_VARIANT 0
#endif

namespace swe {

namespace {

template <int S, int MODE, int XS, int V>
int go_u(StateView v, SumsView sums, float* hn, float* hun, float* hvn, BoundaryInfo bi, DeviceCtl* ctl, Slots sl,
         float dtFixed, int fuseBc, float* dtLog, int logIdx, hipStream_t stream) {
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    const int nxs = (v.nx + XS - 1) / XS;
    const int waves = nyt * nxs;
    const int groups = (waves + kWavesPerGroup - 1) / kWavesPerGroup;
    hipLaunchKernelGGL((sweep_u_kernel<S, MODE, XS, V>), dim3(groups), dim3(64 * kWavesPerGroup), 0, stream, v,
                       sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sweep_u_kernel launch");
    return 0;
}

template <int S, int MODE, int V>
int go_u_xs(int xseg, StateView v, SumsView sums, float* hn, float* hun, float* hvn, BoundaryInfo bi,
            DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog, int logIdx, hipStream_t stream) {
    switch (xseg) {
        case 8:
        case 10:  // width 10 runs the sweeps at 8
            return go_u<S, MODE, 8, V>(v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
        case 12:
            return go_u<S, MODE, 12, V>(v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
        case 32:
            return go_u<S, MODE, 32, V>(v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
        default:
            return go_u<S, MODE, 16, V>(v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
    }
}

template <int S, int V>
__global__ __launch_bounds__(256) void solve_edges_kernel(int64_t n, const float* __restrict__ in,
                                                          float* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const dev::EdgeOut e =
        dev::solve_raw<S>(in[k], in[n + k], in[2 * n + k], in[3 * n + k], in[4 * n + k], in[5 * n + k]);
    out[k] = e.hL;
    out[n + k] = e.hR;
    out[2 * n + k] = e.huL;
    out[3 * n + k] = e.huR;
    out[4 * n + k] = e.speed;
}

}  // namespace

template <int S, int V>
int launch_sweep_u_solver(int mode, int xseg, StateView v, SumsView sums, float* hn, float* hun, float* hvn,
                          BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog,
                          int logIdx, hipStream_t stream) {
    switch (mode) {
        case MODE_MAX:
            return go_u_xs<S, MODE_MAX, V>(xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx,
                                           stream);
        case MODE_STORE:
            return go_u_xs<S, MODE_STORE, V>(xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog,
                                             logIdx, stream);
        default:
            return go_u_xs<S, MODE_UPDATE, V>(xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog,
                                              logIdx, stream);
    }
}

template <int S, int V>
int launch_fused_u_solver(int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn, int slotOut,
                          int slotClear, float* dtLog, int logIdx, hipStream_t stream) {
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    auto go = [&](auto xsTag) {
        constexpr int XS = decltype(xsTag)::value;
        const int nxs = (v.nx + XS - 1) / XS;
        // list mode may list up to one piece per record tile
        const int waves = io.list ? nyt * ((v.nx + kQuietCols - 1) / kQuietCols) : nyt * nxs;
        const int groups = (waves + kWavesPerGroup - 1) / kWavesPerGroup;
        if (g_fusedTimer.start) {
            hipExtLaunchKernelGGL((fused_u_kernel<S, XS, V>), dim3(groups), dim3(64 * kWavesPerGroup), 0, stream,
                                  g_fusedTimer.start, g_fusedTimer.stop, 0, v, io, bi, ctl, slotIn, slotOut, slotClear,
                                  dtLog, logIdx);
            g_fusedTimer = LaunchTimer{};
        } else {
            hipLaunchKernelGGL((fused_u_kernel<S, XS, V>), dim3(groups), dim3(64 * kWavesPerGroup), 0, stream, v, io,
                               bi, ctl, slotIn, slotOut, slotClear, dtLog, logIdx);
        }
    };
    if (xseg == 4)
        go(std::integral_constant<int, 4>());
    else if (xseg == 10)
        go(std::integral_constant<int, 10>());
    else if (xseg == 12)
        go(std::integral_constant<int, 12>());
    else if (xseg == 16)
        go(std::integral_constant<int, 16>());
    else if (xseg == 32)
        go(std::integral_constant<int, 32>());
    else
        go(std::integral_constant<int, 8>());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "fused_u_kernel launch");
    return 0;
}

template <int S, int V>
int launch_solve_edges_solver(int64_t n, const float* in, float* out, hipStream_t stream) {
    const int64_t groups = (n + 255) / 256;
    hipLaunchKernelGGL((solve_edges_kernel<S, V>), dim3((unsigned)groups), dim3(256), 0, stream, n, in, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "solve_edges_kernel launch");
    return 0;
}

template int launch_sweep_u_solver<SWE_SOLVER_ID, SWE_VARIANT>(int, int, StateView, SumsView, float*, float*,
                                                               float*, BoundaryInfo, DeviceCtl*, Slots, float, int,
                                                               float*, int, hipStream_t);
#if SWE_SOLVER_ID != 4
template int launch_solve_edges_solver<SWE_SOLVER_ID, SWE_VARIANT>(int64_t, const float*, float*, hipStream_t);
#endif
template int launch_fused_u_solver<SWE_SOLVER_ID, SWE_VARIANT>(int, StateView, FusedIO, BoundaryInfo, DeviceCtl*, int,
                                                               int, int, float*, int, hipStream_t);

#if SWE_VARIANT == 0 && SWE_SOLVER_ID != 4
template <int S>
int launch_sweep_s_solver(int dir, StateView v, SumsView sums, DeviceCtl* ctl, int par, hipStream_t stream) {
    if (dir == 0) {
        const int nyt = (v.ny + 2 + 63) / 64;
        const int nxs = (v.nx + 1 + kSplitXSeg - 1) / kSplitXSeg;
        const int groups = (nyt * nxs + kWavesPerGroup - 1) / kWavesPerGroup;
        hipLaunchKernelGGL((sweep_s_x_kernel<S>), dim3(groups), dim3(256), 0, stream, v, sums, ctl, par);
    } else {
        const int nyt = (v.ny + kStripRows - 1) / kStripRows;
        const int nxs = (v.nx + kSplitXSeg - 1) / kSplitXSeg;
        const int groups = (nyt * nxs + kWavesPerGroup - 1) / kWavesPerGroup;
        hipLaunchKernelGGL((sweep_s_y_kernel<S>), dim3(groups), dim3(256), 0, stream, v, sums);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sweep_s kernel launch");
    return 0;
}
template int launch_sweep_s_solver<SWE_SOLVER_ID>(int, StateView, SumsView, DeviceCtl*, int, hipStream_t);
#endif

}  // namespace swe
