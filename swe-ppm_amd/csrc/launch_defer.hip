// launch_defer.hip -- the learned middle state with its hard edges deferred,
// for the fused device loop of a single block.
//
// The deferring sweep (fused_u_kernel<SOLVER_AUGRIE_DEFER>) settles the edges
// whose linearised middle-state guess passes the residual check and queues the
// others; it holds no network call and no Newton loop for wet edges, so it
// keeps the exact solver's occupancy.  defer_solve_kernel then solves the
// queued edges with the full learned middle state (network guess, residual
// check, exact iteration where that fails), 64 edges per wave with every lane
// active for the matrix-core network, and adds their parts to the sums and
// their speeds to the CFL maximum.  Every sum has two terms, so the order of
// the additions does not matter: the result equals the learned middle state
// solved inside the sweep.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels_fused.hpp"
#include "swe_internal.h"

namespace swe {

namespace {

__global__ __launch_bounds__(256) void defer_solve_kernel(StateView v, SumsView s, DeviceCtl* ctl, int slot,
                                                          const uint2* __restrict__ q, unsigned cap) {
    const unsigned count = min(ctl->defer[slot], cap);
    const int lane = threadIdx.x & 63;
    const unsigned wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const unsigned stride = gridDim.x * kWavesPerGroup * 64u;
    const dev::SolverCtx<SOLVER_AUGRIE_MLP> solver(ctl->mlp);
    float smaxv = 0.f;
    for (unsigned base = wave * 64u; base < count; base += stride) {  // uniform per wave
        const unsigned i = base + lane;
        const bool valid = i < count;
        const uint2 e = q[valid ? i : count - 1];  // surplus lanes repeat the last edge, unused
        const bool xEdge = (e.x & kDeferY) == 0u;
        const int x = (int)(e.x & ~kDeferY), y = (int)e.y;
        const int xr = xEdge ? x + 1 : x, yr = xEdge ? y : y + 1;
        const size_t cl = idx(v, x, y), cr = idx(v, xr, yr);
        const dev::Side L = dev::cell_side(v.h[cl], xEdge ? v.hu[cl] : v.hv[cl], v.b[cl]);
        const dev::Side R = dev::cell_side(v.h[cr], xEdge ? v.hu[cr] : v.hv[cr], v.b[cr]);
        const dev::EdgeOut o = solver.solve(L, R);
        if (valid) {
            float* sh = xEdge ? s.sx : s.sy;
            float* sm = xEdge ? s.su : s.sv;
            if (xEdge ? x >= 1 : y >= 1) {
                atomicAdd(&sh[cl], o.hL);
                atomicAdd(&sm[cl], o.huL);
            }
            if (xEdge ? xr <= v.nx : yr <= v.ny) {
                atomicAdd(&sh[cr], o.hR);
                atomicAdd(&sm[cr], o.huR);
            }
            smaxv = dev::smax(smaxv, o.speed);
        }
    }
    smaxv = wave_max(smaxv);
    if (lane == 0 && smaxv > 0.f) atomicMax(&ctl->smax[slot], __float_as_uint(smaxv));
}

}  // namespace

int launch_fused_defer(int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn, int slotOut,
                       int slotClear, float* dtLog, int logIdx, hipStream_t stream) {
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    auto go = [&](auto xsTag) {
        constexpr int XS = decltype(xsTag)::value;
        const int nxs = (v.nx + XS - 1) / XS;
        // list mode may list up to one piece per record tile
        const int waves = io.list ? nyt * ((v.nx + kQuietCols - 1) / kQuietCols) : nyt * nxs;
        const int groups = (waves + kWavesPerGroup - 1) / kWavesPerGroup;
        hipLaunchKernelGGL((fused_u_kernel<SOLVER_AUGRIE_DEFER, XS, 0>), dim3(groups), dim3(64 * kWavesPerGroup), 0,
                           stream, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog, logIdx);
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
    if (e != hipSuccess) return hip_fail(e, "fused_u_kernel (deferring) launch");
    return 0;
}

int launch_defer_solve(StateView vNew, SumsView sums, DeviceCtl* ctl, int slot, const uint2* queue, unsigned cap,
                       int groups, hipStream_t stream) {
    hipLaunchKernelGGL(defer_solve_kernel, dim3(groups), dim3(64 * kWavesPerGroup), 0, stream, vNew, sums, ctl, slot,
                       queue, cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "defer_solve_kernel launch");
    return 0;
}

}  // namespace swe
