// launch_batch.hip -- per-solver instantiations of the batched strip sweep.
// Compiled once per solver with -DSWE_SOLVER_ID=<id> (IEEE arithmetic only:
// the batched engine keeps the single-block path's bit-exact results).  The
// deferring learned middle state (id 5) has the fused sweep only, and the
// launch that solves the edges it defers.
#include <hip/hip_runtime.h>

#include "kernels_batch.hpp"
#include "swe_internal.h"

#ifndef SWE_SOLVER_ID
#error "SWE_SOLVER_ID must be defined"
#endif

namespace swe {

template <int S>
int launch_batch_sweep_solver(int xs, int groups, BatchArena A, const BatchBlockDesc* D, BatchBlockState* st,
                              BatchGlobals* G, const int* items, hipStream_t stream) {
    const dim3 grid(groups), block(64 * kWavesPerGroup);
    switch (xs) {
        case 4:
            hipLaunchKernelGGL((batch_sweep_kernel<S, 4>), grid, block, 0, stream, A, D, st, G, items);
            break;
        case 16:
            hipLaunchKernelGGL((batch_sweep_kernel<S, 16>), grid, block, 0, stream, A, D, st, G, items);
            break;
        default:
            hipLaunchKernelGGL((batch_sweep_kernel<S, 8>), grid, block, 0, stream, A, D, st, G, items);
            break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_sweep_kernel launch");
    return 0;
}

#if SWE_SOLVER_ID != 5
template int launch_batch_sweep_solver<SWE_SOLVER_ID>(int, int, BatchArena, const BatchBlockDesc*, BatchBlockState*,
                                                      BatchGlobals*, const int*, hipStream_t);
#endif

template <int S>
int launch_batch_fused_solver(int xs, int groups, int direct, BatchArena A, const BatchBlockDesc* D,
                              BatchBlockState* st, BatchGlobals* G, const int* items, uint2* deferQ,
                              const GhostJob* jobs, const GhostGeom* geo, hipStream_t stream) {
    const dim3 grid(groups), block(64 * kWavesPerGroup);
    switch (xs) {
        case 4:
            hipLaunchKernelGGL((batch_fused_kernel<S, 4>), grid, block, 0, stream, A, D, st, G, items, direct, deferQ,
                               jobs, geo);
            break;
        case 16:
            hipLaunchKernelGGL((batch_fused_kernel<S, 16>), grid, block, 0, stream, A, D, st, G, items, direct, deferQ,
                               jobs, geo);
            break;
        default:
            hipLaunchKernelGGL((batch_fused_kernel<S, 8>), grid, block, 0, stream, A, D, st, G, items, direct, deferQ,
                               jobs, geo);
            break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_fused_kernel launch");
    return 0;
}

template int launch_batch_fused_solver<SWE_SOLVER_ID>(int, int, int, BatchArena, const BatchBlockDesc*,
                                                      BatchBlockState*, BatchGlobals*, const int*, uint2*,
                                                      const GhostJob*, const GhostGeom*, hipStream_t);

#if SWE_SOLVER_ID == 5
namespace {

// The edges the deferring fused sweep queued this iteration, solved with the
// full learned middle state (network guess, residual check, exact iteration)
// on the state the sweep wrote (buffer cur ^ 1, sums sidx ^ 1 of each block,
// before the round flips them); parts added to the two-term sums, speeds to
// the block's CFL maximum.  64 edges per wave, every lane active.
__global__ __launch_bounds__(256) void batch_defer_solve_kernel(BatchArena A, const BatchBlockDesc* D,
                                                                BatchBlockState* st, BatchGlobals* G,
                                                                const uint2* __restrict__ q, unsigned cap) {
    if (G->skip) return;
    const unsigned count = min(G->deferCount, cap);
    const int lane = threadIdx.x & 63;
    const unsigned wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const unsigned stride = gridDim.x * kWavesPerGroup * 64u;
    const dev::SolverCtx<SOLVER_AUGRIE_MLP> solver(G->mlp);
    for (unsigned base = wave * 64u; base < count; base += stride) {  // uniform per wave
        const unsigned i = base + lane;
        const bool valid = i < count;
        const uint2 e = q[valid ? i : count - 1];  // surplus lanes repeat the last edge, unused
        const bool xEdge = (e.x & kDeferY) == 0u;
        const int blk = (int)((e.x & ~kDeferY) >> kDeferBlockShift);
        const int x = (int)(e.x & ((1u << kDeferBlockShift) - 1u)), y = (int)e.y;
        const BatchBlockDesc& d = D[blk];
        const StateView v = batch_view(A, d, st[blk].cur ^ 1);
        const SumsView s = batch_sums(A, d, st[blk].sidx ^ 1);
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
            if (o.speed > 0.f) atomicMax(&st[blk].smax, __float_as_uint(o.speed));
        }
    }
}

}  // namespace

int launch_batch_defer_solve(BatchArena A, const BatchBlockDesc* D, BatchBlockState* st, BatchGlobals* G,
                             const uint2* queue, unsigned cap, int groups, hipStream_t stream) {
    hipLaunchKernelGGL(batch_defer_solve_kernel, dim3(groups), dim3(64 * kWavesPerGroup), 0, stream, A, D, st, G,
                       queue, cap);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_defer_solve_kernel launch");
    return 0;
}
#endif

}  // namespace swe
