// launch_batch.hip -- per-solver instantiations of the batched strip sweep.
// Compiled once per solver with -DSWE_SOLVER_ID=<id> (IEEE arithmetic only:
// the batched engine keeps the single-block path's bit-exact results).
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

template int launch_batch_sweep_solver<SWE_SOLVER_ID>(int, int, BatchArena, const BatchBlockDesc*, BatchBlockState*,
                                                      BatchGlobals*, const int*, hipStream_t);

template <int S>
int launch_batch_fused_solver(int xs, int groups, int direct, BatchArena A, const BatchBlockDesc* D,
                              BatchBlockState* st, BatchGlobals* G, const int* items, hipStream_t stream) {
    const dim3 grid(groups), block(64 * kWavesPerGroup);
    switch (xs) {
        case 4:
            hipLaunchKernelGGL((batch_fused_kernel<S, 4>), grid, block, 0, stream, A, D, st, G, items, direct);
            break;
        case 16:
            hipLaunchKernelGGL((batch_fused_kernel<S, 16>), grid, block, 0, stream, A, D, st, G, items, direct);
            break;
        default:
            hipLaunchKernelGGL((batch_fused_kernel<S, 8>), grid, block, 0, stream, A, D, st, G, items, direct);
            break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_fused_kernel launch");
    return 0;
}

template int launch_batch_fused_solver<SWE_SOLVER_ID>(int, int, int, BatchArena, const BatchBlockDesc*,
                                                      BatchBlockState*, BatchGlobals*, const int*, hipStream_t);

}  // namespace swe
