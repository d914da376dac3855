// swe_internal.h -- declarations shared between the translation units of
// libswe_hip.so (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels_batch.hpp"
#include "kernels_fused.hpp"
#include "kernels_u.hpp"

#include "../../include/swe_hip.h"
#include "../../include/swe_hip_batch.h"
#include "../../include/swe_hip_mlp.h"

namespace swe {

void set_error(const char* fmt, ...);

// Cross-unit accessors: the current state of a flavour-U block and the
// device weight fragments of a network.
int block_view(swe_hip_block* k, StateView* v, hipStream_t* stream);
const unsigned* mlp_frags(swe_hip_mlp* m);
// Attachment bookkeeping of a network: `user` is the attached block's or
// batch's stream (swe_mlp.hip).
void mlp_attach(swe_hip_mlp* m, hipStream_t user);
void mlp_detach(swe_hip_mlp* m, hipStream_t user);
int hip_fail(hipError_t e, const char* what);

// Host copy of the flavour-U dt rule (SWE_DimensionalSplittingMpi.cpp:448-456);
// same float operations as the device version.
inline float host_timestep_u(float s, float dx, float dy) {
    if (s > kMinSpeed) {
        float dt = (dy / s < dx / s) ? dy / s : dx / s;
        dt *= kCfl;
        return dt;
    }
    return FLT_MAX;
}

// Strip-sweep launch (flavour U), dispatched on solver, mode and segment length.
int launch_sweep_u(int solver, int variant, int mode, int xseg, StateView v, SumsView sums, float* hn, float* hun, float* hvn,
                   BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog, int logIdx,
                   hipStream_t stream);
int launch_apply_u(int variant, StateView v, SumsView sums, BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed,
                   int fuseBc, float* dtLog, int logIdx, hipStream_t stream);

int launch_apply_bc(StateView v, BoundaryInfo bi, hipStream_t stream);
int launch_apply_bc_bathymetry(float* b, int nx, int ny, int pitch, BoundaryInfo bi, hipStream_t stream);
int launch_ctl_reset(DeviceCtl* ctl, int par, hipStream_t stream);
// One-shot timing of the next fused launch made on this thread:
// launch_fused_u records these events on the kernel's own dispatch
// (hipExtLaunchKernelGGL) and clears them.  Event markers between launches
// would make the stream pause; these do not.
struct LaunchTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTimer g_fusedTimer;
// Words of the host ring the readback kernel writes (slot seq % kReadbackRing).
constexpr unsigned kReadbackRing = 4;
int launch_readback(const DeviceCtl* ctl, int slot, unsigned seq, hipStream_t stream);
// Quiet-tile plan of fused launch `step` at strip width xseg (a multiple of
// kQuietCols; kernels_fused.hpp): skipped tiles carry their speeds, the busy
// runs are listed for the launch.  slotPrev is the CFL slot of the previous
// launch (its busy-tile count sets the piece length), slotOut the launch's;
// rbSeq != 0 also reports slot slotPrev to the host (launch_readback).
int launch_quiet_plan(const unsigned* quiet, float* speed, int nx, int ny, int xseg, int step, DeviceCtl* ctl,
                      int slotPrev, int slotOut, unsigned* list, unsigned* count, unsigned rbSeq, hipStream_t stream);

// Flavour S.
int launch_sweep_s_x(int solver, StateView v, SumsView sums, DeviceCtl* ctl, int par, hipStream_t stream);
int launch_sweep_s_y(int solver, StateView v, SumsView sums, hipStream_t stream);
int launch_apply_s(StateView v, SumsView sums, float dt, hipStream_t stream);

template <int V>
int launch_apply_u_variant(StateView v, SumsView sums, BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed,
                           int fuseBc, float* dtLog, int logIdx, hipStream_t stream);
// A source block's copy layer (h, hu, hv at sh/shu/shv, stride sstride, its
// pending update's sums at the same offsets) updated with dt on the fly into
// a destination ghost line (dh/dhu/dhv, stride dstride), n cells.
struct LinePending {
    float *dh, *dhu, *dhv;
    size_t dstride;
    const float *sh, *shu, *shv;
    SumsView sums;
    size_t sstride;
    int n;
    float dt, dx, dy;
};
template <int V>
int launch_copy_line_pending_variant(const LinePending& L, hipStream_t stream);
int launch_solve_edges(int solver, int variant, int64_t n, const float* in, float* out, hipStream_t stream);

// Per-solver instantiations (one translation unit each, compiled in parallel).
template <int S, int V>
int launch_sweep_u_solver(int mode, int xseg, StateView v, SumsView sums, float* hn, float* hun, float* hvn,
                          BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog,
                          int logIdx, hipStream_t stream);
template <int S, int V>
int launch_fused_u_solver(int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn, int slotOut,
                          int slotClear, float* dtLog, int logIdx, hipStream_t stream);
int launch_fused_u(int solver, int variant, int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl,
                   int slotIn, int slotOut, int slotClear, float* dtLog, int logIdx, hipStream_t stream);
// Learned middle state with deferred hard edges (launch_defer.hip): the
// deferring fused sweep (io.deferQ set), then the launch that solves the
// edges it queued in ctl->defer[slotOut] and completes the sums of vNew.
int launch_fused_defer(int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn, int slotOut,
                       int slotClear, float* dtLog, int logIdx, hipStream_t stream);
int launch_defer_solve(StateView vNew, SumsView sums, DeviceCtl* ctl, int slot, const uint2* queue, unsigned cap,
                       int groups, hipStream_t stream);
template <int S, int V>
int launch_solve_edges_solver(int64_t n, const float* in, float* out, hipStream_t stream);
template <int S>
int launch_sweep_s_solver(int dir, StateView v, SumsView sums, DeviceCtl* ctl, int par, hipStream_t stream);

// Batched engine: strip sweep over the queued tiles (one translation unit per solver).
template <int S>
int launch_batch_sweep_solver(int xs, int groups, BatchArena A, const BatchBlockDesc* D, BatchBlockState* st,
                              BatchGlobals* G, const int* items, hipStream_t stream);
// The fused sweep also does the ghost-line steps (jobs, geometry: 4 per block).
template <int S>
int launch_batch_fused_solver(int xs, int groups, int direct, BatchArena A, const BatchBlockDesc* D,
                              BatchBlockState* st, BatchGlobals* G, const int* items, uint2* deferQ,
                              const GhostJob* jobs, const GhostGeom* geo, hipStream_t stream);
// The edges the deferring batched sweep queued (G->deferCount of them).
int launch_batch_defer_solve(BatchArena A, const BatchBlockDesc* D, BatchBlockState* st, BatchGlobals* G,
                             const uint2* queue, unsigned cap, int groups, hipStream_t stream);

}  // namespace swe
