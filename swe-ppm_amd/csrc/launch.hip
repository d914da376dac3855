// launch.hip -- solver-independent launches and the solver dispatch.
#include <hip/hip_runtime.h>

#include "kernels_s.hpp"
#include "kernels_u.hpp"
#include "swe_internal.h"

namespace swe {

// setGhostLayer for OUTFLOW/WALL sides plus the corners (SWE_Block.hh:757-911).
__global__ void apply_bc_kernel(StateView v, BoundaryInfo bi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = v.nx, ny = v.ny;
    if (k < ny) {
        const int j = k + 1;
        for (int side = BND_LEFT; side <= BND_RIGHT; ++side) {
            const int t = bi.type[side];
            if (t != OUTFLOW && t != WALL) continue;
            const int gx = side == BND_LEFT ? 0 : nx + 1;
            const int sx = side == BND_LEFT ? 1 : nx;
            const size_t g = idx(v, gx, j), s = idx(v, sx, j);
            v.h[g] = v.h[s];
            v.hu[g] = (t == WALL) ? -v.hu[s] : v.hu[s];
            v.hv[g] = v.hv[s];
        }
    } else if (k < ny + nx) {
        const int i = k - ny + 1;
        for (int side = BND_BOTTOM; side <= BND_TOP; ++side) {
            const int t = bi.type[side];
            if (t != OUTFLOW && t != WALL) continue;
            const int gy = side == BND_BOTTOM ? 0 : ny + 1;
            const int sy = side == BND_BOTTOM ? 1 : ny;
            const size_t g = idx(v, i, gy), s = idx(v, i, sy);
            v.h[g] = v.h[s];
            v.hu[g] = v.hu[s];
            v.hv[g] = (t == WALL) ? -v.hv[s] : v.hv[s];
        }
    } else if (k < ny + nx + 4) {
        const int c = k - ny - nx;
        const int gx = (c < 2) ? 0 : nx + 1;
        const int gy = (c & 1) ? ny + 1 : 0;
        const int sx = (c < 2) ? 1 : nx;
        const int sy = (c & 1) ? ny : 1;
        const size_t g = idx(v, gx, gy), s = idx(v, sx, sy);
        v.h[g] = v.h[s];
        v.hu[g] = v.hu[s];
        v.hv[g] = v.hv[s];
    }
}

// applyBoundaryBathymetry (SWE_Block.hh:724-749): whole ghost lines incl.
// corners for OUTFLOW/WALL sides, then the four corners.
__global__ void apply_bc_bathymetry_kernel(float* b, int nx, int ny, int pitch, BoundaryInfo bi) {
    // single thread block; sequential order matters only for the corners,
    // which are written last
    const int tid = threadIdx.x;
    auto ow = [](int t) { return t == OUTFLOW || t == WALL; };
    if (ow(bi.type[BND_LEFT]))
        for (int j = tid; j < ny + 2; j += blockDim.x) b[j] = b[(size_t)pitch + j];
    if (ow(bi.type[BND_RIGHT]))
        for (int j = tid; j < ny + 2; j += blockDim.x) b[(size_t)(nx + 1) * pitch + j] = b[(size_t)nx * pitch + j];
    __syncthreads();
    if (ow(bi.type[BND_BOTTOM]))
        for (int i = tid; i <= nx + 1; i += blockDim.x) b[(size_t)i * pitch] = b[(size_t)i * pitch + 1];
    if (ow(bi.type[BND_TOP]))
        for (int i = tid; i <= nx + 1; i += blockDim.x) b[(size_t)i * pitch + ny + 1] = b[(size_t)i * pitch + ny];
    __syncthreads();
    if (tid == 0) {
        b[0] = b[(size_t)pitch + 1];
        b[ny + 1] = b[(size_t)pitch + ny];
        b[(size_t)(nx + 1) * pitch] = b[(size_t)nx * pitch + 1];
        b[(size_t)(nx + 1) * pitch + ny + 1] = b[(size_t)nx * pitch + ny];
    }
}

// Skip decisions for fused launch `step` (>= 3) over the wave segments of
// XS = fine * kQuietCols columns.  One lane per record tile: the `lanes`
// lanes of a segment (plan_lanes(fine)) sit side by side in a wave,
// and a ballot gives each segment its busy tiles.  A skipped record tile
// carries its speed into slot step & 1 and into the CFL maximum of slot
// slotOut.  The record tiles of a segment from its first to its last busy
// one are listed in pieces of at most `piece` tiles,
// (segment << 6) | (first << 3) | (length - 1); workgroup g keeps its
// entries in segment order at list[g * kPlanGroup * kPlanPieces ...] and
// their number in count[g].  piece adapts to the load: busy tiles found by
// the previous launch's plan over kPlanWaves, between 1 and `fine` (with few
// busy tiles short pieces spread them over more waves, the busiest wave sets
// the launch time; with many, long pieces solve fewer halo edges).  The
// number of busy tiles goes to ctl->busy[slotOut].  rbSeq != 0: the plan also
// reports the previous launch's CFL maximum to the host, as readback_kernel
// does (one kernel less between a flux pass and the next).
__global__ __launch_bounds__(kPlanGroup) void quiet_plan_kernel(const unsigned* __restrict__ quiet, float* speed,
                                                                int nyt, int nxs, int nxf, int fine, int lanes,
                                                                int step, DeviceCtl* ctl, int slotPrev3, int slotOut,
                                                                unsigned* list, unsigned* count, unsigned rbSeq) {
    __shared__ unsigned wavePieces[kPlanGroup / 64];
    __shared__ unsigned waveBusy[kPlanGroup / 64];
    __shared__ unsigned waveMax[kPlanGroup / 64];
    const int tiles = nyt * nxf;
    const int g = blockIdx.x * kPlanGroup + threadIdx.x;
    const int seg = g / lanes, i = g % lanes;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slotNow = (step & 1) * tiles, slotPrev = ((step - 1) & 1) * tiles;
    const unsigned prevBusy = ctl->busy[slotPrev3];
    if (rbSeq != 0u && blockIdx.x == 0 && threadIdx.x == 0)  // the previous launch's maximum to the host
        __hip_atomic_store(&ctl->ring[rbSeq % kReadbackRing], ((unsigned long long)rbSeq << 32) | ctl->smax[slotPrev3],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const int piece = prevBusy == 0u ? fine : (int)min((unsigned)fine, max(1u, (prevBusy + kPlanWaves - 1) / kPlanWaves));
    const int yt = seg % nyt, xs = seg / nyt;
    const int f = xs * fine + i;
    bool busy = false;
    float sp = 0.f;
    if (seg < nyt * nxs && i < fine && f < nxf) {
        const int t = yt + nyt * f;
        const unsigned* q = quiet;
        // every load of the tile at once: the flags and the speed to carry
        const unsigned self = q[slotPrev + t], selfOld = q[slotNow + t];
        const unsigned left = f == 0 ? kQuietRight : q[slotPrev + t - nyt];
        const unsigned right = f == nxf - 1 ? kQuietLeft : q[slotPrev + t + nyt];
        const unsigned below = yt == 0 ? kQuietTop : q[slotPrev + t - 1];
        const unsigned above = yt == nyt - 1 ? kQuietBottom : q[slotPrev + t + 1];
        const float spPrev = speed[slotPrev + t];
        const bool skip = (self & kQuietAll) && (selfOld & kQuietAll) && (left & kQuietRight) &&
                          (right & kQuietLeft) && (below & kQuietTop) && (above & kQuietBottom);
        if (skip) {
            sp = spPrev;
            speed[slotNow + t] = sp;
        } else {
            busy = true;
        }
    }
    // the segment's busy tiles, from the ballot bits of its lanes
    const unsigned long long m = __ballot(busy);
    const unsigned segBits = (unsigned)(m >> (lane & ~(lanes - 1))) & ((1u << lanes) - 1u);
    const bool leader = i == 0 && segBits != 0u;
    const int firstBusy = leader ? __builtin_ctz(segBits) : 0;
    const int runLen = leader ? 32 - __builtin_clz(segBits) - firstBusy : 0;
    const unsigned pieces = (unsigned)((runLen + piece - 1) / piece);
    // exclusive prefix of the pieces over the wave, then over the workgroup
    unsigned incl = pieces;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned o = __shfl_up(incl, off, 64);
        incl += (lane >= off) ? o : 0u;
    }
    sp = wave_max(sp);
    if (lane == 63) wavePieces[w] = incl;
    if (lane == 0) {
        waveBusy[w] = (unsigned)__popcll(m);
        waveMax[w] = __float_as_uint(sp);
    }
    __syncthreads();
    unsigned before = 0, total = 0, nbusy = 0, mx = 0;
    for (int k = 0; k < kPlanGroup / 64; ++k) {
        const unsigned c = wavePieces[k];
        before += (k < w) ? c : 0u;
        total += c;
        nbusy += waveBusy[k];
        mx = max(mx, waveMax[k]);
    }
    const unsigned pos = (unsigned)blockIdx.x * kPlanGroup * kPlanPieces + before + incl - pieces;
    for (unsigned p = 0; p < pieces; ++p) {
        const int f0 = firstBusy + (int)p * piece;
        const int len = min(piece, runLen - (int)p * piece);
        list[pos + p] = ((unsigned)seg << 6) | ((unsigned)f0 << 3) | (unsigned)(len - 1);
    }
    if (threadIdx.x == 0) {
        count[blockIdx.x] = total;
        if (nbusy) atomicAdd(&ctl->busy[slotOut], nbusy);
        if (mx > 0u) atomicMax(&ctl->smax[slotOut], mx);
    }
}

// The CFL maximum of a slot to the host: one 64-bit word (sequence number
// above, float bits below) stored at system scope into coherent host memory,
// ctl->ring[seq % kReadbackRing], which the host polls for that number.
// A kernel of its own after the flux pass: it needs no event in the stream
// (an event there costs the stream a pause before its next kernel), and
// the host learns the maximum while the stream runs on.
__global__ void readback_kernel(const DeviceCtl* ctl, int slot, unsigned seq) {
    if (threadIdx.x == 0) {
        const unsigned long long w = ((unsigned long long)seq << 32) | ctl->smax[slot];
        __hip_atomic_store(&ctl->ring[seq % kReadbackRing], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void ctl_reset_kernel(DeviceCtl* ctl, int par) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->smax[par] = 0u;
        ctl->defer[par] = 0u;
        ctl->busy[par] = 0u;
    }
}

__global__ __launch_bounds__(256) void apply_s_kernel(StateView v, SumsView sums, float dt) {
    const int y = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
    const int x0 = 1 + (blockIdx.y * 4 + (threadIdx.x >> 6)) * kSplitXSeg;
    if (y > v.ny || x0 > v.nx) return;
    const int x1 = min(x0 + kSplitXSeg - 1, v.nx);
    const float dtdx = dt / v.dx;
    for (int x = x0; x <= x1; ++x) {
        const size_t c = idx(v, x, y);
        const float hStar = v.h[c] - dtdx * sums.sx[c];
        const float huStar = v.hu[c] - dtdx * sums.su[c];
        v.h[c] = hStar - dtdx * sums.sy[c];
        v.hu[c] = huStar;
        v.hv[c] = v.hv[c] - dtdx * sums.sv[c];
    }
}


template <int V>
static int sweep_u_v(int solver, int mode, int xseg, StateView v, SumsView sums, float* hn, float* hun, float* hvn,
                     BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog, int logIdx,
                     hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_sweep_u_solver<SOLVER_HLLE, V>(mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed,
                                                         fuseBc, dtLog, logIdx, stream);
        case SOLVER_FWAVE:
            return launch_sweep_u_solver<SOLVER_FWAVE, V>(mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed,
                                                          fuseBc, dtLog, logIdx, stream);
        case SOLVER_AUGRIE:
            return launch_sweep_u_solver<SOLVER_AUGRIE, V>(mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed,
                                                           fuseBc, dtLog, logIdx, stream);
        case SOLVER_HYBRID:
            return launch_sweep_u_solver<SOLVER_HYBRID, V>(mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed,
                                                           fuseBc, dtLog, logIdx, stream);
        case SOLVER_AUGRIE_MLP:
            // the learned middle state is built with IEEE arithmetic only
            return launch_sweep_u_solver<SOLVER_AUGRIE_MLP, 0>(mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl,
                                                               dtFixed, fuseBc, dtLog, logIdx, stream);
        default:
            set_error("unknown solver %d", solver);
            return -1;
    }
}

int launch_sweep_u(int solver, int variant, int mode, int xseg, StateView v, SumsView sums, float* hn, float* hun,
                   float* hvn, BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed, int fuseBc, float* dtLog,
                   int logIdx, hipStream_t stream) {
    if (variant == 1)
        return sweep_u_v<1>(solver, mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx,
                            stream);
    return sweep_u_v<0>(solver, mode, xseg, v, sums, hn, hun, hvn, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx,
                        stream);
}

int launch_apply_u(int variant, StateView v, SumsView sums, BoundaryInfo bi, DeviceCtl* ctl, Slots sl, float dtFixed,
                   int fuseBc, float* dtLog, int logIdx, hipStream_t stream) {
    if (variant == 1) return launch_apply_u_variant<1>(v, sums, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
    return launch_apply_u_variant<0>(v, sums, bi, ctl, sl, dtFixed, fuseBc, dtLog, logIdx, stream);
}

template <int V>
static int solve_edges_v(int solver, int64_t n, const float* in, float* out, hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_solve_edges_solver<SOLVER_HLLE, V>(n, in, out, stream);
        case SOLVER_FWAVE:
            return launch_solve_edges_solver<SOLVER_FWAVE, V>(n, in, out, stream);
        case SOLVER_AUGRIE:
            return launch_solve_edges_solver<SOLVER_AUGRIE, V>(n, in, out, stream);
        default:
            return launch_solve_edges_solver<SOLVER_HYBRID, V>(n, in, out, stream);
    }
}

template <int V>
static int fused_u_v(int solver, int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn,
                     int slotOut, int slotClear, float* dtLog, int logIdx, hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_fused_u_solver<SOLVER_HLLE, V>(xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog,
                                                         logIdx, stream);
        case SOLVER_FWAVE:
            return launch_fused_u_solver<SOLVER_FWAVE, V>(xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog,
                                                          logIdx, stream);
        case SOLVER_AUGRIE:
            return launch_fused_u_solver<SOLVER_AUGRIE, V>(xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog,
                                                           logIdx, stream);
        case SOLVER_AUGRIE_MLP:
            return launch_fused_u_solver<SOLVER_AUGRIE_MLP, 0>(xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog,
                                                               logIdx, stream);
        default:
            return launch_fused_u_solver<SOLVER_HYBRID, V>(xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog,
                                                           logIdx, stream);
    }
}

int launch_fused_u(int solver, int variant, int xseg, StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl,
                   int slotIn, int slotOut, int slotClear, float* dtLog, int logIdx, hipStream_t stream) {
    if (variant == 1) return fused_u_v<1>(solver, xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog, logIdx, stream);
    return fused_u_v<0>(solver, xseg, v, io, bi, ctl, slotIn, slotOut, slotClear, dtLog, logIdx, stream);
}

int launch_solve_edges(int solver, int variant, int64_t n, const float* in, float* out, hipStream_t stream) {
    return variant == 1 ? solve_edges_v<1>(solver, n, in, out, stream) : solve_edges_v<0>(solver, n, in, out, stream);
}

static int dispatch_s(int solver, int dir, StateView v, SumsView sums, DeviceCtl* ctl, int par,
                      hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_sweep_s_solver<SOLVER_HLLE>(dir, v, sums, ctl, par, stream);
        case SOLVER_FWAVE:
            return launch_sweep_s_solver<SOLVER_FWAVE>(dir, v, sums, ctl, par, stream);
        case SOLVER_AUGRIE:
            return launch_sweep_s_solver<SOLVER_AUGRIE>(dir, v, sums, ctl, par, stream);
        case SOLVER_HYBRID:
            return launch_sweep_s_solver<SOLVER_HYBRID>(dir, v, sums, ctl, par, stream);
        default:
            set_error("unknown solver %d", solver);
            return -1;
    }
}

int launch_sweep_s_x(int solver, StateView v, SumsView sums, DeviceCtl* ctl, int par, hipStream_t stream) {
    return dispatch_s(solver, 0, v, sums, ctl, par, stream);
}

int launch_sweep_s_y(int solver, StateView v, SumsView sums, hipStream_t stream) {
    return dispatch_s(solver, 1, v, sums, nullptr, 0, stream);
}

int launch_apply_bc(StateView v, BoundaryInfo bi, hipStream_t stream) {
    const int n = v.nx + v.ny + 4;
    hipLaunchKernelGGL(apply_bc_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, v, bi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "apply_bc_kernel launch");
    return 0;
}

int launch_apply_bc_bathymetry(float* b, int nx, int ny, int pitch, BoundaryInfo bi, hipStream_t stream) {
    hipLaunchKernelGGL(apply_bc_bathymetry_kernel, dim3(1), dim3(256), 0, stream, b, nx, ny, pitch, bi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "apply_bc_bathymetry_kernel launch");
    return 0;
}

int launch_quiet_plan(const unsigned* quiet, float* speed, int nx, int ny, int xseg, int step, DeviceCtl* ctl,
                      int slotPrev, int slotOut, unsigned* list, unsigned* count, unsigned rbSeq,
                      hipStream_t stream) {
    const int nyt = (ny + kStripRows - 1) / kStripRows, nxs = (nx + xseg - 1) / xseg;
    const int nxf = (nx + kQuietCols - 1) / kQuietCols;
    const int fine = xseg / kQuietCols;
    hipLaunchKernelGGL(quiet_plan_kernel, dim3((unsigned)plan_groups(nyt, nxs, fine)), dim3(kPlanGroup), 0, stream,
                       quiet, speed, nyt, nxs, nxf, fine, plan_lanes(fine), step, ctl, slotPrev, slotOut, list,
                       count, rbSeq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "quiet_plan_kernel launch");
    return 0;
}

int launch_readback(const DeviceCtl* ctl, int slot, unsigned seq, hipStream_t stream) {
    hipLaunchKernelGGL(readback_kernel, dim3(1), dim3(64), 0, stream, ctl, slot, seq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "readback_kernel launch");
    return 0;
}

int launch_ctl_reset(DeviceCtl* ctl, int par, hipStream_t stream) {
    hipLaunchKernelGGL(ctl_reset_kernel, dim3(1), dim3(64), 0, stream, ctl, par);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "ctl_reset_kernel launch");
    return 0;
}

int launch_apply_s(StateView v, SumsView sums, float dt, hipStream_t stream) {
    const int nxs = (v.nx + kSplitXSeg - 1) / kSplitXSeg;
    dim3 grid((v.ny + 63) / 64, (nxs + 3) / 4);
    hipLaunchKernelGGL(apply_s_kernel, grid, dim3(256), 0, stream, v, sums, dt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "apply_s_kernel launch");
    return 0;
}

}  // namespace swe
