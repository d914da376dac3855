// kernels_s.hpp -- flavour-S step kernels: swe_simple's split sweep
// (src/blocks/SWE_DimensionalSplitting.cpp:105-210):
//   x-sweep over x in [0, nx], all rows incl. ghosts -> dt = 0.4 dx / max
//   Q* = Q - dt/dx (L + R)                 for x in [1, nx]
//   y-sweep over y in [0, ny] on the OLD h, hv (not Q*)
//   h = h* - dt/dx (B + A), hu = hu*, hv = hv - dt/dx (B + A)
// (the y-part uses dx, as the reference does).  The net updates are kept as
// per-cell sums in the same four device arrays flavour U uses.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_u.hpp"

namespace swe {

constexpr int kSplitXSeg = 16;

// x-sweep: one lane per row y in [0, ny+1] (64 rows per wave), marching x
// over a segment; stores sx/su = L(x|x+1) + R(x-1|x) for interior cells and
// reduces the maximum over every x-edge (ghost rows included).
template <int S>
__global__ __launch_bounds__(256) void sweep_s_x_kernel(StateView v, SumsView sums, DeviceCtl* ctl, int par) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int nyt = (v.ny + 2 + 63) / 64;
    const int nxs = (v.nx + 1 + kSplitXSeg - 1) / kSplitXSeg;  // edge columns 0..nx
    if (wave >= nyt * nxs) return;
    const int yt = wave % nyt;
    const int xs = wave / nyt;
    const int y = 64 * yt + lane;
    const bool rowValid = y <= v.ny + 1;
    const int yc = rowValid ? y : v.ny + 1;
    const int e0 = xs * kSplitXSeg;                       // first edge (e | e+1)
    const int e1 = min(e0 + kSplitXSeg - 1, v.nx);        // last edge
    size_t o = idx(v, e0, yc);
    float ch = v.h[o], chu = v.hu[o], cb = v.b[o];
    float carryH = 0.f, carryHu = 0.f;  // R of edge (e-1 | e), valid from the second edge on
    if (e0 >= 1) {
        const size_t p = o - v.pitch;
        const dev::EdgeOut e = dev::solve_raw<S>(v.h[p], ch, v.hu[p], chu, v.b[p], cb);
        carryH = e.hR;
        carryHu = e.huR;
    }
    float smaxv = 0.f;
    for (int e = e0; e <= e1; ++e) {
        o += v.pitch;
        const float nh = v.h[o], nhu = v.hu[o], nb = v.b[o];
        const dev::EdgeOut ev = dev::solve_raw<S>(ch, nh, chu, nhu, cb, nb);
        if (rowValid) smaxv = dev::smax(smaxv, ev.speed);
        // cell e is complete: L from edge (e|e+1), R from edge (e-1|e)
        if (e >= 1 && rowValid && y >= 1 && y <= v.ny) {
            const size_t c = idx(v, e, y);
            sums.sx[c] = ev.hL + carryH;
            sums.su[c] = ev.huL + carryHu;
        }
        carryH = ev.hR;
        carryHu = ev.huR;
        ch = nh;
        chu = nhu;
        cb = nb;
    }
    smaxv = wave_max(smaxv);
    if (lane == 0) atomicMax(&ctl->smax[par], __float_as_uint(smaxv));
}

// y-sweep on the old state: strip decomposition as in flavour U (62 owned
// rows per wave, shuffles for the upper neighbour); stores sy/sv = B + A.
template <int S>
__global__ __launch_bounds__(256) void sweep_s_y_kernel(StateView v, SumsView sums) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    const int nxs = (v.nx + kSplitXSeg - 1) / kSplitXSeg;
    if (wave >= nyt * nxs) return;
    const int yt = wave % nyt;
    const int xs = wave / nyt;
    const int y = kStripRows * yt + lane;
    const bool rowValid = y <= v.ny + 1;
    const int yc = rowValid ? y : v.ny + 1;
    const bool owned = (lane >= 1) && (lane <= kStripRows) && (y <= v.ny);
    const int x0 = 1 + xs * kSplitXSeg;
    const int x1 = min(x0 + kSplitXSeg - 1, v.nx);
    for (int x = x0; x <= x1; ++x) {
        const size_t o = idx(v, x, yc);
        const float ch = v.h[o], chv = v.hv[o], cb = v.b[o];
        const float uh = __shfl_down(ch, 1, 64);
        const float uhv = __shfl_down(chv, 1, 64);
        const float ub = __shfl_down(cb, 1, 64);
        const dev::EdgeOut eh = dev::solve_raw<S>(ch, uh, chv, uhv, cb, ub);
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);
        if (owned) {
            const size_t c = idx(v, x, y);
            sums.sy[c] = eh.hL + aH;
            sums.sv[c] = eh.huL + aHv;
        }
    }
}

}  // namespace swe
