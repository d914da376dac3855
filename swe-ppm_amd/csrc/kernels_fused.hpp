// kernels_fused.hpp -- flavour-U step with the previous step's update fused
// into the load ("apply on load"): one launch per time step.
//
// Launch k (k >= 1) reads Q(k-1) and the net-update sums S(k-1) of step k-1,
// forms Q(k) = update(Q(k-1), S(k-1), dt(k-1)) cell by cell as it loads them
// (dry clamp included), resolves the ghost cells of Q(k) the way
// setGhostLayer would (OUTFLOW copies, WALL negates the normal momentum,
// corners take the diagonal interior cell, other sides keep their ghosts),
// writes Q(k) to the other state buffer, then solves step k's edges from
// Q(k) and writes S(k) and the CFL maximum M(k).  dt(k-1) comes from M(k-1).
// A run of N steps is: sweep(STORE) on Q(0), N-1 fused launches, one apply.
//
// The CFL maxima rotate through three slots: launch k reads slot (k-1)%3,
// accumulates into slot k%3 and clears slot (k+1)%3, which the previous
// launch was the last to read -- so no memset is ever queued between steps.
//
// Work decomposition and edge order are those of the strip sweep
// (kernels_u.hpp), so S(k) and M(k) are bit-identical to the two-kernel step.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_u.hpp"

namespace swe {

struct FusedIO {
    const float* h;   // Q(k-1)
    const float* hu;
    const float* hv;
    const float* b;
    float* hn;  // Q(k)
    float* hun;
    float* hvn;
    SumsView sIn;   // S(k-1)
    SumsView sOut;  // S(k)
    uint2* deferQ = nullptr;  // queue of deferred edges (SOLVER_AUGRIE_DEFER only)
};


struct Q3 {
    float h, hu, hv;
};

__device__ __forceinline__ bool is_ow(int t) { return t == OUTFLOW || t == WALL; }

template <int S, int XS, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fused_waves_per_simd(S), 8))) void fused_u_kernel(
    StateView v, FusedIO io, BoundaryInfo bi, DeviceCtl* ctl, int slotIn, int slotOut, int slotClear, float* dtLog,
    int logIdx) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    const int nxs = (v.nx + XS - 1) / XS;
    const int nx = v.nx, ny = v.ny;

    const float dt = timestep_from_max(__uint_as_float(ctl->smax[slotIn]), v.dx, v.dy);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->smax[slotClear] = 0u;
        ctl->defer[slotClear] = 0u;
        ctl->dt = dt;
        ctl->t += dt;
        ctl->steps += 1;
        if (dtLog) dtLog[logIdx] = dt;
    }
    if (wave >= nyt * nxs) return;  // uniform per wave
    // the Newton loop's first iteration from the bracket (middle_state<true>):
    // measured 3% faster at seven waves per SIMD, same results
    const dev::SolverCtx<S, true> solver(S == SOLVER_AUGRIE_MLP ? ctl->mlp : nullptr);

    const float dtdx = dt / v.dx;
    const float dtdy = dt / v.dy;
    const int yt = wave % nyt;
    const int xs = wave / nyt;
    const int y = kStripRows * yt + lane;
    const bool rowValid = y <= ny + 1;
    const int yc = rowValid ? y : ny + 1;
    const bool owned = (lane >= 1) && (lane <= kStripRows) && (y <= ny);
    const bool rowInterior = (y >= 1) && (y <= ny);
    const bool hEdgeValid = (lane <= kStripRows) && (y <= ny);
    const bool rowBottom = (y == 0);
    const bool rowTop = (y == ny + 1);
    const bool carryRow = owned || (rowBottom && lane == 0) || (rowTop && lane >= 1);
    const int x0 = 1 + xs * XS;
    const int x1 = min(x0 + XS - 1, nx);
    const int tB = bi.type[BND_BOTTOM], tT = bi.type[BND_TOP], tL = bi.type[BND_LEFT], tR = bi.type[BND_RIGHT];

    // Q(k) of the cell (x, yc): interior cells updated, ghosts as stored
    auto load_updated = [&](int x) -> Q3 {
        const size_t c = idx(v, x, yc);
        Q3 q{io.h[c], io.hu[c], io.hv[c]};
        if (x >= 1 && x <= nx && rowInterior) {
            const Cell u = update_cell(q.h, q.hu, q.hv, dtdx, dtdy, io.sIn.sx[c], io.sIn.sy[c], io.sIn.su[c],
                                       io.sIn.sv[c]);
            q = Q3{u.h, u.hu, u.hv};
        }
        return q;
    };
    // ghost rows of an interior column: the boundary value of the row next to
    // it (lane+1 for the bottom row, lane-1 for the top row)
    auto fix_rows = [&](Q3 q) -> Q3 {
        const float dh = __shfl_down(q.h, 1, 64), dhu = __shfl_down(q.hu, 1, 64), dhv = __shfl_down(q.hv, 1, 64);
        const float uh = __shfl_up(q.h, 1, 64), uhu = __shfl_up(q.hu, 1, 64), uhv = __shfl_up(q.hv, 1, 64);
        if (rowBottom && is_ow(tB)) q = Q3{dh, dhu, (tB == WALL) ? -dhv : dhv};
        if (rowTop && is_ow(tT)) q = Q3{uh, uhu, (tT == WALL) ? -uhv : uhv};
        return q;
    };
    // a ghost column next to the interior column value `in` (same lane);
    // corners take the diagonal interior cell (from the neighbouring lane)
    auto fix_column = [&](Q3 raw, Q3 in, int type) -> Q3 {
        const float dh = __shfl_down(in.h, 1, 64), dhu = __shfl_down(in.hu, 1, 64), dhv = __shfl_down(in.hv, 1, 64);
        const float uh = __shfl_up(in.h, 1, 64), uhu = __shfl_up(in.hu, 1, 64), uhv = __shfl_up(in.hv, 1, 64);
        Q3 q = raw;
        if (rowInterior && is_ow(type)) q = Q3{in.h, (type == WALL) ? -in.hu : in.hu, in.hv};
        if (rowBottom) q = Q3{dh, dhu, dhv};
        if (rowTop) q = Q3{uh, uhu, uhv};
        return q;
    };
    auto store = [&](int x, const Q3& q) {
        const size_t c = idx(v, x, y);
        io.hn[c] = q.h;
        io.hun[c] = q.hu;
        io.hvn[c] = q.hv;
    };

    const float* B = v.b;
    // columns x0-1 and x0
    Q3 cur = fix_rows(load_updated(x0));
    Q3 prev = load_updated(x0 - 1);
    if (x0 == 1) prev = fix_column(prev, cur, tL);
    if (x0 == 1 && carryRow) store(0, prev);
    float cb = B[idx(v, x0 - 1, yc)];
    float nb = B[idx(v, x0, yc)];

    CellQ pq = cell_q(prev.h, prev.hu, prev.hv, cb);
    CellQ cq = cell_q(cur.h, cur.hu, cur.hv, nb);
    dev::EdgeOut e = solver.solve(pq.xside(), cq.xside());
    if constexpr (S == SOLVER_AUGRIE_DEFER)
        e = defer_edge(e, owned && x0 == 1, make_uint2(0u, (unsigned)y), io.deferQ, &ctl->defer[slotOut]);
    float carryH = e.hR, carryHu = e.huR;
    float smaxv = rowInterior ? e.speed : 0.f;

    for (int x = x0; x <= x1; ++x) {
        // column x+1 of Q(k)
        Q3 nxt = fix_rows(load_updated(x + 1));
        if (x + 1 == nx + 1) nxt = fix_column(nxt, cur, tR);
        const float bn = B[idx(v, x + 1, yc)];
        const CellQ nq = cell_q(nxt.h, nxt.hu, nxt.hv, bn);

        if (carryRow) store(x, cur);
        if (x + 1 == nx + 1 && carryRow) store(nx + 1, nxt);

        dev::EdgeOut ev = solver.solve(cq.xside(), nq.xside());
        if constexpr (S == SOLVER_AUGRIE_DEFER)
            ev = defer_edge(ev, owned, make_uint2((unsigned)x, (unsigned)y), io.deferQ, &ctl->defer[slotOut]);
        if (rowInterior) smaxv = dev::smax(smaxv, ev.speed);
        dev::EdgeOut eh = solver.solve(cq.yside(), cq.yside_from_lane_above());
        if constexpr (S == SOLVER_AUGRIE_DEFER)
            eh = defer_edge(eh, owned || y == 0, make_uint2((unsigned)x | kDeferY, (unsigned)y), io.deferQ,
                            &ctl->defer[slotOut]);
        if (hEdgeValid) smaxv = dev::smax(smaxv, eh.speed);
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);
        if (owned) {
            const size_t c = idx(v, x, y);
            io.sOut.sx[c] = carryH + ev.hL;
            io.sOut.su[c] = carryHu + ev.huL;
            io.sOut.sy[c] = aH + eh.hL;
            io.sOut.sv[c] = aHv + eh.huL;
        }
        carryH = ev.hR;
        carryHu = ev.huR;
        cur = nxt;
        cq = nq;
    }

    smaxv = wave_max(smaxv);
    if (lane == 0) atomicMax(&ctl->smax[slotOut], __float_as_uint(smaxv));
}

}  // namespace swe
