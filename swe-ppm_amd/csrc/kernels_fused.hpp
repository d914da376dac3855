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
//
// Quiet tiles.  The quiet record works on tiles of one strip (62 rows) by
// kQuietCols columns.  Launch k records per tile which of its parts had all
// four sums of every owned cell +0 (bit pattern 0) with no deferred edge --
// all its cells, its bottom and top owned rows, its first and last columns --
// and its CFL maximum, speed(k): the maximum over its x-edges on owned rows
// (those right of its columns; the left halo edge too where the tile starts a
// run of computed tiles) and its y-edges.  The tile's results depend
// only on Q(k) over its footprint: its own cells, the last column of its left
// neighbour, the first column of its right one, the top row of the tile below
// and the bottom row of the tile above (x-edge speeds count on owned rows
// only, so the corner cells of the halo feed nothing the tile reports).  If
// those parts were quiet at k-1, then Q(k) = Q(k-1) over the footprint -- an
// update by +0 sums leaves h, hu, hv bit for bit unchanged once the state has
// passed one dry clamp -- so S(k) and speed(k) repeat S(k-1) = +0 and
// speed(k-1).  If the tile was also quiet at k-2, the buffers launch k writes
// (state and sums of step k-2) already hold exactly those values.  Such a
// tile is skipped: it only carries its speed forward and adds it to the CFL
// maximum.  Launches 1 and 2 of a run never skip (their inputs may predate
// the first clamp), so every flag read was written earlier in the same run.
// Results are bit-identical with and without skipping.
//
// A wave still owns a segment of XS columns (a whole number of record
// tiles).  The skip decisions of launch k are taken by quiet_plan_kernel
// between launches k-1 and k: it carries the skipped tiles' speeds and, per
// segment with work left, lists the run of record tiles from its first to its
// last busy one (quiet tiles inside the run are recomputed, to the same
// values).  The list is compact and in domain order, per group of kPlanGroup
// segments.  Launch k then runs in list mode: wave w takes the w-th listed
// run, and waves past the count leave at once.  The busy runs thus take the
// first waves, which the dispatcher spreads over all CUs, instead of sitting
// wherever they lie in the domain (with one wave per segment in domain order,
// the CUs that hold the busy region would set the launch time while the
// others idle); with every tile busy the launch is the one without a plan.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_u.hpp"

namespace swe {

// Parts of a tile whose owned sums the quiet record tracks (one bit each):
// all cells, the bottom and top owned rows and the first and last columns --
// the halo a neighbouring tile reads.
constexpr unsigned kQuietAll = 1u, kQuietBottom = 2u, kQuietTop = 4u, kQuietLeft = 8u, kQuietRight = 16u;
// Record tiles are kQuietCols columns wide; a wave's segment of XS columns
// spans XS / kQuietCols of them.
constexpr int kQuietCols = 4;
// Waves the quiet-tile plan aims to fill with busy record tiles before it
// lengthens the pieces it lists: four per SIMD of the chip.  Measured on the
// 2048^2 islands, the 4096^2 trench and the 4096^2 islands (us per step):
// 1024: 92.9 / - / -; 2048: 85.6 / 92.4 / 169.4; 3072: 81.3 / 89.2 / 166.9;
// 4096: 79.7 / 89.0 / 165.6; 5120: 81.0 / 90.3 / 165.0; 6144: 82.5 / 93.3 /
// 166.1; 8192: 88.3 / - / -.  Shorter pieces give more waves to hide the
// busy tiles' latency but solve more halo edges.
constexpr unsigned kPlanWaves = 4096;

// Lanes the quiet-tile plan gives a wave segment of `fine` record tiles: the
// smallest power of two >= fine, so that a segment's lanes share one wave.
__host__ __device__ constexpr int plan_lanes(int fine) { return fine <= 1 ? 1 : fine <= 2 ? 2 : fine <= 4 ? 4 : 8; }
// Workgroups of the plan (and entries of its count array) for nyt x nxs
// segments of `fine` record tiles.
__host__ __device__ constexpr int plan_groups(int nyt, int nxs, int fine) {
    return (nyt * nxs * plan_lanes(fine) + kPlanGroup - 1) / kPlanGroup;
}

// Which fused kernels keep the quiet record: the augmented solver's, exact
// and deferring, whose still edges give +0 sums (HLLE and f-wave leave the
// round-off of a lake at rest, so their tiles are never quiet; the in-sweep
// network variant is at its three-wave register limit and the device loop
// runs the deferring variant instead), at widths that are whole record tiles.
template <int S>
constexpr bool quiet_record(int xs) {
    return (S == SOLVER_AUGRIE || S == SOLVER_AUGRIE_DEFER) && xs % kQuietCols == 0;
}

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
    // quiet-tile record: quiet[slot * tiles + t], speed[slot * tiles + t] with
    // slot = launch index & 1 (null: no record, no skipping)
    unsigned* quiet = nullptr;
    float* speed = nullptr;
    int step = 0;  // launch index within the run (1, 2, ...)
    // list mode: workgroup g of the plan lists count[g] runs of record tiles
    // to compute at list[g * kPlanGroup * kPlanPieces ...]
    const unsigned* list = nullptr;
    const unsigned* count = nullptr;
    // >= 0: the previous update's dt, given by the caller (the per-call
    // drop-in API: updateUnknowns' dt applied on the next flux pass's load)
    float dtFixed = -1.f;
};

// The list entry of wave w in list mode, or -1 past the last listed entry: a
// wave scan over the group counts finds the group holding the w-th entry.
__device__ __forceinline__ int listed_item(const FusedIO& io, int groups, int w) {
    const int lane = threadIdx.x & 63;
    for (int base = 0; base < groups; base += 64) {  // uniform per wave
        const unsigned c = (base + lane < groups) ? io.count[base + lane] : 0u;
        unsigned incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned o = __shfl_up(incl, off, 64);
            incl += (lane >= off) ? o : 0u;
        }
        const unsigned total = __shfl(incl, 63, 64);
        if ((unsigned)w < total) {
            const unsigned long long m = __ballot(incl > (unsigned)w);
            const int g = __builtin_ctzll(m);
            const unsigned excl = __shfl(incl - c, g, 64);
            return (int)__builtin_amdgcn_readfirstlane(
                io.list[(size_t)(base + g) * kPlanGroup * kPlanPieces + (w - excl)]);
        }
        w -= (int)total;
    }
    return -1;
}

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

    const float dt = io.dtFixed >= 0.f ? io.dtFixed : timestep_from_max(__uint_as_float(ctl->smax[slotIn]), v.dx, v.dy);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->smax[slotClear] = 0u;
        ctl->defer[slotClear] = 0u;
        ctl->busy[slotClear] = 0u;
        ctl->dt = dt;
        ctl->t += dt;
        ctl->steps += 1;
        if (dtLog) dtLog[logIdx] = dt;
    }
    constexpr bool kRecord = quiet_record<S>(XS);
    const int items = nyt * nxs;
    int item = wave, first = 0, ncols = XS;
    if (kRecord && io.list) {
        const int e = listed_item(io, plan_groups(nyt, nxs, XS / kQuietCols), wave);
        if (e < 0) return;  // uniform per wave
        item = e >> 6;
        first = (e >> 3) & 7;
        ncols = ((e & 7) + 1) * kQuietCols;
    }
    if (item >= items) return;  // uniform per wave
    const int yt = item % nyt;
    const int xs = item / nyt;
    const int slotNow = (io.step & 1) * nyt * ((nx + kQuietCols - 1) / kQuietCols);
    // the Newton loop's first iteration from the bracket (middle_state<true>):
    // measured 3% faster at seven waves per SIMD, same results
    const dev::SolverCtx<S, true> solver(S == SOLVER_AUGRIE_MLP ? ctl->mlp : nullptr);

    const float dtdx = dt / v.dx;
    const float dtdy = dt / v.dy;
    const int y = kStripRows * yt + lane;
    const bool rowValid = y <= ny + 1;
    const int yc = rowValid ? y : ny + 1;
    const bool owned = (lane >= 1) && (lane <= kStripRows) && (y <= ny);
    const bool rowInterior = (y >= 1) && (y <= ny);
    const bool hEdgeValid = (lane <= kStripRows) && (y <= ny);
    const bool rowBottom = (y == 0);
    const bool rowTop = (y == ny + 1);
    const bool carryRow = owned || (rowBottom && lane == 0) || (rowTop && lane >= 1);
    const int x0 = 1 + xs * XS + first * kQuietCols;
    const int x1 = min(x0 + ncols - 1, nx);
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
    // record tile of the columns being solved: a deferred edge touching it
    // makes it not quiet; the halo edge (x0-1 | x0) counts for the first
    bool loud = false;
    if constexpr (S == SOLVER_AUGRIE_DEFER) {
        loud = e.speed < 0.f;
        e = defer_edge(e, owned && x0 == 1, make_uint2(0u, (unsigned)y), io.deferQ, &ctl->defer[slotOut]);
    }
    float carryH = e.hR, carryHu = e.huR;
    // x-edges count on owned rows (each interior row is owned by one strip),
    // so a record tile's maximum depends on its footprint only
    float smaxTile = owned ? e.speed : 0.f;
    // the segment's maximum over its record tiles, wave-uniform (float bits:
    // non-negative floats order like unsigned integers)
    unsigned smaxRun = 0u;
    // kQuiet* bits of the record tile's parts whose owned sums were not all +0
    unsigned loudParts = 0u;

    // Column x+1's loads are issued one column ahead (a software pipeline):
    // their latency overlaps column x's edges.  The loads past x1 + 1 read
    // column nx + 1 at most and are not used.  Not for the learned middle
    // state, whose network call leaves no registers for them.
    struct Raw {
        float h, hu, hv, sx, sy, su, sv, b;
    };
    auto load_raw = [&](int x) -> Raw {
        const size_t c = idx(v, min(x, nx + 1), yc);
        return Raw{io.h[c], io.hu[c], io.hv[c], io.sIn.sx[c], io.sIn.sy[c], io.sIn.su[c], io.sIn.sv[c], B[c]};
    };
    constexpr bool kAhead = S != SOLVER_AUGRIE_MLP && S != SOLVER_HYBRID && S != SOLVER_AUGRIE_DEFER;
    Raw ahead{};
    if constexpr (kAhead) ahead = load_raw(x0 + 1);
    for (int x = x0; x <= x1; ++x) {
        // column x+1 of Q(k)
        const Raw r = kAhead ? ahead : load_raw(x + 1);
        if constexpr (kAhead) ahead = load_raw(x + 2);
        Q3 nxt{r.h, r.hu, r.hv};
        if (x + 1 <= nx && rowInterior) {
            const Cell u = update_cell(nxt.h, nxt.hu, nxt.hv, dtdx, dtdy, r.sx, r.sy, r.su, r.sv);
            nxt = Q3{u.h, u.hu, u.hv};
        }
        nxt = fix_rows(nxt);
        if (x + 1 == nx + 1) nxt = fix_column(nxt, cur, tR);
        const CellQ nq = cell_q(nxt.h, nxt.hu, nxt.hv, r.b);

        if (carryRow) store(x, cur);
        if (x + 1 == nx + 1 && carryRow) store(nx + 1, nxt);

        dev::EdgeOut ev = solver.solve(cq.xside(), nq.xside());
        bool loudEdge = false;
        if constexpr (S == SOLVER_AUGRIE_DEFER) {
            loudEdge = ev.speed < 0.f;
            ev = defer_edge(ev, owned, make_uint2((unsigned)x, (unsigned)y), io.deferQ, &ctl->defer[slotOut]);
        }
        if (owned) smaxTile = dev::smax(smaxTile, ev.speed);
        dev::EdgeOut eh = solver.solve(cq.yside(), cq.yside_from_lane_above());
        if constexpr (S == SOLVER_AUGRIE_DEFER) {
            loud = loud || loudEdge || eh.speed < 0.f;
            eh = defer_edge(eh, owned || y == 0, make_uint2((unsigned)x | kDeferY, (unsigned)y), io.deferQ,
                            &ctl->defer[slotOut]);
        }
        if (hEdgeValid) smaxTile = dev::smax(smaxTile, eh.speed);
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);
        if (owned) {
            const size_t c = idx(v, x, y);
            const float sx = carryH + ev.hL, su = carryHu + ev.huL, sy = aH + eh.hL, sv = aHv + eh.huL;
            io.sOut.sx[c] = sx;
            io.sOut.su[c] = su;
            io.sOut.sy[c] = sy;
            io.sOut.sv[c] = sv;
            if constexpr (kRecord)
                if ((__float_as_uint(sx) | __float_as_uint(su) | __float_as_uint(sy) | __float_as_uint(sv)) != 0u)
                    loudParts |= kQuietAll | (lane == 1 ? kQuietBottom : 0u) | (lane == kStripRows ? kQuietTop : 0u) |
                                 (((x - 1) & (kQuietCols - 1)) == 0 ? kQuietLeft : 0u) |
                                 ((x & (kQuietCols - 1)) == 0 || x == nx ? kQuietRight : 0u);
        }
        carryH = ev.hR;
        carryHu = ev.huR;
        cur = nxt;
        cq = nq;
        // the end of a record tile: its parts' flags and its maximum.  A tile
        // reports its right edge (x | x+1) and, first in a run, its left halo
        // edge; the tile of column 1 always starts a run, so every x-edge is
        // in some record, and each record's edges lie in its tile's footprint.
        // A deferred edge at the boundary touches both tiles.
        if (kRecord && ((x & (kQuietCols - 1)) == 0 || x == x1)) {  // uniform per wave
            const float tileMax = wave_max(smaxTile);
            smaxRun = max(smaxRun, __builtin_amdgcn_readfirstlane(__float_as_uint(tileMax)));
            if (io.quiet) {
                unsigned quiet = 0u;
#pragma unroll
                for (unsigned part = 1u; part <= kQuietRight; part <<= 1)
                    if (!__any((loudParts & part) != 0u || loud)) quiet |= part;
                if (lane == 0) {
                    const int rec = slotNow + yt + nyt * ((x - 1) / kQuietCols);
                    io.quiet[rec] = quiet;
                    io.speed[rec] = tileMax;
                }
            }
            smaxTile = 0.f;
            loud = loudEdge;
            loudParts = 0u;
        }
    }

    if constexpr (!kRecord) smaxRun = __float_as_uint(wave_max(smaxTile));
    if (lane == 0) atomicMax(&ctl->smax[slotOut], smaxRun);
}

}  // namespace swe
