// kernels_u.hpp -- flavour-U step kernels (every edge from one state, unsplit
// update with dry clamp), the step of all of the reference's parallel blocks
// (src/blocks/SWE_DimensionalSplittingMpi.cpp:383-518).
//
// Data layout in HBM: each field is one column-major array like the
// reference's Float2D (a[x][y] at x*pitch + y, y contiguous), so a wavefront
// reads 64 consecutive y at one x: one 256-byte coalesced load per field.
//
// Work decomposition ("strip"): one wavefront owns 62 consecutive rows
// y0..y0+61 (lanes 1..62; lane 0 and lane 63 hold the rows just below and
// above) and marches along x over a segment of XS columns.  At column x a
// lane holds cell (x, y) and loads cell (x+1, y):
//   * the vertical edge (x | x+1) is solved in-lane; its right-going part is
//     carried in registers to the next column;
//   * the horizontal edge (y | y+1) takes the upper cell from lane+1 with a
//     cross-lane shuffle; the up-going part of edge (y-1 | y) comes back from
//     lane-1 the same way.
// Each interior edge is solved once per strip (plus one halo x-edge per
// segment), net updates never leave registers, and the CFL maximum is
// reduced in-wave before one atomic per wave.
//
// Modes:
//   MODE_MAX    edges -> max speed only (first half of the two-pass step);
//   MODE_STORE  edges -> per-cell net-update sums (4 floats) + max speed;
//   MODE_UPDATE edges -> update with dt from the device max, written to the
//               other state buffer (second half of the two-pass step).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "mlp.hpp"
#include "solvers_dev.hpp"
#include "swe_common.h"

namespace swe {

enum SweepMode : int { MODE_MAX = 0, MODE_STORE = 1, MODE_UPDATE = 2 };

constexpr int kStripRows = 62;  // owned rows per wavefront
constexpr int kWavesPerGroup = 4;
constexpr int kPlanGroup = 1024;  // wave segments per workgroup of the quiet-tile plan
constexpr int kPlanPieces = 8;    // list entries per segment at most (32 columns in runs of 4)

// ---- strip tiles of the batched engine --------------------------------------
// The strips of kStripRows rows below a block's last strip are cut into
// column segments of XS columns, one wave per (strip, segment).  The last
// strip (the tail, R = ny - kStripRows * (nyt - 1) rows) follows: when R is
// small, its inner segments (full XS columns, away from the left and right
// boundary columns) share waves, `pack` segments per wave, each on a group of
// R + 2 lanes (R rows with a halo lane on either side) laid side by side.
// Every lane computes what the lane of an unpacked strip computes for its row
// and column, so the packing does not change any result.  (The single-block
// kernels keep one wave per (strip, segment): on a large block the tail is a
// small share of the waves, and its clamped lanes make those waves cheap.)
struct StripTiling {
    int nyt, nxs, pack, items;
};

__host__ __device__ inline StripTiling strip_tiling(int nx, int ny, int xs) {
    StripTiling g;
    g.nyt = (ny + kStripRows - 1) / kStripRows;
    g.nxs = (nx + xs - 1) / xs;
    const int perWave = 64 / (ny - kStripRows * (g.nyt - 1) + 2);
    g.pack = (perWave >= 2 && g.nxs >= 4) ? perWave : 1;
    g.items = (g.nyt - 1) * g.nxs + (g.pack == 1 ? g.nxs : 2 + (g.nxs - 2 + g.pack - 1) / g.pack);
    return g;
}

struct TileLanes {
    int yt;            // strip (uniform)
    int xs0;           // first column segment of the item (uniform)
    int xs;            // column segment of this lane's group
    int r;             // row of the lane in the strip: 0 and R+1 are halo rows; 63 on idle lanes
    bool first, last;  // the item's segments are the block's first / last (uniform)
};

// item `tile` of a block's strip tiling: order (strip, segment) with the
// strip fastest below the tail, then the tail's first and last segments,
// then its packed inner segments
__device__ __forceinline__ TileLanes tile_lanes(int nyt, int nxs, int pack, int ny, int tile, int lane) {
    TileLanes t;
    const int body = (nyt - 1) * nxs;
    int lanesPerGroup = 64, groups = 1;
    if (tile < body) {
        t.yt = tile % (nyt - 1);
        t.xs0 = tile / (nyt - 1);
    } else {
        const int u = tile - body;
        t.yt = nyt - 1;
        if (pack == 1) {
            t.xs0 = u;
        } else if (u < 2) {
            t.xs0 = u ? nxs - 1 : 0;
        } else {
            t.xs0 = 1 + (u - 2) * pack;
            groups = min(pack, nxs - 1 - t.xs0);
            lanesPerGroup = ny - kStripRows * t.yt + 2;
        }
    }
    const int g = lane / lanesPerGroup;
    t.xs = t.xs0 + g;
    t.r = lane - g * lanesPerGroup;
    if (g >= groups) {
        t.xs = t.xs0;
        t.r = 63;
    }
    t.first = t.xs0 == 0;
    t.last = t.xs0 == nxs - 1;
    return t;
}

// columns of every segment of the item: XS, fewer for the last
__device__ __forceinline__ int tile_columns(const TileLanes& t, int nx, int xs) {
    return min(xs, nx - t.xs0 * xs);
}

struct StateView {
    float* h;
    float* hu;
    float* hv;
    const float* b;
    int nx, ny, pitch;
    float dx, dy;
};

struct SumsView {  // per-cell net-update sums, same layout as the state
    float* sx;     // h, x-edges:  R(x-1|x) + L(x|x+1)
    float* sy;     // h, y-edges:  A(y-1|y) + B(y|y+1)
    float* su;     // hu, x-edges
    float* sv;     // hv, y-edges
};

struct BoundaryInfo {
    int type[4];  // BoundaryType per side
};

// Device control block of one SWE block: the CFL maxima (two, used
// alternately so that a step can clear the next one without a memset),
// the current dt and simulated time and a per-step dt log.
struct DeviceCtl {
    unsigned int smax[3];  // float bits of non-negative speeds: unsigned order == float order
    float dt;
    float t;
    int steps;
    int pad[2];
    const unsigned* mlp;  // weight fragments of the learned middle state (solver SOLVER_AUGRIE_MLP)
    unsigned int defer[3];  // edges queued by the deferring sweep, per CFL slot (SOLVER_AUGRIE_DEFER)
    unsigned int busy[3];   // busy record tiles found by the quiet-tile plan of each CFL slot's launch
    // coherent host ring the readback kernel writes (kept here rather than
    // passed as a kernel argument, so the launch is an ordinary one)
    unsigned long long* ring;
};

// Which CFL-maximum slot a launch reads or accumulates into, and which slot
// it clears for a later step (the last reader of that slot has finished).
struct Slots {
    int use;
    int clear;
};

__device__ __forceinline__ size_t idx(const StateView& v, int x, int y) { return (size_t)x * v.pitch + y; }

// dt rule of flavour U (SWE_DimensionalSplittingMpi.cpp:448-456)
__device__ __forceinline__ float timestep_from_max(float s, float dx, float dy) {
    if (s > kMinSpeed) {
        float dt = dev::smin(dx / s, dy / s);
        dt *= kCfl;
        return dt;
    }
    return FLT_MAX;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dev::smax(v, __shfl_xor(v, off, 64));
    return v;
}

// Deferred edges (SOLVER_AUGRIE_DEFER).  An entry is (x, y) of the edge's left
// or lower cell, with kDeferY set in .x for the y-edge (y|y+1); the x-edge is
// (x|x+1).  The batched engine keeps the block index in bits kDeferBlockShift
// and up of .x (blocks < 2^11, x < 2^20).  Each edge is queued once, by the
// lane that owns it: x-edges by the wave whose segment holds x (edge (0|1) by
// the first segment), y-edges by the lane of row y (edge (0|1) by lane 0 of
// the first strip); rows 1..ny only for x-edges.  A deferred edge contributes
// nothing in the sweep: its parts are -0 (x + -0 == x for every x, zeros of
// both signs included, so the sweep stores exactly the other term of each
// two-term sum) and its speed 0.  A second launch adds its parts to the sums
// and its speed to the CFL maximum; IEEE addition is commutative, so every sum
// ends bit for bit as when the edge is solved in the sweep.
constexpr unsigned kDeferY = 0x80000000u;
constexpr int kDeferBlockShift = 20;

__device__ __forceinline__ dev::EdgeOut defer_edge(const dev::EdgeOut& e, bool owns, uint2 entry, uint2* q,
                                                   unsigned* count) {
    const bool d = e.speed < 0.f;
    const unsigned long long m = __ballot(d && owns);
    if (m) {  // uniform over the wave; every lane is active here
        unsigned base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(count, (unsigned)__popcll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        const unsigned below =
            __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (d && owns) q[base + below] = entry;
    }
    return d ? dev::EdgeOut{-0.f, -0.f, -0.f, -0.f, 0.f} : e;
}

// True for the ghost cells that the fused boundary write refreshes after an
// update: corners, and ghost lines of OUTFLOW/WALL sides.
__device__ __forceinline__ bool bc_handles(const StateView& v, const BoundaryInfo& bi, int gx, int gy) {
    auto ow = [](int t) { return t == OUTFLOW || t == WALL; };
    const bool cx = (gx == 0 || gx == v.nx + 1);
    const bool cy = (gy == 0 || gy == v.ny + 1);
    if (cx && cy) return true;
    if (gx == 0) return ow(bi.type[BND_LEFT]);
    if (gx == v.nx + 1) return ow(bi.type[BND_RIGHT]);
    if (gy == 0) return ow(bi.type[BND_BOTTOM]);
    if (gy == v.ny + 1) return ow(bi.type[BND_TOP]);
    return false;
}

// Boundary ghost values of cell (gx, gy) derived from the interior after an
// update: OUTFLOW copies, WALL negates the normal momentum
// (SWE_Block.hh:757-911).  Writes only sides whose type is OUTFLOW/WALL and
// the four corners.
__device__ __forceinline__ void write_boundary_ghosts(const StateView& v, const BoundaryInfo& bi, int x, int y,
                                                      float h, float hu, float hv, float* oh, float* ohu,
                                                      float* ohv) {
    const int nx = v.nx, ny = v.ny;
    if (x == 1) {
        const int t = bi.type[BND_LEFT];
        if (t == OUTFLOW || t == WALL) {
            const size_t g = (size_t)0 * v.pitch + y;
            oh[g] = h;
            ohu[g] = (t == WALL) ? -hu : hu;
            ohv[g] = hv;
        }
    }
    if (x == nx) {
        const int t = bi.type[BND_RIGHT];
        if (t == OUTFLOW || t == WALL) {
            const size_t g = (size_t)(nx + 1) * v.pitch + y;
            oh[g] = h;
            ohu[g] = (t == WALL) ? -hu : hu;
            ohv[g] = hv;
        }
    }
    if (y == 1) {
        const int t = bi.type[BND_BOTTOM];
        if (t == OUTFLOW || t == WALL) {
            const size_t g = (size_t)x * v.pitch + 0;
            oh[g] = h;
            ohu[g] = hu;
            ohv[g] = (t == WALL) ? -hv : hv;
        }
    }
    if (y == ny) {
        const int t = bi.type[BND_TOP];
        if (t == OUTFLOW || t == WALL) {
            const size_t g = (size_t)x * v.pitch + ny + 1;
            oh[g] = h;
            ohu[g] = hu;
            ohv[g] = (t == WALL) ? -hv : hv;
        }
    }
    // corners take the diagonal interior cell whatever the side types
    if (x == 1 && y == 1) {
        oh[0] = h;
        ohu[0] = hu;
        ohv[0] = hv;
    }
    if (x == 1 && y == ny) {
        const size_t g = (size_t)0 * v.pitch + ny + 1;
        oh[g] = h;
        ohu[g] = hu;
        ohv[g] = hv;
    }
    if (x == nx && y == 1) {
        const size_t g = (size_t)(nx + 1) * v.pitch + 0;
        oh[g] = h;
        ohu[g] = hu;
        ohv[g] = hv;
    }
    if (x == nx && y == ny) {
        const size_t g = (size_t)(nx + 1) * v.pitch + ny + 1;
        oh[g] = h;
        ohu[g] = hu;
        ohv[g] = hv;
    }
}

// Flavour-U update of one cell with the dry clamp
// (SWE_DimensionalSplittingMpi.cpp:498-515).
struct Cell {
    float h, hu, hv;
};
__device__ __forceinline__ Cell update_cell(float h, float hu, float hv, float dtdx, float dtdy, float sx, float sy,
                                            float su, float sv) {
    Cell c;
    c.h = h - (dtdx * sx + dtdy * sy);
    c.hu = hu - dtdx * su;
    c.hv = hv - dtdy * sv;
    if (c.h < 0.f) {
        c.h = 0.f;
        c.hu = 0.f;
        c.hv = 0.f;
    } else if (c.h < kClampTol) {
        c.hu = 0.f;
        c.hv = 0.f;
    }
    return c;
}

// One loaded cell: state plus the quantities its edges share.
struct CellQ {
    float h, hu, hv, b, u, v, c, sq, inv;
    __device__ __forceinline__ dev::Side xside() const { return dev::Side{h, hu, b, u, c, sq, inv}; }
    __device__ __forceinline__ dev::Side yside() const { return dev::Side{h, hv, b, v, c, sq, inv}; }
    // the y-side of the cell held by lane + 1 (the cell above), by shuffle
    __device__ __forceinline__ dev::Side yside_from_lane_above() const {
        dev::Side s;
        s.h = __shfl_down(h, 1, 64);
        s.hu = __shfl_down(hv, 1, 64);
        s.b = __shfl_down(b, 1, 64);
        s.u = __shfl_down(v, 1, 64);
        s.c = __shfl_down(c, 1, 64);
        s.sq = __shfl_down(sq, 1, 64);
        s.inv = __shfl_down(inv, 1, 64);
        return s;
    }
};

__device__ __forceinline__ CellQ cell_q(float h, float hu, float hv, float b) {
    CellQ q;
    q.h = h;
    q.hu = hu;
    q.hv = hv;
    q.b = b;
    q.inv = 1.f / h;
    q.u = hu * q.inv;
    q.v = hv * q.inv;
    q.sq = sqrtf(h);
    q.c = kSqrtGravity * q.sq;
    return q;
}

// Strip sweep.  One wave per (y-tile, x-segment); 4 waves per workgroup.
// dtFixed >= 0 forces dt (the per-call API passes the host's maxTimestep);
// otherwise dt comes from ctl->smax[par].
// The max and store sweeps hold the fused kernel's occupancy target; the
// update sweep of the recompute strategy keeps its natural register use
// (at seven waves it would spill).
template <int S, int MODE, int XS, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == MODE_UPDATE ? 1 : fused_waves_per_simd(S),
                                                                     8))) void sweep_u_kernel(
    StateView v, SumsView sums, float* hn, float* hun, float* hvn, BoundaryInfo bi, DeviceCtl* ctl, Slots sl,
    float dtFixed, int fuseBc, float* dtLog, int logIdx) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
    const int nyt = (v.ny + kStripRows - 1) / kStripRows;
    const int nxs = (v.nx + XS - 1) / XS;
    if (wave >= nyt * nxs) return;  // uniform per wave
    // the Newton loop's first iteration from the bracket, as in the fused kernel
    const dev::SolverCtx<S, true> solver(S == SOLVER_AUGRIE_MLP ? ctl->mlp : nullptr);
    const int yt = wave % nyt;
    const int xs = wave / nyt;
    const int y = kStripRows * yt + lane;  // rows y0-1 .. y0+62 with y0 = 62*yt + 1
    const bool rowValid = y <= v.ny + 1;
    const int yc = rowValid ? y : v.ny + 1;
    const bool owned = (lane >= 1) && (lane <= kStripRows) && (y <= v.ny);
    const bool rowInterior = (y >= 1) && (y <= v.ny);
    const bool hEdgeValid = (lane <= kStripRows) && (y <= v.ny);  // edge (y|y+1), y in [0, ny]
    const int x0 = 1 + xs * XS;
    const int x1 = min(x0 + XS - 1, v.nx);

    float dtdx = 0.f, dtdy = 0.f, dt = 0.f;
    if (MODE == MODE_UPDATE) {
        dt = dtFixed >= 0.f ? dtFixed : timestep_from_max(__uint_as_float(ctl->smax[sl.use]), v.dx, v.dy);
        dtdx = dt / v.dx;
        dtdy = dt / v.dy;
    }

    const float* H = v.h;
    const float* HU = v.hu;
    const float* HV = v.hv;
    const float* B = v.b;
    size_t o = idx(v, x0 - 1, yc);
    float ch = H[o], chu = HU[o], chv = HV[o], cb = B[o];
    o += v.pitch;
    float nh = H[o], nhu = HU[o], nhv = HV[o], nb = B[o];
    // HLLE, f-wave and AugRie load column x+1 one column ahead (a software
    // pipeline: the loads' latency overlaps column x's edges); past x1 + 1
    // the loads stay on column nx + 1 and are not used.
    constexpr bool kAhead = S == SOLVER_HLLE || S == SOLVER_FWAVE || S == SOLVER_AUGRIE;
    float ah = 0.f, ahu = 0.f, ahv = 0.f, ab = 0.f;
    if constexpr (kAhead) {
        o += v.pitch;  // column x0 + 1 <= nx + 1
        ah = H[o];
        ahu = HU[o];
        ahv = HV[o];
        ab = B[o];
    }

    // In MODE_UPDATE the new buffer must end up a complete state: ghost cells
    // not rewritten by the fused boundary conditions are carried over.  Each
    // ghost row is carried by exactly one lane: row 0 by lane 0 of tile 0,
    // row ny+1 by the tile where it is lane >= 1, other rows by their owner.
    const bool carryRow = owned || (y == 0 && lane == 0) || (y == v.ny + 1 && lane >= 1);
    if (MODE == MODE_UPDATE && x0 == 1 && carryRow && !(fuseBc && bc_handles(v, bi, 0, y))) {
        const size_t g = idx(v, 0, y);
        hn[g] = ch;
        hun[g] = chu;
        hvn[g] = chv;
    }

    // per-cell quantities (velocities, celerity, sqrt h) are computed once
    // when a cell is loaded and reused by its four edges
    CellQ cq = cell_q(ch, chu, chv, cb);
    CellQ nq = cell_q(nh, nhu, nhv, nb);

    // halo vertical edge (x0-1 | x0)
    dev::EdgeOut e = solver.solve(cq.xside(), nq.xside());
    float carryH = e.hR, carryHu = e.huR;
    float smaxv = rowInterior ? e.speed : 0.f;

    for (int x = x0; x <= x1; ++x) {
        ch = nh;
        chu = nhu;
        chv = nhv;
        cb = nb;
        cq = nq;
        if constexpr (kAhead) {
            nh = ah;
            nhu = ahu;
            nhv = ahv;
            nb = ab;
            if (x + 2 <= v.nx + 1) o += v.pitch;  // column x+2, else stay on nx+1
            ah = H[o];
            ahu = HU[o];
            ahv = HV[o];
            ab = B[o];
        } else {
            o += v.pitch;  // column x+1 (<= nx+1)
            nh = H[o];
            nhu = HU[o];
            nhv = HV[o];
            nb = B[o];
        }
        nq = cell_q(nh, nhu, nhv, nb);

        // vertical edge (x | x+1)
        const dev::EdgeOut ev = solver.solve(cq.xside(), nq.xside());
        if (rowInterior) smaxv = dev::smax(smaxv, ev.speed);

        // horizontal edge (y | y+1): upper cell from lane+1
        const dev::EdgeOut eh = solver.solve(cq.yside(), cq.yside_from_lane_above());
        if (hEdgeValid) smaxv = dev::smax(smaxv, eh.speed);
        // up-going part of edge (y-1 | y) from lane-1
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);

        const float sx = carryH + ev.hL;
        const float su = carryHu + ev.huL;
        const float sy = aH + eh.hL;
        const float sv = aHv + eh.huL;
        carryH = ev.hR;
        carryHu = ev.huR;

        if (MODE == MODE_STORE && owned) {
            const size_t c = idx(v, x, y);
            sums.sx[c] = sx;
            sums.sy[c] = sy;
            sums.su[c] = su;
            sums.sv[c] = sv;
        }
        if (MODE == MODE_UPDATE) {
            const size_t c = idx(v, x, y);
            if (owned) {
                const Cell nc = update_cell(ch, chu, chv, dtdx, dtdy, sx, sy, su, sv);
                hn[c] = nc.h;
                hun[c] = nc.hu;
                hvn[c] = nc.hv;
                if (fuseBc) write_boundary_ghosts(v, bi, x, y, nc.h, nc.hu, nc.hv, hn, hun, hvn);
            } else if (carryRow && (y == 0 || y == v.ny + 1) && !(fuseBc && bc_handles(v, bi, x, y))) {
                hn[c] = ch;  // carry ghost rows into the new buffer
                hun[c] = chu;
                hvn[c] = chv;
            }
        }
    }
    if (MODE == MODE_UPDATE && x1 == v.nx && carryRow && !(fuseBc && bc_handles(v, bi, v.nx + 1, y))) {
        const size_t g = idx(v, v.nx + 1, y);
        hn[g] = nh;
        hun[g] = nhu;
        hvn[g] = nhv;
    }

    if (MODE != MODE_UPDATE) {
        smaxv = wave_max(smaxv);
        if (lane == 0) atomicMax(&ctl->smax[sl.use], __float_as_uint(smaxv));
    } else {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->smax[sl.clear] = 0u;  // a later step's maximum starts from zero
            if (dtFixed < 0.f) {
                ctl->dt = dt;
                ctl->t += dt;
                ctl->steps += 1;
                if (dtLog) dtLog[logIdx] = dt;
            }
        }
    }
}

// Apply stored net-update sums in place (second half of the store strategy).
template <int XS, int V>
__global__ __launch_bounds__(256) void apply_u_kernel(StateView v, SumsView sums, BoundaryInfo bi, DeviceCtl* ctl,
                                                       Slots sl, float dtFixed, int fuseBc, float* dtLog,
                                                       int logIdx) {
    const int y = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
    const int xseg = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int x0 = 1 + xseg * XS;
    const float dt = dtFixed >= 0.f ? dtFixed : timestep_from_max(__uint_as_float(ctl->smax[sl.use]), v.dx, v.dy);
    const float dtdx = dt / v.dx;
    const float dtdy = dt / v.dy;
    if (y <= v.ny && x0 <= v.nx) {
        const int x1 = min(x0 + XS - 1, v.nx);
        for (int x = x0; x <= x1; ++x) {
            const size_t c = idx(v, x, y);
            const Cell nc = update_cell(v.h[c], v.hu[c], v.hv[c], dtdx, dtdy, sums.sx[c], sums.sy[c], sums.su[c],
                                        sums.sv[c]);
            v.h[c] = nc.h;
            v.hu[c] = nc.hu;
            v.hv[c] = nc.hv;
            if (fuseBc) write_boundary_ghosts(v, bi, x, y, nc.h, nc.hu, nc.hv, v.h, v.hu, v.hv);
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ctl->smax[sl.clear] = 0u;
        if (dtFixed < 0.f) {
            ctl->dt = dt;
            ctl->t += dt;
            ctl->steps += 1;
            if (dtLog) dtLog[logIdx] = dt;
        }
    }
}

}  // namespace swe
