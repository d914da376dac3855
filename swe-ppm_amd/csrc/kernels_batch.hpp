// kernels_batch.hpp -- batched flavour-U engine: many blocks of uneven size
// in one device arena, driven by a device-side work queue.
//
// Three uses share it:
//   * ensemble: independent instances, each with its own CFL dt and time;
//   * one domain split into connected blocks with a global dt (the
//     overdecomposed driver without LTS, swe_mpi_overdecomp.cpp:297-307);
//   * multi-rate local time stepping across connected blocks (the driver's
//     LTS loop, swe_mpi_overdecomp.cpp:263-343, with the per-block state
//     machine of SWE_Block.hh:311-562).
//
// One driver iteration is two launches with the fused strategy (sweep,
// round), five with the store strategy (lines, sweep, round, apply, fill).
// Every launch reads its parameters from device memory, so a chunk of
// iterations is captured once in a hipGraph and replayed:
//   lines : setGhostLayer's CONNECT_WITHIN_RANK part
//           (SWE_DimensionalSplittingMPIOverdecomp.cpp:273-304): the
//           neighbour's copy layer straight into the ghost line (global dt),
//           or, for LTS, into the side's receive line when the side receives
//           this iteration, followed by checkAllGhostlayers' copy or
//           interpolation (SWE_Block.hh:375-436).  The previous round wrote
//           each side's job (what to do, from which buffers).  The fused
//           sweep does this step itself: a tile of a block that computes
//           takes its ghost cells from the job as it loads them, and the
//           sides of the blocks that wait are extra items of its queue.  The
//           store strategy runs it as a launch of its own, one workgroup per
//           (block, side, piece of the side).
//   sweep : waves take (block, tile) items from the queue and solve the
//           tile's edges: per-cell net-update sums + per-block CFL maximum;
//           with the fused strategy a block's pending update is applied as
//           its cells are loaded (and, for LTS, a waiting block's ghost-line
//           steps are items too).  By default the fused sweep's grid has one
//           wave per possible item and a wave takes the item of its index
//           (the queue is compact, so the waves past its count leave at
//           once); a persistent grid that pulls items in a loop is the
//           alternative (swe_hip_batch_set_tuning).
//   round : one workgroup: per-block dt (CFL; LTS: getRoundTimestep; global:
//           minimum over blocks), the driver's completion test
//           (hasMaxLocalTimestep), resetStepSizeCounter at the end of a macro
//           step; then, as every block's time is now known, the next
//           iteration's LTS bookkeeping: which sides receive and with what
//           time stamp (receiveGhostLayer, SWE_Block.hh:311-374) and the
//           ghost-layer state transitions of checkAllGhostlayers; the next
//           iteration's ghost-line jobs and work queue (the tiles of the
//           blocks in sync).
//   apply : (store strategy only) waves pull the same items and update their
//           cells, writing the OUTFLOW / WALL ghosts that the next
//           setGhostLayer would write; a one-workgroup fill launch then
//           builds the next queue.
// Persistent grids are sized to the machine, not to the work: each wave
// loops until the queue is drained.  No workgroup ever waits for another
// inside a launch; the launch boundaries order the phases.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels_u.hpp"

namespace swe {

constexpr int kTileBits = 20;  // work item = block << kTileBits | tile
constexpr int kLineTile = (1 << kTileBits) - 4;  // item tile kLineTile + side: the side's ghost-line step
constexpr int kMaxBatchBlocks = 2047;
constexpr int kMaxDivisor = 1024;  // SWE_Block.hh:538 (maxDivisor)

enum BatchMode : int { BATCH_ENSEMBLE = 0, BATCH_GLOBAL = 1, BATCH_LTS = 2 };

struct BatchBlockDesc {
    int nx, ny, pitch;
    float dx, dy;
    int pad0;
    long long off;      // offset of the block's fields in the arena (floats)
    int bnd[4];         // BoundaryType per side
    int nbr[4];         // neighbour block index or -1
    int nyt, ntiles;    // strips of kStripRows rows, work items (tile_lanes)
    int nxs, pack;      // column segments of xs columns, tail segments per packed item
    long long line[4];  // LTS receive lines (h, hu, hv of one side) in the line arena
};

struct BatchBlockState {
    unsigned int smax;  // CFL maximum of this iteration (float bits)
    float maxTimestep;  // dt of the block's last update (pending or applied)
    int stepSize, stepSizeCounter, timestepCounter;
    int gl[4];
    float borderTs[4];
    int notified;
    int active;  // in sync this iteration: computes and updates
    int cur;     // state buffer (0/1) holding the block
    int pend;    // fused strategy: the last update (sums set sidx, dt maxTimestep) is not yet applied
    int sidx;    // sums set of the pending update
    int recv;    // LTS: sides (bit per side) whose receive line takes the neighbour's layer this iteration
    long long updates;  // updates performed
    float t;            // simulated time of an independent block
    float total;        // LTS: lts_total at the end of the last iteration (what neighbours read)
};

// State double-buffered per block (the fused sweep reads one buffer and
// writes the other), bathymetry single, two sets of net-update sums.
struct BatchArena {
    float *h[2], *hu[2], *hv[2], *b;
    float* s[2][4];  // sums sets: sx, sy, su, sv
    float* lines;
};

struct BatchGlobals {
    float T;         // LTS macro step (maxTimestepLocal)
    float duration;  // simulation end (GL_DONE / notifiedLastTimestep)
    int mode;
    int nblocks;
    int count;        // items in the queue
    int head;         // next item of the sweep
    int headApply;    // next item of the apply
    int skip;         // this iteration does nothing (the run's target is reached)
    int stop;         // target reached: later iterations are skipped
    int target;       // units of this run: steps (ensemble, global) or macro steps (LTS)
    int done;         // units completed
    int itersUnit;    // iterations of the current macro step
    float gdt;        // global dt of this iteration
    int logCap;
    long long iterations;  // iterations that did work
    const unsigned* mlp;   // weight fragments of the learned middle state (solver SOLVER_AUGRIE_MLP)
    int fused;             // 1: fused sweep (update applied on the next load), 0: sweep + apply
    int allItems;          // work items of all blocks (the queue of an iteration without LTS)
    unsigned deferCount;   // edges queued by the deferring sweep this iteration (SOLVER_AUGRIE_DEFER)
};

// ---- ghost lines of an iteration ---------------------------------------------
// Fixed geometry of one (block, side), set at create: n == 0 unless the side
// is CONNECT_WITHIN_RANK.
struct GhostGeom {
    long long src;   // arena offset of the neighbour's copy-layer element 1
    long long dst;   // arena offset of the block's ghost-line element 1
    long long line;  // offset of the side's receive line in the line arena (LTS)
    int n, srcStride, dstStride;
    float ndx, ndy;  // the neighbour's cell sizes
};
constexpr int kLineReceive = 1;   // read the neighbour's copy layer this iteration
constexpr int kLineKeep = 2;      // LTS: through the side's receive line
constexpr int kLineActShift = 2;  // ghost action: 0 none, 1 copy, 2 interpolate
// What the lines launch does for one (block, side) in the next iteration,
// written by the round (op == 0: nothing).
struct GhostJob {
    int op;
    int srcBuf, pend, sidx;  // the neighbour's state buffer, pending flag and sums set
    float ndt;               // the neighbour's pending dt
    int dstBuf;              // the block's state buffer
    float w;                 // interpolation weight t_me / t_neighbour
};

// Outputs of the round besides the block states and globals.
struct RoundIO {
    int* items;         // the next iteration's queue
    float* dtLog;       // per completed unit: global dt (0 otherwise)
    int* iterLog;       // per completed unit: driver iterations
    int* hostStat;      // pinned: stop, units done, iterations of the unit
    GhostJob* jobs;     // 4 per block
};

struct BQ3 {
    float h, hu, hv;
};

// One element of a side's ghost-line step: the neighbour's copy-layer value
// (its pending update applied) or the side's receive line; into the receive
// line when the side keeps one and receives (writeLine: this thread owns the
// element), and the new ghost value: the copy, old + (new - old) * w, or the
// ghost's value `old` unchanged.
__device__ __forceinline__ BQ3 ghost_element(const BatchArena& A, const GhostJob& j, const GhostGeom& g, int i,
                                             BQ3 old, bool writeLine) {
    float* L = A.lines + g.line;
    const int n = g.n;
    BQ3 nv;
    if (j.op & kLineReceive) {
        const long long c = g.src + (long long)i * g.srcStride;
        Cell q{A.h[j.srcBuf][c], A.hu[j.srcBuf][c], A.hv[j.srcBuf][c]};
        if (j.pend)
            q = update_cell(q.h, q.hu, q.hv, j.ndt / g.ndx, j.ndt / g.ndy, A.s[j.sidx][0][c], A.s[j.sidx][1][c],
                            A.s[j.sidx][2][c], A.s[j.sidx][3][c]);
        nv = BQ3{q.h, q.hu, q.hv};
        if ((j.op & kLineKeep) && writeLine) {
            L[i] = nv.h;
            L[n + i] = nv.hu;
            L[2 * n + i] = nv.hv;
        }
    } else {
        nv = BQ3{L[i], L[n + i], L[2 * n + i]};
    }
    const int act = j.op >> kLineActShift;
    if (act == 1) return nv;
    if (act == 2)
        return BQ3{old.h + (nv.h - old.h) * j.w, old.hu + (nv.hu - old.hu) * j.w, old.hv + (nv.hv - old.hv) * j.w};
    return old;
}

// A side's ghost-line step done in place by one wave (a waiting block's queue
// item, or a launch of the store strategy): the ghost line of buffer dstBuf.
__device__ __forceinline__ void ghost_line_step(const BatchArena& A, const GhostJob& j, const GhostGeom& g, int i0,
                                                int stride) {
    const int act = j.op >> kLineActShift;
    float* f[3] = {A.h[j.dstBuf] + g.dst, A.hu[j.dstBuf] + g.dst, A.hv[j.dstBuf] + g.dst};
    for (int i = i0; i < g.n; i += stride) {
        const long long gi = (long long)i * g.dstStride;
        BQ3 old{0.f, 0.f, 0.f};
        if (act == 2) old = BQ3{f[0][gi], f[1][gi], f[2][gi]};
        const BQ3 q = ghost_element(A, j, g, i, old, true);
        if (act) {
            f[0][gi] = q.h;
            f[1][gi] = q.hu;
            f[2][gi] = q.hv;
        }
    }
}

__device__ __forceinline__ StateView batch_view(const BatchArena& A, const BatchBlockDesc& d, int buf) {
    StateView v;
    v.h = A.h[buf] + d.off;
    v.hu = A.hu[buf] + d.off;
    v.hv = A.hv[buf] + d.off;
    v.b = A.b + d.off;
    v.nx = d.nx;
    v.ny = d.ny;
    v.pitch = d.pitch;
    v.dx = d.dx;
    v.dy = d.dy;
    return v;
}
__device__ __forceinline__ SumsView batch_sums(const BatchArena& A, const BatchBlockDesc& d, int set) {
    return SumsView{A.s[set][0] + d.off, A.s[set][1] + d.off, A.s[set][2] + d.off, A.s[set][3] + d.off};
}

// ---- LTS scalars of one block (SWE_Block.hh:311-367, 530-562) -------------
__device__ __forceinline__ float lts_total(float T, const BatchBlockState& s) {
    return (float)(T * s.timestepCounter) + (s.stepSize > 0 ? ((float)(T * s.stepSizeCounter) / (float)s.stepSize) : 0.f);
}
__device__ __forceinline__ bool lts_in_sync(const BatchBlockState& s) {
    return s.gl[BND_LEFT] != GL_UNVALID && s.gl[BND_RIGHT] != GL_UNVALID && s.gl[BND_TOP] != GL_UNVALID &&
           s.gl[BND_BOTTOM] != GL_UNVALID && (s.stepSizeCounter <= s.stepSize);
}
__device__ __forceinline__ bool lts_receivable(const BatchBlockState& s, int side) {
    if (s.notified) return false;
    return s.gl[side] == GL_NEXT || s.gl[side] == GL_UNVALID || s.gl[side] == GL_SYNC;
}
// getRoundTimestep's divisor (SWE_Block.hh:536-555): the clamp, then
// 2^-ceil(log(ts)/log(2)) with the natural log taken in single precision and
// the quotient in double, doubled once if 1/divisor still exceeds ts.  The
// single-precision log is the double log rounded to float.
__device__ __forceinline__ int lts_step_size(float T, float ts) {
    if (ts > T) ts = T;
    if (ts < T / kMaxDivisor) ts = T / kMaxDivisor;
    const float lf = (float)log((double)ts);
    int divisor = (int)ldexp(1.0, -(int)ceil((double)lf / log(2.0)));  // pow(2, -ceil(..)), exact
    if (1.f / divisor > ts) divisor *= 2;
    if (divisor > kMaxDivisor) divisor = kMaxDivisor;
    return divisor;
}

// geometry of one ghost line (ghost = 1) or copy layer (ghost = 0), elements 1..n
__host__ __device__ __forceinline__ void side_line(const BatchBlockDesc& d, int side, int ghost, long long& off,
                                                   int& stride, int& n) {
    switch (side) {
        case BND_LEFT:
            off = (long long)(ghost ? 0 : 1) * d.pitch + 1;
            stride = 1;
            n = d.ny;
            break;
        case BND_RIGHT:
            off = (long long)(ghost ? d.nx + 1 : d.nx) * d.pitch + 1;
            stride = 1;
            n = d.ny;
            break;
        case BND_BOTTOM:
            off = (long long)d.pitch + (ghost ? 0 : 1);
            stride = d.pitch;
            n = d.nx;
            break;
        default:  // BND_TOP
            off = (long long)d.pitch + (ghost ? d.ny + 1 : d.ny);
            stride = d.pitch;
            n = d.nx;
            break;
    }
}

}  // namespace swe

#include "kernels_round.hpp"

namespace swe {

// ---- queue pull ------------------------------------------------------------
// One fetch per workgroup: thread 0 takes kWavesPerGroup consecutive items
// (one per wave) and the base reaches every wave through LDS between two
// barriers.  The result is made wave-uniform (an SGPR), so the loop exit and
// everything derived from the item are uniform branches and the cross-lane
// operations of the sweep always run with the full wave.  (Broadcasting a
// lane-0 fetch with a shuffle inside the loop is not safe: the compiler
// treats the shuffled value as divergent and may run the shuffle for lanes
// 1..63 while lane 0 is masked off, which reads 0 and never ends the loop.)
__device__ __forceinline__ int next_items(int* head, int* sBase) {
    if (threadIdx.x == 0) *sBase = atomicAdd(head, kWavesPerGroup);
    __syncthreads();
    const int base = __builtin_amdgcn_readfirstlane(*sBase);
    __syncthreads();  // every wave has read the base before the next fetch overwrites it
    return base;
}

// The batched kernels take the Newton loop's first iteration from the bracket
// (middle_state<true>), as the fused single-block kernel does: measured faster
// under their 72-register budget, same results.
template <int S>
using BatchSolver = dev::SolverCtx<S, true>;

// ---- tiles of a block (strip_tiling / tile_lanes) ---------------------------
__device__ __forceinline__ TileLanes tile_lanes(const BatchBlockDesc& d, int tile, int lane) {
    return tile_lanes(d.nyt, d.nxs, d.pack, d.ny, tile, lane);
}

// ---- sweep: net-update sums of one strip tile ------------------------------
// Same edge order and arithmetic as the single-block strip sweep: the lanes
// of a group own rows 62*yt+1 .. 62*yt+62 (rows r = 1..62; r = 0 and the
// group's top lane are halo) and march over columns 1+xs*XS .. xs*XS+n.
template <int S, int XS>
__device__ __forceinline__ float strip_sums(const BatchSolver<S>& solver, const StateView& v, const SumsView& sums,
                                            const TileLanes& t) {
    const int y = kStripRows * t.yt + t.r;
    const bool rowValid = y <= v.ny + 1;
    const int yc = rowValid ? y : v.ny + 1;
    const bool owned = (t.r >= 1) && (t.r <= kStripRows) && (y <= v.ny);
    const bool rowInterior = (y >= 1) && (y <= v.ny);
    const bool hEdgeValid = (t.r <= kStripRows) && (y <= v.ny);
    const int x0 = 1 + t.xs0 * XS;  // the item's first column; packed groups add a per-lane offset
    const int n = tile_columns(t, v.nx, XS);
    const size_t colOff = (size_t)(t.xs - t.xs0) * XS * v.pitch;
    size_t o = idx(v, x0 - 1, yc) + colOff;
    CellQ cq = cell_q(v.h[o], v.hu[o], v.hv[o], v.b[o]);
    o += v.pitch;
    CellQ nq = cell_q(v.h[o], v.hu[o], v.hv[o], v.b[o]);
    const dev::EdgeOut e = solver.solve(cq.xside(), nq.xside());
    float carryH = e.hR, carryHu = e.huR;
    float smaxv = rowInterior ? e.speed : 0.f;
    for (int k = 0; k < n; ++k) {
        const int x = x0 + k;
        cq = nq;
        o += v.pitch;
        nq = cell_q(v.h[o], v.hu[o], v.hv[o], v.b[o]);
        const dev::EdgeOut ev = solver.solve(cq.xside(), nq.xside());
        if (rowInterior) smaxv = dev::smax(smaxv, ev.speed);
        const dev::EdgeOut eh = solver.solve(cq.yside(), cq.yside_from_lane_above());
        if (hEdgeValid) smaxv = dev::smax(smaxv, eh.speed);
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);
        if (owned) {
            const size_t c = idx(v, x, y) + colOff;
            sums.sx[c] = carryH + ev.hL;
            sums.su[c] = carryHu + ev.huL;
            sums.sy[c] = aH + eh.hL;
            sums.sv[c] = aHv + eh.huL;
        }
        carryH = ev.hR;
        carryHu = ev.huR;
    }
    return wave_max(smaxv);
}

template <int S, int XS>
__global__ __launch_bounds__(256) void batch_sweep_kernel(BatchArena A, const BatchBlockDesc* D, BatchBlockState* st,
                                                          BatchGlobals* G, const int* items) {
    if (G->skip) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int count = G->count;
    const BatchSolver<S> solver(S == SOLVER_AUGRIE_MLP ? G->mlp : nullptr);
    __shared__ int sBase;
    for (;;) {
        const int item = next_items(&G->head, &sBase) + wave;
        if (item - wave >= count) break;  // uniform over the workgroup
        if (item >= count) continue;
        const int code = __builtin_amdgcn_readfirstlane(items[item]);
        const int blk = code >> kTileBits;
        const int tile = code & ((1 << kTileBits) - 1);
        const BatchBlockDesc& d = D[blk];
        const StateView v = batch_view(A, d, st[blk].cur);
        const float m = strip_sums<S, XS>(solver, v, batch_sums(A, d, 0), tile_lanes(d, tile, lane));
        if (lane == 0) atomicMax(&st[blk].smax, __float_as_uint(m));
    }
}

// ---- apply: update the queued tiles' cells --------------------------------
template <int XS>
__global__ __launch_bounds__(256) void batch_apply_kernel(BatchArena A, const BatchBlockDesc* D,
                                                          const BatchBlockState* st, BatchGlobals* G,
                                                          const int* items) {
    if (G->skip) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int count = G->count;
    const int mode = G->mode;
    const float gdt = G->gdt;
    __shared__ int sBase;
    for (;;) {
        const int item = next_items(&G->headApply, &sBase) + wave;
        if (item - wave >= count) break;  // uniform over the workgroup
        if (item >= count) continue;
        const int code = __builtin_amdgcn_readfirstlane(items[item]);
        const int blk = code >> kTileBits;
        const int tile = code & ((1 << kTileBits) - 1);
        const BatchBlockDesc& d = D[blk];
        const StateView v = batch_view(A, d, st[blk].cur);
        const SumsView sums = batch_sums(A, d, 0);
        const float dt = (mode == BATCH_GLOBAL) ? gdt : st[blk].maxTimestep;
        const float dtdx = dt / v.dx;
        const float dtdy = dt / v.dy;
        BoundaryInfo bi;
        for (int k = 0; k < 4; ++k) bi.type[k] = d.bnd[k];
        const TileLanes t = tile_lanes(d, tile, lane);
        const int y = kStripRows * t.yt + t.r;
        const bool owned = (t.r >= 1) && (t.r <= kStripRows) && (y <= v.ny);
        const int x0 = 1 + t.xs * XS;
        const int x1 = x0 + tile_columns(t, v.nx, XS) - 1;
        if (owned) {
            for (int x = x0; x <= x1; ++x) {
                const size_t c = idx(v, x, y);
                const Cell nc = update_cell(v.h[c], v.hu[c], v.hv[c], dtdx, dtdy, sums.sx[c], sums.sy[c], sums.su[c],
                                            sums.sv[c]);
                v.h[c] = nc.h;
                v.hu[c] = nc.hu;
                v.hv[c] = nc.hv;
                write_boundary_ghosts(v, bi, x, y, nc.h, nc.hu, nc.hv, v.h, v.hu, v.hv);
            }
        }
    }
}

// ---- fused strategy: previous update applied on load ----------------------

// One strip tile of a block with its pending update (sums sIn, dt) applied as
// the cells are loaded: the updated state goes to buffer `o` (interior cells,
// OUTFLOW/WALL ghosts derived from the updated interior, CONNECT_WITHIN_RANK
// ghosts from this iteration's ghost-line step (jobs jb, geometry gg of the
// block's four sides), corners from the diagonal interior cell), the edges are
// solved on it and the new sums go to sOut.  The same loads, edge order and
// arithmetic as the single-block fused kernel, so sums and maxima equal the
// sweep + apply pair.
template <int S, int XS>
__device__ __forceinline__ float fused_tile(const BatchSolver<S>& solver, const StateView& v, const StateView& o,
                                            const SumsView& sIn, const SumsView& sOut, bool pend, float dt,
                                            const int* bnd, const TileLanes& t, int blk, uint2* deferQ,
                                            unsigned* deferCount, const BatchArena& A, const GhostJob* jb,
                                            const GhostGeom* gg) {
    const int nx = v.nx, ny = v.ny;
    const float dtdx = dt / v.dx;
    const float dtdy = dt / v.dy;
    const int y = kStripRows * t.yt + t.r;
    const bool rowValid = y <= ny + 1;
    const int yc = rowValid ? y : ny + 1;
    const bool owned = (t.r >= 1) && (t.r <= kStripRows) && (y <= ny);
    const bool rowInterior = (y >= 1) && (y <= ny);
    const bool hEdgeValid = (t.r <= kStripRows) && (y <= ny);
    const bool rowBottom = (y == 0);
    const bool rowTop = (y == ny + 1);
    const bool carryRow = owned || (rowBottom && t.r == 0) || (rowTop && t.r >= 1);
    // the item's first column (uniform); a packed group's lanes address their
    // own segment through a constant per-lane offset
    const int x0 = 1 + t.xs0 * XS;
    const int n = tile_columns(t, nx, XS);
    const size_t colOff = (size_t)(t.xs - t.xs0) * XS * v.pitch;
    const size_t rowC = (size_t)yc + colOff, rowY = (size_t)y + colOff;
    auto at = [&](int x, size_t row) { return (size_t)x * v.pitch + row; };
    const int tB = bnd[BND_BOTTOM], tT = bnd[BND_TOP], tL = bnd[BND_LEFT], tR = bnd[BND_RIGHT];
    auto ow = [](int t) { return t == OUTFLOW || t == WALL; };

    auto load_updated = [&](int x) -> BQ3 {
        const size_t c = at(x, rowC);
        BQ3 q{v.h[c], v.hu[c], v.hv[c]};
        // x is the item's column: a packed item's segments are all inner, so
        // the test has the same outcome for every lane
        if (pend && x >= 1 && x <= nx && rowInterior) {
            const Cell u = update_cell(q.h, q.hu, q.hv, dtdx, dtdy, sIn.sx[c], sIn.sy[c], sIn.su[c], sIn.sv[c]);
            q = BQ3{u.h, u.hu, u.hv};
        }
        return q;
    };
    auto fix_rows = [&](BQ3 q) -> BQ3 {
        const float dh = __shfl_down(q.h, 1, 64), dhu = __shfl_down(q.hu, 1, 64), dhv = __shfl_down(q.hv, 1, 64);
        const float uh = __shfl_up(q.h, 1, 64), uhu = __shfl_up(q.hu, 1, 64), uhv = __shfl_up(q.hv, 1, 64);
        if (rowBottom && ow(tB)) q = BQ3{dh, dhu, (tB == WALL) ? -dhv : dhv};
        if (rowTop && ow(tT)) q = BQ3{uh, uhu, (tT == WALL) ? -uhv : uhv};
        return q;
    };
    auto fix_column = [&](BQ3 raw, BQ3 in, int type) -> BQ3 {
        const float dh = __shfl_down(in.h, 1, 64), dhu = __shfl_down(in.hu, 1, 64), dhv = __shfl_down(in.hv, 1, 64);
        const float uh = __shfl_up(in.h, 1, 64), uhu = __shfl_up(in.hu, 1, 64), uhv = __shfl_up(in.hv, 1, 64);
        BQ3 q = raw;
        if (rowInterior && ow(type)) q = BQ3{in.h, (type == WALL) ? -in.hu : in.hu, in.hv};
        if (rowBottom) q = BQ3{dh, dhu, dhv};
        if (rowTop) q = BQ3{uh, uhu, uhv};
        return q;
    };
    auto store = [&](int x, const BQ3& q) {
        const size_t c = at(x, rowY);
        o.h[c] = q.h;
        o.hu[c] = q.hu;
        o.hv[c] = q.hv;
    };

    // CONNECT_WITHIN_RANK sides: this iteration's ghost-line step, in place
    // in v before the cells are loaded, each ghost cell by the tile that owns
    // it: a ghost row's cells by the tile of their column (the lanes spread
    // over the row's columns), a ghost column's cells by the owned lanes of
    // their row.  Other tiles may load a ghost cell before or after the
    // update only as a halo cell, whose edges add neither sums nor speeds
    // (the speeds of x-edges are taken on owned rows; every interior row is
    // owned by one strip, so the maximum is that over all edges).
    bool ghostRows = false;
    auto row_step = [&](int side, bool isRow) {
        const GhostJob j = jb[side];
        if (j.op == 0) return;
        const GhostGeom g = gg[side];
        const unsigned long long rows = __ballot(isRow);  // one lane per group of a packed item
        const int groups = __popcll(rows), row = (side == BND_BOTTOM) ? 0 : ny + 1;
        for (int p = (int)(threadIdx.x & 63); p < groups * n; p += 64) {
            const int x = x0 + (p / n) * XS + p % n;
            const size_t c = (size_t)x * v.pitch + row;
            const BQ3 q = ghost_element(A, j, g, x - 1, BQ3{v.h[c], v.hu[c], v.hv[c]}, true);
            v.h[c] = q.h;
            v.hu[c] = q.hu;
            v.hv[c] = q.hv;
        }
        ghostRows = true;
    };
    if (tB == CONNECT_WITHIN_RANK && t.yt == 0) row_step(BND_BOTTOM, rowBottom);
    if (tT == CONNECT_WITHIN_RANK && kStripRows * t.yt + 63 >= ny + 1) row_step(BND_TOP, rowTop);
    auto column_step = [&](int side, int x) {
        const GhostJob j = jb[side];
        if (j.op == 0 || !owned) return;
        const size_t c = at(x, rowC);
        const BQ3 q = ghost_element(A, j, gg[side], y - 1, BQ3{v.h[c], v.hu[c], v.hv[c]}, true);
        v.h[c] = q.h;
        v.hu[c] = q.hu;
        v.hv[c] = q.hv;
    };
    if (t.first && tL == CONNECT_WITHIN_RANK) column_step(BND_LEFT, 0);
    if (t.last && tR == CONNECT_WITHIN_RANK) column_step(BND_RIGHT, nx + 1);
    if (ghostRows) {  // other lanes of the wave load the row cells
        __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }

    BQ3 cur = fix_rows(load_updated(x0));
    BQ3 prev = load_updated(x0 - 1);
    if (t.first) prev = fix_column(prev, cur, tL);  // x0 == 1
    if (t.first && carryRow) store(0, prev);
    CellQ pq = cell_q(prev.h, prev.hu, prev.hv, v.b[at(x0 - 1, rowC)]);
    CellQ cq = cell_q(cur.h, cur.hu, cur.hv, v.b[at(x0, rowC)]);
    // deferred-edge entries: the lane's own column, the block in the high bits
    const unsigned xLane = (unsigned)(t.xs - t.xs0) * XS, blkBits = (unsigned)blk << kDeferBlockShift;
    dev::EdgeOut e = solver.solve(pq.xside(), cq.xside());
    if constexpr (S == SOLVER_AUGRIE_DEFER)  // edge (0|1): only the first segment's items have t.first
        e = defer_edge(e, owned && t.first, make_uint2(blkBits, (unsigned)y), deferQ, deferCount);
    float carryH = e.hR, carryHu = e.huR;
    float smaxv = owned ? e.speed : 0.f;
    // Column x+1's loads are issued one column ahead (a software pipeline),
    // as in the single-block fused kernel; past the tile they read column
    // nx + 1 at most (a packed item's segments are inner ones) and are not
    // used.  Not for the deferring sweep, which has no registers to spare.
    struct Raw {
        float h, hu, hv, sx, sy, su, sv, b;
    };
    auto load_raw = [&](int x) -> Raw {
        const size_t c = at(min(x, nx + 1), rowC);
        return Raw{v.h[c], v.hu[c], v.hv[c], sIn.sx[c], sIn.sy[c], sIn.su[c], sIn.sv[c], v.b[c]};
    };
    constexpr bool kAhead = S != SOLVER_AUGRIE_DEFER;
    Raw ahead{};
    if constexpr (kAhead) ahead = load_raw(x0 + 1);
    for (int k = 0; k < n; ++k) {
        const int x = x0 + k;
        const bool rightGhost = t.last && k == n - 1;  // x + 1 == nx + 1
        const Raw r = kAhead ? ahead : load_raw(x + 1);
        if constexpr (kAhead) ahead = load_raw(x + 2);
        BQ3 nxt{r.h, r.hu, r.hv};
        if (pend && x + 1 <= nx && rowInterior) {
            const Cell u = update_cell(nxt.h, nxt.hu, nxt.hv, dtdx, dtdy, r.sx, r.sy, r.su, r.sv);
            nxt = BQ3{u.h, u.hu, u.hv};
        }
        nxt = fix_rows(nxt);
        if (rightGhost) nxt = fix_column(nxt, cur, tR);
        const CellQ nq = cell_q(nxt.h, nxt.hu, nxt.hv, r.b);
        if (carryRow) store(x, cur);
        if (rightGhost && carryRow) store(nx + 1, nxt);
        dev::EdgeOut ev = solver.solve(cq.xside(), nq.xside());
        if constexpr (S == SOLVER_AUGRIE_DEFER)
            ev = defer_edge(ev, owned, make_uint2(blkBits | ((unsigned)x + xLane), (unsigned)y), deferQ, deferCount);
        if (owned) smaxv = dev::smax(smaxv, ev.speed);
        dev::EdgeOut eh = solver.solve(cq.yside(), cq.yside_from_lane_above());
        if constexpr (S == SOLVER_AUGRIE_DEFER)
            eh = defer_edge(eh, owned || y == 0, make_uint2(kDeferY | blkBits | ((unsigned)x + xLane), (unsigned)y),
                            deferQ, deferCount);
        if (hEdgeValid) smaxv = dev::smax(smaxv, eh.speed);
        const float aH = __shfl_up(eh.hR, 1, 64);
        const float aHv = __shfl_up(eh.huR, 1, 64);
        if (owned) {
            const size_t c = at(x, rowY);
            sOut.sx[c] = carryH + ev.hL;
            sOut.su[c] = carryHu + ev.huL;
            sOut.sy[c] = aH + eh.hL;
            sOut.sv[c] = aHv + eh.huL;
        }
        carryH = ev.hR;
        carryHu = ev.huR;
        cur = nxt;
        cq = nq;
    }
    return wave_max(smaxv);
}

// direct = 0: persistent grid, waves pull items until the queue is drained;
// direct = 1: one wave per possible item (the grid covers every block's
// tiles and every side's ghost-line step), a wave takes item
// blockIdx.x * kWavesPerGroup + wave or leaves.  An item is a strip tile of a
// block that computes this iteration or, for LTS, the ghost-line step of a
// side of a block that waits (tile kLineTile + side).
// Six waves per SIMD (at most 80 VGPRs, no spills) with the early column
// loads; at seven without them the LTS tiling ran 5% and the ensemble 4%
// slower.  The deferring sweep runs at six without the early loads (77
// registers; at seven it spilled), the learned-middle-state solver at its
// own register budget (its network call needs far more).
template <int S, int XS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S == SOLVER_AUGRIE_MLP ? 1 : 6, 8))) void
batch_fused_kernel(BatchArena A, const BatchBlockDesc* D, BatchBlockState* st, BatchGlobals* G, const int* items,
                   int direct, uint2* deferQ, const GhostJob* jobs, const GhostGeom* geo) {
    if (G->skip) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int count = G->count;
    const BatchSolver<S> solver(S == SOLVER_AUGRIE_MLP ? G->mlp : nullptr);
    auto run_item = [&](int item) {
        const int code = __builtin_amdgcn_readfirstlane(items[item]);
        const int blk = code >> kTileBits;
        const int tile = code & ((1 << kTileBits) - 1);
        if (tile >= kLineTile) {  // a waiting block's side
            const int js = 4 * blk + tile - kLineTile;
            ghost_line_step(A, jobs[js], geo[js], lane, 64);
            return;
        }
        const BatchBlockDesc& d = D[blk];
        const int cur = __builtin_amdgcn_readfirstlane(st[blk].cur);
        const int sidx = __builtin_amdgcn_readfirstlane(st[blk].sidx);
        const bool pend = __builtin_amdgcn_readfirstlane(st[blk].pend) != 0;
        const float dt = st[blk].maxTimestep;
        const float m = fused_tile<S, XS>(solver, batch_view(A, d, cur), batch_view(A, d, cur ^ 1),
                                          batch_sums(A, d, sidx), batch_sums(A, d, sidx ^ 1), pend, dt, d.bnd,
                                          tile_lanes(d, tile, lane), blk, deferQ, &G->deferCount, A, jobs + 4 * blk,
                                          geo + 4 * blk);
        if (lane == 0) atomicMax(&st[blk].smax, __float_as_uint(m));
    };
    if (direct) {
        const int item = blockIdx.x * kWavesPerGroup + wave;  // uniform per wave
        if (item < count) run_item(item);
        return;
    }
    __shared__ int sBase;
    for (;;) {
        const int item = next_items(&G->head, &sBase) + wave;
        if (item - wave >= count) break;  // uniform over the workgroup
        if (item >= count) continue;
        run_item(item);
    }
}

}  // namespace swe
