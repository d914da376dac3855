// swe_batch.hip -- batched engine of libswe_hip.so: the edges kernel of the
// driver iteration (ghost exchange, LTS checks, queue fill), the persistent
// apply, start / end-of-run kernels, hipGraph capture / replay, the autotuner
// and the C ABI of include/swe_hip_batch.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/swe_hip_batch.h"
#include "../../include/swe_hip_mlp.h"
#include "kernels_batch.hpp"
#include "swe_internal.h"

using namespace swe;

#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

namespace {

constexpr int kCtlThreads = 256;

// ---- initial ghost cells: applyBoundaryConditions (SWE_Block.hh:757-911)
// for every block and the in-rank bathymetry copy
// (SWE_DimensionalSplittingMPIOverdecomp.cpp:208-235).  One workgroup per block.
__global__ __launch_bounds__(kCtlThreads) void batch_init_kernel(BatchArena A, const BatchBlockDesc* D) {
    const BatchBlockDesc d = D[blockIdx.x];
    float* h = A.h[0] + d.off;
    float* hu = A.hu[0] + d.off;
    float* hv = A.hv[0] + d.off;
    float* b = A.b + d.off;
    const int nx = d.nx, ny = d.ny, p = d.pitch;
    for (int side = 0; side < 4; ++side) {
        const int t = d.bnd[side];
        if (t == OUTFLOW || t == WALL) {
            long long go, so;
            int gs, ss, n;
            side_line(d, side, 1, go, gs, n);
            side_line(d, side, 0, so, ss, n);
            const bool vertical = side == BND_LEFT || side == BND_RIGHT;
            const float su = (t == WALL && vertical) ? -1.f : 1.f;
            const float sv = (t == WALL && !vertical) ? -1.f : 1.f;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const long long g = go + (long long)i * gs, s = so + (long long)i * ss;
                h[g] = h[s];
                hu[g] = (su < 0.f) ? -hu[s] : hu[s];
                hv[g] = (sv < 0.f) ? -hv[s] : hv[s];
            }
        } else if (t == CONNECT_WITHIN_RANK) {
            const BatchBlockDesc dn = D[d.nbr[side]];
            long long go, so;
            int gs, ss, n, nn;
            side_line(d, side, 1, go, gs, n);
            side_line(dn, side ^ 1, 0, so, ss, nn);
            const float* bn = A.b + dn.off;
            for (int i = threadIdx.x; i < n; i += blockDim.x) b[go + (long long)i * gs] = bn[so + (long long)i * ss];
        }
    }
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        const int cx = (k < 2) ? 0 : nx + 1, cy = (k & 1) ? ny + 1 : 0;
        const int sx = (k < 2) ? 1 : nx, sy = (k & 1) ? ny : 1;
        const long long c = (long long)cx * p + cy, s = (long long)sx * p + sy;
        h[c] = h[s];
        hu[c] = hu[s];
        hv[c] = hv[s];
    }
}

// The lines launch runs one workgroup per (block, side, piece of the side):
// the lines of a block are short and strided (bottom / top), so spreading
// them keeps each thread at one or two elements.
constexpr int kSideChunks = 4;

// ---- lines: the ghost-line data of an iteration ----------------------------
// Per (block, side, piece) workgroup, the job the previous round (or the
// start of the run) wrote for the side:
//   global dt: setGhostLayer's in-rank copy
//     (SWE_DimensionalSplittingMPIOverdecomp.cpp:273-304): the neighbour's
//     copy layer, with its pending update applied, into the ghost line;
//   LTS: the same copy into the side's receive line when the side receives
//     this iteration, then checkAllGhostlayers' ghost-line step
//     (SWE_Block.hh:375-436): copy when the time stamps agree,
//     old + (new - old) * (t_me / t_neighbour) when the neighbour is ahead,
//     nothing when it is behind.
// Grid (blocks, 4, kSideChunks).
__global__ __launch_bounds__(kCtlThreads) void batch_lines_kernel(BatchArena A, const GhostGeom* geo,
                                                                  const GhostJob* jobs) {
    const int js = 4 * blockIdx.x + blockIdx.y;
    const GhostJob j = jobs[js];
    if (j.op == 0) return;
    ghost_line_step(A, j, geo[js], blockIdx.z * blockDim.x + threadIdx.x, blockDim.x * gridDim.z);
}

// ---- work queue of an iteration: the tiles of every active block -----------
// One workgroup: a prefix sum over the blocks gives each block's first item,
// then each wave writes the items of every fourth block.
__device__ void fill_queue(const BatchBlockDesc* D, const BatchBlockState* st, BatchGlobals* G, int* items,
                           bool all, const int* lineMask = nullptr) {
    __shared__ int sBase[kMaxBatchBlocks];
    const int nb = G->nblocks;
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
    auto active = [&](int b) { return all || st[b].active != 0; };
    auto mask = [&](int b) { return lineMask ? lineMask[b] : 0; };
    int mine = 0;
    for (int b = b0; b < b1; ++b) mine += queue_entries(active(b), D[b].ntiles, mask(b));
    int total;
    int run = wg_exclusive_scan<kCtlThreads>(mine, total);
    for (int b = b0; b < b1; ++b) {
        sBase[b] = run;
        run += queue_entries(active(b), D[b].ntiles, mask(b));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    for (int b = wave; b < nb; b += waves) queue_write(items, sBase[b], b, active(b), D[b].ntiles, mask(b), lane);
    if (threadIdx.x == 0) {
        G->count = total;
        G->head = 0;
        G->headApply = 0;
        if (all) G->allItems = total;
    }
}

// the queue of the next iteration and the iteration's skip flag
__device__ void next_queue(const BatchBlockDesc* D, const BatchBlockState* st, BatchGlobals* G, int* items) {
    if (G->mode == BATCH_LTS) {
        fill_queue(D, st, G, items, false);
    } else if (threadIdx.x == 0) {  // every block computes every iteration: the queue stays as filled
        G->count = G->allItems;
        G->head = 0;
        G->headApply = 0;
    }
    if (threadIdx.x == 0) G->skip = G->stop;
}

// ---- round: one workgroup after the sweep (round_body) -----------------------
__global__ __launch_bounds__(kRoundThreads) void batch_round_kernel(const BatchBlockDesc* D, BatchBlockState* st,
                                                                    BatchGlobals* G, RoundIO io) {
    if (G->skip) return;
    round_body<kRoundThreads, kMaxBatchBlocks>(threadIdx.x, D, st, G, io);
}

// ---- store strategy: the next queue, after the apply launch ------------------
__global__ __launch_bounds__(kCtlThreads) void batch_fill_kernel(const BatchBlockDesc* D, const BatchBlockState* st,
                                                                 BatchGlobals* G, int* items) {
    if (G->skip) return;
    next_queue(D, st, G, items);
}

// ---- after an upload: the first iteration's bookkeeping ----------------------
__global__ __launch_bounds__(kCtlThreads) void batch_prepare_kernel(const BatchBlockDesc* D, BatchBlockState* st,
                                                                    int nblocks, int mode, float duration) {
    for (int blk = threadIdx.x; blk < nblocks; blk += blockDim.x) {
        if (mode == BATCH_LTS)
            lts_transitions(D[blk], st[blk], [&](int k) { return st[k].total; }, duration);
        else
            st[blk].active = 1;
    }
}

// ---- end of a fused run: apply the pending updates ---------------------------
// Pass 1 writes the updated state of every block with a pending update into
// its other buffer (interior updated, OUTFLOW / WALL ghosts from the updated
// interior, other ghosts copied, corners from the updated diagonal cell).
// Pass 2 copies every block whose state then sits in buffer 1 to buffer 0.
// Pass 3 resets the flags.  Grid: (64, nblocks).
__device__ __forceinline__ Cell pending_cell(const BatchArena& A, const BatchBlockDesc& d, const BatchBlockState& s,
                                             int x, int y) {
    const StateView v = batch_view(A, d, s.cur);
    const size_t c = idx(v, x, y);
    Cell q{v.h[c], v.hu[c], v.hv[c]};
    if (x >= 1 && x <= d.nx && y >= 1 && y <= d.ny) {
        const SumsView ps = batch_sums(A, d, s.sidx);
        q = update_cell(q.h, q.hu, q.hv, s.maxTimestep / d.dx, s.maxTimestep / d.dy, ps.sx[c], ps.sy[c], ps.su[c],
                        ps.sv[c]);
    }
    return q;
}

__global__ __launch_bounds__(256) void batch_normalize_kernel(BatchArena A, const BatchBlockDesc* D,
                                                              const BatchBlockState* st) {
    const int blk = blockIdx.y;
    const BatchBlockState s = st[blk];
    if (!s.pend) return;
    const BatchBlockDesc d = D[blk];
    const int nx = d.nx, ny = d.ny, p = d.pitch;
    const long long cells = (long long)(nx + 2) * p;
    const StateView o = batch_view(A, d, s.cur ^ 1);
    auto ow = [](int t) { return t == OUTFLOW || t == WALL; };
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(c / p), y = (int)(c % p);
        const bool gx = x == 0 || x == nx + 1, gy = y == 0 || y == ny + 1;
        Cell q;
        if (gx && gy) {
            q = pending_cell(A, d, s, x == 0 ? 1 : nx, y == 0 ? 1 : ny);  // corner: diagonal interior cell
        } else if (gx) {
            const int t = d.bnd[x == 0 ? BND_LEFT : BND_RIGHT];
            if (ow(t)) {
                q = pending_cell(A, d, s, x == 0 ? 1 : nx, y);
                if (t == WALL) q.hu = -q.hu;
            } else {
                q = pending_cell(A, d, s, x, y);  // ghost: as stored
            }
        } else if (gy) {
            const int t = d.bnd[y == 0 ? BND_BOTTOM : BND_TOP];
            if (ow(t)) {
                q = pending_cell(A, d, s, x, y == 0 ? 1 : ny);
                if (t == WALL) q.hv = -q.hv;
            } else {
                q = pending_cell(A, d, s, x, y);
            }
        } else {
            q = pending_cell(A, d, s, x, y);
        }
        o.h[c] = q.h;
        o.hu[c] = q.hu;
        o.hv[c] = q.hv;
    }
}

__global__ __launch_bounds__(256) void batch_gather_kernel(BatchArena A, const BatchBlockDesc* D,
                                                           const BatchBlockState* st) {
    const int blk = blockIdx.y;
    const BatchBlockState s = st[blk];
    const int where = s.pend ? (s.cur ^ 1) : s.cur;
    if (where == 0) return;
    const BatchBlockDesc d = D[blk];
    const long long cells = (long long)(d.nx + 2) * d.pitch;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        A.h[0][d.off + c] = A.h[1][d.off + c];
        A.hu[0][d.off + c] = A.hu[1][d.off + c];
        A.hv[0][d.off + c] = A.hv[1][d.off + c];
    }
}

__global__ void batch_flags_kernel(BatchBlockState* st, int nblocks) {
    for (int blk = threadIdx.x; blk < nblocks; blk += blockDim.x) {
        st[blk].cur = 0;
        st[blk].pend = 0;
        st[blk].sidx = 0;
    }
}

// start of a run: `target` more units, the queue and the ghost-line jobs of
// the first iteration (the tiling or the block states may have changed since
// the last run).  One workgroup.
__global__ __launch_bounds__(kCtlThreads) void batch_begin_kernel(const BatchBlockDesc* D, const BatchBlockState* st,
                                                                  BatchGlobals* G, int* items, int target, int logCap,
                                                                  int* hostStat, GhostJob* jobs) {
    if (threadIdx.x == 0) {
        G->target = target;
        G->done = 0;
        G->stop = target <= 0 ? 1 : 0;
        G->skip = G->stop;
        G->itersUnit = 0;
        G->logCap = logCap;
        hostStat[0] = G->stop;
        hostStat[1] = 0;
        hostStat[2] = 0;
    }
    // the jobs; with the fused strategy, LTS queues the steps of the waiting blocks' sides
    __shared__ int sMask[kMaxBatchBlocks];
    const int mode = G->mode;
    const bool fusedLts = mode == BATCH_LTS && G->fused != 0;
    for (int blk = threadIdx.x; blk < G->nblocks; blk += blockDim.x) {
        int m = 0;
        if (mode != BATCH_ENSEMBLE) {
            for (int side = 0; side < 4; ++side) {
                const int bnd = D[blk].bnd[side], nb = D[blk].nbr[side];
                const bool link = bnd == CONNECT_WITHIN_RANK;
                const GhostJob j = line_job(st[blk], side, bnd, mode == BATCH_LTS, target <= 0,
                                            link ? line_flags(st[nb]) : 0, link ? st[nb].maxTimestep : 0.f);
                jobs[4 * blk + side] = j;
                if (j.op != 0 && !st[blk].active && fusedLts) m |= 1 << side;
            }
        }
        sMask[blk] = m;
    }
    __syncthreads();
    fill_queue(D, st, G, items, mode != BATCH_LTS, sMask);
}

template <int XS>
int launch_apply(int groups, BatchArena A, const BatchBlockDesc* D, const BatchBlockState* st, BatchGlobals* G,
                 const int* items, hipStream_t stream) {
    hipLaunchKernelGGL((batch_apply_kernel<XS>), dim3(groups), dim3(64 * kWavesPerGroup), 0, stream, A, D, st, G,
                       items);
    return 0;
}

int launch_fused(int solver, int xs, int groups, int direct, BatchArena A, const BatchBlockDesc* D,
                 BatchBlockState* st, BatchGlobals* G, const int* items, uint2* deferQ, const GhostJob* jobs,
                 const GhostGeom* geo, hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_batch_fused_solver<SOLVER_HLLE>(xs, groups, direct, A, D, st, G, items, nullptr, jobs,
                                                          geo, stream);
        case SOLVER_FWAVE:
            return launch_batch_fused_solver<SOLVER_FWAVE>(xs, groups, direct, A, D, st, G, items, nullptr, jobs,
                                                           geo, stream);
        case SOLVER_AUGRIE:
            return launch_batch_fused_solver<SOLVER_AUGRIE>(xs, groups, direct, A, D, st, G, items, nullptr, jobs,
                                                            geo, stream);
        case SOLVER_AUGRIE_MLP:
            return launch_batch_fused_solver<SOLVER_AUGRIE_MLP>(xs, groups, direct, A, D, st, G, items, nullptr, jobs,
                                                                geo, stream);
        case SOLVER_AUGRIE_DEFER:
            return launch_batch_fused_solver<SOLVER_AUGRIE_DEFER>(xs, groups, direct, A, D, st, G, items, deferQ, jobs,
                                                                  geo, stream);
        default:
            return launch_batch_fused_solver<SOLVER_HYBRID>(xs, groups, direct, A, D, st, G, items, nullptr, jobs,
                                                            geo, stream);
    }
}

int launch_sweep(int solver, int xs, int groups, BatchArena A, const BatchBlockDesc* D, BatchBlockState* st,
                 BatchGlobals* G, const int* items, hipStream_t stream) {
    switch (solver) {
        case SOLVER_HLLE:
            return launch_batch_sweep_solver<SOLVER_HLLE>(xs, groups, A, D, st, G, items, stream);
        case SOLVER_FWAVE:
            return launch_batch_sweep_solver<SOLVER_FWAVE>(xs, groups, A, D, st, G, items, stream);
        case SOLVER_AUGRIE:
            return launch_batch_sweep_solver<SOLVER_AUGRIE>(xs, groups, A, D, st, G, items, stream);
        case SOLVER_AUGRIE_MLP:
            return launch_batch_sweep_solver<SOLVER_AUGRIE_MLP>(xs, groups, A, D, st, G, items, stream);
        default:
            return launch_batch_sweep_solver<SOLVER_HYBRID>(xs, groups, A, D, st, G, items, stream);
    }
}

}  // namespace

struct swe_hip_batch {
    swe_hip_mlp* net = nullptr;  // the attached network's handle
    int nblocks = 0, solver = SOLVER_AUGRIE, mode = BATCH_ENSEMBLE;
    float T = 1.f, duration = 0.f;
    std::vector<BatchBlockDesc> desc;
    long long total = 0;      // floats per field over all blocks
    long long lineTotal = 0;  // floats of the LTS receive lines
    int itemCap = 0;
    BatchArena A{};
    BatchBlockDesc* dDesc = nullptr;
    BatchBlockState* dState = nullptr;
    BatchGlobals* dG = nullptr;
    int* dItems = nullptr;
    GhostGeom* dGeom = nullptr;  // 4 per block, fixed at create
    GhostJob* dJobs = nullptr;   // 4 per block, written by the round and the start of a run
    float* dDtLog = nullptr;
    int* dIterLog = nullptr;
    int logCap = 0;
    int* hostStat = nullptr;  // pinned, written by the round kernel
    int* dHostStat = nullptr;
    int xs = 8, groups = 0, chunk = 16, useGraph = 1;
    long long itemsAll = 0;   // work items of all blocks at the current xs
    long long itemsMost = 0;  // the longest queue an iteration can have (the direct grid)
    int cus = 256;
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;
    bool uploaded = false;
    const unsigned* mlp = nullptr;  // attached learned middle state (weight fragments), or null
    int fused = 1;                  // 1: fused sweep (update applied on load), 0: sweep + apply
    uint2* deferQ = nullptr;        // deferred-edge queue of the learned middle state's fused sweep, or null
    unsigned deferCap = 0;
};

namespace {

void free_graph(swe_hip_batch* k) {
    if (k->exec) (void)hipGraphExecDestroy(k->exec);
    k->exec = nullptr;
}

int tiles_for(swe_hip_batch* k, int xs, bool upload) {
    long long count = 0, most = 0;
    for (auto& d : k->desc) {
        const StripTiling g = strip_tiling(d.nx, d.ny, xs);
        d.nyt = g.nyt;
        d.nxs = g.nxs;
        d.pack = g.pack;
        d.ntiles = g.items;
        count += d.ntiles;
        // an LTS queue holds a block's tiles or the steps of its four sides
        most += (k->mode == BATCH_LTS) ? std::max(d.ntiles, 4) : d.ntiles;
    }
    k->itemsAll = count;
    k->itemsMost = most;
    if (most > k->itemCap) {
        set_error("batch: %lld work items exceed the queue (%d)", most, k->itemCap);
        return SWE_HIP_EINVAL;
    }
    if (upload) {
        HIP_TRY(hipMemcpyAsync(k->dDesc, k->desc.data(), sizeof(BatchBlockDesc) * k->desc.size(),
                               hipMemcpyHostToDevice, k->stream));
        HIP_TRY(hipStreamSynchronize(k->stream));
    }
    return 0;
}

int default_groups(swe_hip_batch* k) {
    long long items = 0;
    for (auto& d : k->desc) items += d.ntiles;
    const long long want = (items + kWavesPerGroup - 1) / kWavesPerGroup;
    return (int)std::max<long long>(1, std::min<long long>(want, 8LL * k->cus));
}

// the launches of one driver iteration
int enqueue_iteration(swe_hip_batch* k) {
    const dim3 ctl(kCtlThreads);
    // the fused sweep does the ghost-line steps itself
    if (k->mode != BATCH_ENSEMBLE && !k->fused)
        hipLaunchKernelGGL(batch_lines_kernel, dim3(k->nblocks, 4, kSideChunks), ctl, 0, k->stream, k->A, k->dGeom,
                           k->dJobs);
    const int solver = (k->mlp && k->solver == SOLVER_AUGRIE) ? SOLVER_AUGRIE_MLP : k->solver;
    // the learned middle state's fused sweep defers its hard edges to a second launch
    const bool defer = k->fused && solver == SOLVER_AUGRIE_MLP && k->deferQ;
    // groups == 0: the direct form (one wave per possible item) for the fused
    // sweep; the store strategy's persistent grids then take the default size
    const bool direct = k->groups == 0;
    const int persistent = direct ? default_groups(k) : k->groups;
    const int directGrid = (int)((k->itemsMost + kWavesPerGroup - 1) / kWavesPerGroup);
    int rc = k->fused ? launch_fused(defer ? SOLVER_AUGRIE_DEFER : solver, k->xs,
                                     direct ? std::max(1, directGrid) : k->groups, direct ? 1 : 0, k->A, k->dDesc,
                                     k->dState, k->dG, k->dItems, k->deferQ, k->dJobs, k->dGeom, k->stream)
                      : launch_sweep(solver, k->xs, persistent, k->A, k->dDesc, k->dState, k->dG, k->dItems, k->stream);
    if (rc) return rc;
    if (defer) {
        // one round of waves at the learned solver's three waves per SIMD
        rc = launch_batch_defer_solve(k->A, k->dDesc, k->dState, k->dG, k->deferQ, k->deferCap, 3 * k->cus,
                                      k->stream);
        if (rc) return rc;
    }
    const RoundIO io{k->dItems, k->dDtLog, k->dIterLog, k->dHostStat, k->dJobs};
    hipLaunchKernelGGL(batch_round_kernel, dim3(1), dim3(kRoundThreads), 0, k->stream, k->dDesc, k->dState, k->dG,
                       io);
    if (!k->fused) {
        switch (k->xs) {
            case 4:
                launch_apply<4>(persistent, k->A, k->dDesc, k->dState, k->dG, k->dItems, k->stream);
                break;
            case 16:
                launch_apply<16>(persistent, k->A, k->dDesc, k->dState, k->dG, k->dItems, k->stream);
                break;
            default:
                launch_apply<8>(persistent, k->A, k->dDesc, k->dState, k->dG, k->dItems, k->stream);
                break;
        }
        hipLaunchKernelGGL(batch_fill_kernel, dim3(1), ctl, 0, k->stream, k->dDesc, k->dState, k->dG, k->dItems);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch iteration launch");
    return 0;
}

int ensure_graph(swe_hip_batch* k) {
    if (k->exec) return 0;
    HIP_TRY(hipStreamBeginCapture(k->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < k->chunk && rc == 0; ++i) rc = enqueue_iteration(k);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(k->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "batch graph capture");
    const hipError_t e2 = hipGraphInstantiate(&k->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        k->exec = nullptr;
        return hip_fail(e2, "batch graph instantiate");
    }
    return 0;
}

// one chunk of driver iterations
int enqueue_chunk(swe_hip_batch* k) {
    if (k->useGraph) {
        int rc = ensure_graph(k);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(k->exec, k->stream));
        return 0;
    }
    for (int i = 0; i < k->chunk; ++i) {
        int rc = enqueue_iteration(k);
        if (rc) return rc;
    }
    return 0;
}

int ensure_log(swe_hip_batch* k, int units) {
    if (units <= k->logCap) return 0;
    const int cap = std::max(units, 1024);
    float* dl = nullptr;
    int* il = nullptr;
    HIP_TRY(hipStreamSynchronize(k->stream));
    HIP_TRY(hipMalloc(&dl, sizeof(float) * cap));
    const hipError_t e = hipMalloc(&il, sizeof(int) * cap);
    if (e != hipSuccess) {
        (void)hipFree(dl);
        return hip_fail(e, "batch log");
    }
    if (k->dDtLog) (void)hipFree(k->dDtLog);
    if (k->dIterLog) (void)hipFree(k->dIterLog);
    k->dDtLog = dl;
    k->dIterLog = il;
    k->logCap = cap;
    free_graph(k);  // captured launches hold the old pointers
    return 0;
}

// Advance until `units` are done; returns the iterations replayed.
int advance(swe_hip_batch* k, int units, long long maxIter, long long* itersOut) {
    hipLaunchKernelGGL(batch_begin_kernel, dim3(1), dim3(kCtlThreads), 0, k->stream, k->dDesc, k->dState, k->dG,
                       k->dItems, units, k->logCap, k->dHostStat, k->dJobs);
    HIP_TRY(hipGetLastError());
    long long iters = 0;
    if (k->mode != BATCH_LTS) {
        // the unit count fixes the iterations; the device skips the overshoot
        while (iters < units) {
            int rc = enqueue_chunk(k);
            if (rc) return rc;
            iters += k->chunk;
        }
        HIP_TRY(hipStreamSynchronize(k->stream));
    } else {
        // two chunks in flight; the round kernel reports completion in pinned memory
        hipEvent_t ev[2];
        HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        int rc = 0;
        for (long long r = 0;; ++r) {
            rc = enqueue_chunk(k);
            if (rc) break;
            iters += k->chunk;
            hipError_t e = hipEventRecord(ev[r & 1], k->stream);
            if (e != hipSuccess) {
                rc = hip_fail(e, "batch event");
                break;
            }
            if (r >= 1) {
                e = hipEventSynchronize(ev[(r - 1) & 1]);
                if (e != hipSuccess) {
                    rc = hip_fail(e, "batch event sync");
                    break;
                }
                if (__atomic_load_n(&k->hostStat[0], __ATOMIC_ACQUIRE)) break;
            }
            if (iters > maxIter) {
                set_error("batch: LTS macro step did not complete within %lld driver iterations", maxIter);
                rc = SWE_HIP_ETIMEOUT;
                break;
            }
        }
        const hipError_t e = hipStreamSynchronize(k->stream);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "batch sync");
    }
    if (itersOut) *itersOut = iters;
    return 0;
}

// Apply the pending updates of a fused run and bring every block back to
// buffer 0 (downloads, snapshots and the next upload see one layout).
int finish_run(swe_hip_batch* k) {
    if (!k->fused) return 0;
    const dim3 grid(64, k->nblocks);
    hipLaunchKernelGGL(batch_normalize_kernel, grid, dim3(256), 0, k->stream, k->A, k->dDesc, k->dState);
    hipLaunchKernelGGL(batch_gather_kernel, grid, dim3(256), 0, k->stream, k->A, k->dDesc, k->dState);
    hipLaunchKernelGGL(batch_flags_kernel, dim3(1), dim3(256), 0, k->stream, k->dState, k->nblocks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(k->stream));
    return 0;
}

}  // namespace

extern "C" {

int swe_hip_batch_create(int nblocks, const int* nx, const int* ny, const float* dx, const float* dy, const int* bnd,
                         const int* nbr, int solver, int mode, float max_timestep_local, float duration,
                         swe_hip_batch** out) {
    if (!out || nblocks < 1 || nblocks > kMaxBatchBlocks || !nx || !ny || !dx || !dy || !bnd || !nbr || solver < 0 ||
        solver > 3 || mode < BATCH_ENSEMBLE || mode > BATCH_LTS) {
        set_error("swe_hip_batch_create: invalid argument (nblocks=%d solver=%d mode=%d)", nblocks, solver, mode);
        return SWE_HIP_EINVAL;
    }
    if (mode == BATCH_LTS && !(max_timestep_local > 0.f)) {
        set_error("swe_hip_batch_create: LTS needs max_timestep_local > 0");
        return SWE_HIP_EINVAL;
    }
    std::vector<BatchBlockDesc> desc(nblocks);
    long long off = 0, lines = 0, items = 0;
    for (int k = 0; k < nblocks; ++k) {
        BatchBlockDesc& d = desc[k];
        std::memset(&d, 0, sizeof(d));
        if (nx[k] < 1 || ny[k] < 1 || !(dx[k] > 0.f) || !(dy[k] > 0.f) ||
            (long long)(nx[k] + 2) * (ny[k] + 2) > (1LL << 31)) {
            set_error("swe_hip_batch_create: block %d: invalid size or mesh (nx=%d ny=%d dx=%g dy=%g)", k, nx[k],
                      ny[k], dx[k], dy[k]);
            return SWE_HIP_EINVAL;
        }
        d.nx = nx[k];
        d.ny = ny[k];
        d.pitch = ny[k] + 2;
        d.dx = dx[k];
        d.dy = dy[k];
        d.off = off;
        off += (long long)(nx[k] + 2) * (ny[k] + 2);
        const int nyt = (ny[k] + kStripRows - 1) / kStripRows;
        const long long tiles4 = (long long)nyt * ((nx[k] + 3) / 4);
        if (tiles4 > kLineTile) {
            set_error("swe_hip_batch_create: block %d too large for the work queue", k);
            return SWE_HIP_EINVAL;
        }
        items += (mode == BATCH_LTS) ? std::max<long long>(tiles4, 4) : tiles4;  // LTS: or its sides' steps
        for (int s = 0; s < 4; ++s) {
            const int t = bnd[4 * k + s];
            if (t != OUTFLOW && t != WALL && t != CONNECT_WITHIN_RANK && t != PASSIVE) {
                set_error("swe_hip_batch_create: block %d side %d: boundary type %d not supported", k, s, t);
                return SWE_HIP_EINVAL;
            }
            d.bnd[s] = t;
            d.nbr[s] = nbr[4 * k + s];
            d.line[s] = 0;
        }
    }
    // neighbour consistency: the edge is shared whole and the link is mutual
    for (int k = 0; k < nblocks; ++k) {
        for (int s = 0; s < 4; ++s) {
            if (desc[k].bnd[s] != CONNECT_WITHIN_RANK) continue;
            const int m = desc[k].nbr[s];
            const bool vertical = s == BND_LEFT || s == BND_RIGHT;
            if (m < 0 || m >= nblocks || m == k || desc[m].bnd[s ^ 1] != CONNECT_WITHIN_RANK ||
                desc[m].nbr[s ^ 1] != k || (vertical ? desc[m].ny != desc[k].ny : desc[m].nx != desc[k].nx)) {
                set_error("swe_hip_batch_create: block %d side %d: neighbour %d does not share the edge", k, s, m);
                return SWE_HIP_EINVAL;
            }
            if (mode == BATCH_ENSEMBLE) {
                set_error("swe_hip_batch_create: ensemble blocks cannot be connected (block %d)", k);
                return SWE_HIP_EINVAL;
            }
            if (mode == BATCH_LTS) {
                desc[k].line[s] = lines;
                lines += 3LL * (vertical ? desc[k].ny : desc[k].nx);
            }
        }
    }
    if (items >= (1LL << 31)) {
        set_error("swe_hip_batch_create: too many work items");
        return SWE_HIP_EINVAL;
    }
    swe_hip_batch* k = new (std::nothrow) swe_hip_batch();
    if (!k) {
        set_error("swe_hip_batch_create: out of host memory");
        return SWE_HIP_EINVAL;
    }
    k->nblocks = nblocks;
    k->solver = solver;
    k->mode = mode;
    k->T = max_timestep_local;
    k->duration = duration;
    k->desc = desc;
    k->total = off;
    k->lineTotal = lines;
    k->itemCap = (int)items;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
        k->cus = prop.multiProcessorCount;
    const size_t bytes = sizeof(float) * (size_t)off;
    hipError_t e = hipStreamCreateWithFlags(&k->stream, hipStreamNonBlocking);
    float** fields[15] = {&k->A.h[0],    &k->A.hu[0],   &k->A.hv[0],   &k->A.h[1],    &k->A.hu[1],
                          &k->A.hv[1],   &k->A.b,       &k->A.s[0][0], &k->A.s[0][1], &k->A.s[0][2],
                          &k->A.s[0][3], &k->A.s[1][0], &k->A.s[1][1], &k->A.s[1][2], &k->A.s[1][3]};
    for (int i = 0; i < 15 && e == hipSuccess; ++i) e = hipMalloc(fields[i], bytes);
    if (e == hipSuccess) e = hipMalloc(&k->A.lines, sizeof(float) * (size_t)std::max<long long>(lines, 1));
    if (e == hipSuccess) e = hipMalloc(&k->dDesc, sizeof(BatchBlockDesc) * nblocks);
    if (e == hipSuccess) e = hipMalloc(&k->dState, sizeof(BatchBlockState) * nblocks);
    if (e == hipSuccess) e = hipMalloc(&k->dG, sizeof(BatchGlobals));
    if (e == hipSuccess) e = hipMalloc(&k->dItems, sizeof(int) * (size_t)std::max(k->itemCap, 1));
    if (e == hipSuccess) e = hipMalloc(&k->dGeom, sizeof(GhostGeom) * 4 * (size_t)nblocks);
    if (e == hipSuccess) e = hipMalloc(&k->dJobs, sizeof(GhostJob) * 4 * (size_t)nblocks);
    if (e == hipSuccess) e = hipMemsetAsync(k->dJobs, 0, sizeof(GhostJob) * 4 * (size_t)nblocks, k->stream);
    if (e == hipSuccess) {
        // the fixed part of every side's ghost-line job
        std::vector<GhostGeom> geo(4 * (size_t)nblocks);
        for (int b = 0; b < nblocks; ++b) {
            const BatchBlockDesc& d = desc[b];
            for (int side = 0; side < 4; ++side) {
                GhostGeom& g = geo[4 * (size_t)b + side];
                std::memset(&g, 0, sizeof(g));
                if (d.bnd[side] != CONNECT_WITHIN_RANK) continue;
                const BatchBlockDesc& dn = desc[d.nbr[side]];
                long long so, go;
                int ss, gs, n, nn;
                side_line(dn, side ^ 1, 0, so, ss, nn);  // neighbour's copy layer at the shared edge
                side_line(d, side, 1, go, gs, n);        // the block's ghost line
                g.src = dn.off + so;
                g.dst = d.off + go;
                g.line = d.line[side];
                g.n = n;
                g.srcStride = ss;
                g.dstStride = gs;
                g.ndx = dn.dx;
                g.ndy = dn.dy;
            }
        }
        e = hipMemcpyAsync(k->dGeom, geo.data(), sizeof(GhostGeom) * geo.size(), hipMemcpyHostToDevice, k->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(k->stream);
    }
    if (e == hipSuccess) e = hipHostMalloc((void**)&k->hostStat, 4 * sizeof(int), hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&k->dHostStat, k->hostStat, 0);
    if (e == hipSuccess) {
        for (int i = 3; i < 15 && e == hipSuccess; ++i)
            if (i != 6) e = hipMemsetAsync(*fields[i], 0, bytes, k->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(k->stream);
    if (e != hipSuccess) {
        hip_fail(e, "swe_hip_batch_create");
        swe_hip_batch_destroy(k);
        return SWE_HIP_EDEVICE;
    }
    int rc = tiles_for(k, k->xs, true);
    if (rc == 0) rc = ensure_log(k, 1024);
    if (rc) {
        swe_hip_batch_destroy(k);
        return rc;
    }
    k->groups = 0;  // the direct form: fastest on the LTS tiling and the ensemble (swe_hip_batch_autotune)
    *out = k;
    return SWE_HIP_OK;
}

int swe_hip_batch_destroy(swe_hip_batch* k) {
    if (!k) return SWE_HIP_OK;
    if (k->stream) (void)hipStreamSynchronize(k->stream);
    if (k->net) swe::mlp_detach(k->net, k->stream);
    free_graph(k);
    float* fields[16] = {k->A.h[0],    k->A.hu[0],   k->A.hv[0],   k->A.h[1],    k->A.hu[1],   k->A.hv[1],
                         k->A.b,       k->A.s[0][0], k->A.s[0][1], k->A.s[0][2], k->A.s[0][3], k->A.s[1][0],
                         k->A.s[1][1], k->A.s[1][2], k->A.s[1][3], k->A.lines};
    for (float* f : fields)
        if (f) (void)hipFree(f);
    if (k->dDesc) (void)hipFree(k->dDesc);
    if (k->dState) (void)hipFree(k->dState);
    if (k->dG) (void)hipFree(k->dG);
    if (k->dItems) (void)hipFree(k->dItems);
    if (k->dGeom) (void)hipFree(k->dGeom);
    if (k->dJobs) (void)hipFree(k->dJobs);
    if (k->deferQ) (void)hipFree(k->deferQ);
    if (k->dDtLog) (void)hipFree(k->dDtLog);
    if (k->dIterLog) (void)hipFree(k->dIterLog);
    if (k->hostStat) (void)hipHostFree(k->hostStat);
    if (k->stream) (void)hipStreamDestroy(k->stream);
    delete k;
    return SWE_HIP_OK;
}

int swe_hip_batch_upload(swe_hip_batch* k, const float* h, const float* hu, const float* hv, const float* b) {
    if (!k || !h || !hu || !hv || !b) {
        set_error("swe_hip_batch_upload: null argument");
        return SWE_HIP_EINVAL;
    }
    const size_t bytes = sizeof(float) * (size_t)k->total;
    HIP_TRY(hipMemcpyAsync(k->A.h[0], h, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->A.hu[0], hu, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->A.hv[0], hv, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->A.b, b, bytes, hipMemcpyHostToDevice, k->stream));
    std::vector<BatchBlockState> st(k->nblocks);
    for (auto& s : st) {
        std::memset(&s, 0, sizeof(s));
        for (int i = 0; i < 4; ++i) s.gl[i] = GL_NEXT;  // SWE_Block.hh:297-300
    }
    HIP_TRY(hipMemcpyAsync(k->dState, st.data(), sizeof(BatchBlockState) * st.size(), hipMemcpyHostToDevice,
                           k->stream));
    BatchGlobals g;
    std::memset(&g, 0, sizeof(g));
    g.T = k->T;
    g.duration = k->duration;
    g.mode = k->mode;
    g.nblocks = k->nblocks;
    g.stop = 1;
    g.mlp = k->mlp;
    g.fused = k->fused;
    HIP_TRY(hipMemcpyAsync(k->dG, &g, sizeof(g), hipMemcpyHostToDevice, k->stream));
    hipLaunchKernelGGL(batch_init_kernel, dim3(k->nblocks), dim3(kCtlThreads), 0, k->stream, k->A, k->dDesc);
    hipLaunchKernelGGL(batch_prepare_kernel, dim3(1), dim3(kCtlThreads), 0, k->stream, k->dDesc, k->dState,
                       k->nblocks, k->mode, k->duration);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(k->stream));
    k->uploaded = true;
    return SWE_HIP_OK;
}

int swe_hip_batch_download(swe_hip_batch* k, float* h, float* hu, float* hv, float* b) {
    if (!k || !h || !hu || !hv) {
        set_error("swe_hip_batch_download: null argument");
        return SWE_HIP_EINVAL;
    }
    const size_t bytes = sizeof(float) * (size_t)k->total;
    HIP_TRY(hipMemcpyAsync(h, k->A.h[0], bytes, hipMemcpyDeviceToHost, k->stream));
    HIP_TRY(hipMemcpyAsync(hu, k->A.hu[0], bytes, hipMemcpyDeviceToHost, k->stream));
    HIP_TRY(hipMemcpyAsync(hv, k->A.hv[0], bytes, hipMemcpyDeviceToHost, k->stream));
    if (b) HIP_TRY(hipMemcpyAsync(b, k->A.b, bytes, hipMemcpyDeviceToHost, k->stream));
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_batch_run(swe_hip_batch* k, int units, float* dt_log, int* iter_log, float* elapsed_ms) {
    if (!k || units < 0) {
        set_error("swe_hip_batch_run: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (!k->uploaded) {
        set_error("swe_hip_batch_run: no state uploaded");
        return SWE_HIP_ESTATE;
    }
    if (units == 0) {
        if (elapsed_ms) *elapsed_ms = 0.f;
        return SWE_HIP_OK;
    }
    int rc = ensure_log(k, units);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, k->stream));
    }
    const long long maxIter = (long long)units * 4 * (kMaxDivisor + 1) + 2LL * k->chunk;
    rc = advance(k, units, maxIter, nullptr);
    if (rc == 0) rc = finish_run(k);
    if (elapsed_ms && rc == 0) {
        HIP_TRY(hipEventRecord(e1, k->stream));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(elapsed_ms, e0, e1));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (dt_log) HIP_TRY(hipMemcpy(dt_log, k->dDtLog, sizeof(float) * units, hipMemcpyDeviceToHost));
    if (iter_log) HIP_TRY(hipMemcpy(iter_log, k->dIterLog, sizeof(int) * units, hipMemcpyDeviceToHost));
    return SWE_HIP_OK;
}

int swe_hip_batch_stats(swe_hip_batch* k, int64_t* updates, float* t) {
    if (!k) {
        set_error("swe_hip_batch_stats: null handle");
        return SWE_HIP_EINVAL;
    }
    std::vector<BatchBlockState> st(k->nblocks);
    HIP_TRY(hipMemcpyAsync(st.data(), k->dState, sizeof(BatchBlockState) * st.size(), hipMemcpyDeviceToHost,
                           k->stream));
    HIP_TRY(hipStreamSynchronize(k->stream));
    for (int i = 0; i < k->nblocks; ++i) {
        if (updates) updates[i] = st[i].updates;
        if (t) t[i] = st[i].t;
    }
    return SWE_HIP_OK;
}

int swe_hip_batch_set_tuning(swe_hip_batch* k, int xs, int groups, int chunk, int graph, int fused) {
    if (!k || (xs != -1 && xs != 4 && xs != 8 && xs != 16) || (groups != -1 && (groups < 0 || groups > 65535)) ||
        (chunk != -1 && (chunk < 1 || chunk > 256)) || (graph != -1 && graph != 0 && graph != 1) ||
        (fused != -1 && fused != 0 && fused != 1)) {
        set_error("swe_hip_batch_set_tuning: invalid argument (xs=%d groups=%d chunk=%d graph=%d fused=%d)", xs, groups,
                  chunk, graph, fused);
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipStreamSynchronize(k->stream));
    if (fused != -1 && fused != k->fused) {
        // the state between runs has no pending update, so the strategy may change here
        HIP_TRY(hipMemcpy(&k->dG->fused, &fused, sizeof(int), hipMemcpyHostToDevice));
        k->fused = fused;
    }
    if (xs != -1 && xs != k->xs) {
        int rc = tiles_for(k, xs, true);
        if (rc) return rc;
        k->xs = xs;
    }
    if (groups != -1) k->groups = groups;
    if (chunk != -1) k->chunk = chunk;
    if (graph != -1) k->useGraph = graph;
    free_graph(k);
    return SWE_HIP_OK;
}

int swe_hip_batch_set_surrogate(swe_hip_batch* k, swe_hip_mlp* m) {
    if (!k) {
        set_error("swe_hip_batch_set_surrogate: null handle");
        return SWE_HIP_EINVAL;
    }
    if (m && k->solver != SOLVER_AUGRIE) {
        set_error("swe_hip_batch_set_surrogate: the learned middle state serves AugRie batches only");
        return SWE_HIP_EINVAL;
    }
    const unsigned* frags = m ? swe::mlp_frags(m) : nullptr;
    HIP_TRY(hipStreamSynchronize(k->stream));
    HIP_TRY(hipMemcpy(&k->dG->mlp, &frags, sizeof(frags), hipMemcpyHostToDevice));
    if (k->net != m) {
        if (k->net) swe::mlp_detach(k->net, k->stream);
        if (m) swe::mlp_attach(m, k->stream);
        k->net = m;
    }
    k->mlp = frags;
    free_graph(k);  // the captured sweep launches the other instantiation
    if (m && !k->deferQ) {
        // queue for the fused sweep's deferred edges: one entry per edge it
        // solves; entries hold x < 2^20 beside the block index (< 2^11)
        size_t edges = 0;
        bool fits = k->nblocks <= (1 << (31 - kDeferBlockShift));
        for (const auto& d : k->desc) {
            edges += (size_t)(d.nx + 1) * d.ny + (size_t)d.nx * (d.ny + 1);
            fits = fits && d.nx + 1 < (1 << kDeferBlockShift);
        }
        if (fits && edges <= 0xffffffffull) {  // otherwise the sweep solves the hard edges itself
            HIP_TRY(hipMalloc(&k->deferQ, edges * sizeof(uint2)));
            k->deferCap = (unsigned)edges;
        }
    }
    return SWE_HIP_OK;
}

int swe_hip_batch_get_tuning(swe_hip_batch* k, int* xs, int* groups, int* chunk, int* graph, int* fused) {
    if (!k) {
        set_error("swe_hip_batch_get_tuning: null handle");
        return SWE_HIP_EINVAL;
    }
    if (xs) *xs = k->xs;
    if (groups) *groups = k->groups;
    if (chunk) *chunk = k->chunk;
    if (graph) *graph = k->useGraph;
    if (fused) *fused = k->fused;
    return SWE_HIP_OK;
}

int swe_hip_batch_autotune(swe_hip_batch* k, int iterations, float* cand_ms, int* ncand) {
    if (!k || iterations < 1) {
        set_error("swe_hip_batch_autotune: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (!k->uploaded) {
        set_error("swe_hip_batch_autotune: no state uploaded");
        return SWE_HIP_ESTATE;
    }
    // snapshot of everything an iteration changes
    const size_t bytes = sizeof(float) * (size_t)k->total;
    const size_t lbytes = sizeof(float) * (size_t)std::max<long long>(k->lineTotal, 1);
    float *sh = nullptr, *shu = nullptr, *shv = nullptr, *sl = nullptr;
    BatchBlockState* sst = nullptr;
    BatchGlobals* sg = nullptr;
    hipError_t e = hipMalloc(&sh, bytes);
    if (e == hipSuccess) e = hipMalloc(&shu, bytes);
    if (e == hipSuccess) e = hipMalloc(&shv, bytes);
    if (e == hipSuccess) e = hipMalloc(&sl, lbytes);
    if (e == hipSuccess) e = hipMalloc(&sst, sizeof(BatchBlockState) * k->nblocks);
    if (e == hipSuccess) e = hipMalloc(&sg, sizeof(BatchGlobals));
    auto copy_all = [&](bool save) -> hipError_t {
        hipError_t r = hipSuccess;
        auto cp = [&](void* a, void* b, size_t n) {
            if (r == hipSuccess) r = hipMemcpyAsync(save ? a : b, save ? b : a, n, hipMemcpyDeviceToDevice, k->stream);
        };
        cp(sh, k->A.h[0], bytes);
        cp(shu, k->A.hu[0], bytes);
        cp(shv, k->A.hv[0], bytes);
        cp(sl, k->A.lines, lbytes);
        cp(sst, k->dState, sizeof(BatchBlockState) * k->nblocks);
        cp(sg, k->dG, sizeof(BatchGlobals));
        if (r == hipSuccess) r = hipStreamSynchronize(k->stream);
        return r;
    };
    bool saved = false;
    if (e == hipSuccess) {
        e = copy_all(true);
        saved = e == hipSuccess;
    }
    const int xsC[3] = {4, 8, 16};
    const int gC[4] = {2, 4, 8, 0};  // x CUs; 0 = the direct form
    const int oldXs = k->xs, oldGroups = k->groups, oldChunk = k->chunk;
    int bestXs = oldXs, bestGroups = oldGroups, n = 0, rc = 0;
    float best = FLT_MAX;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int units = 1 << 30;  // never reached: the timed iterations run through macro-step ends
    for (int a = 0; a < 3 && e == hipSuccess && rc == 0; ++a) {
        for (int g = 0; g < 4 && e == hipSuccess && rc == 0; ++g) {
            rc = swe_hip_batch_set_tuning(k, xsC[a], gC[g] * k->cus, std::min(iterations, 256), -1, -1);
            if (rc) break;
            // warm-up chunk (also captures the graph), then the timed chunks
            hipLaunchKernelGGL(batch_begin_kernel, dim3(1), dim3(kCtlThreads), 0, k->stream, k->dDesc, k->dState,
                               k->dG, k->dItems, units, k->logCap, k->dHostStat, k->dJobs);
            rc = enqueue_chunk(k);
            if (rc) break;
            e = hipEventRecord(e0, k->stream);
            int done = 0;
            while (rc == 0 && e == hipSuccess && done < iterations) {
                rc = enqueue_chunk(k);
                done += k->chunk;
            }
            if (rc || e != hipSuccess) break;
            e = hipEventRecord(e1, k->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) break;
            const float per = ms / done;
            if (cand_ms) cand_ms[n] = per;
            ++n;
            if (per < best) {
                best = per;
                bestXs = xsC[a];
                bestGroups = gC[g] * k->cus;
            }
            e = copy_all(false);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    // the snapshot goes back also after a failed timed run, which leaves the
    // batch half-advanced; the first error is the one reported
    if (saved) {
        const hipError_t first = e;
        e = copy_all(false);
        if (first != hipSuccess) e = first;
    }
    for (void* p : {(void*)sh, (void*)shu, (void*)shv, (void*)sl, (void*)sst, (void*)sg})
        if (p) (void)hipFree(p);
    if (rc || e != hipSuccess) {
        (void)swe_hip_batch_set_tuning(k, oldXs, oldGroups, oldChunk, -1, -1);
        return rc ? rc : hip_fail(e, "swe_hip_batch_autotune");
    }
    if (ncand) *ncand = n;
    return swe_hip_batch_set_tuning(k, bestXs, bestGroups, oldChunk, -1, -1);
}

}  // extern "C"
