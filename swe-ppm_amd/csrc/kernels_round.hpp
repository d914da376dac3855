// kernels_round.hpp -- the batched engine's round: per-block dt, the global
// minimum, the completion test, the LTS bookkeeping of the next iteration, its
// ghost-line jobs and its work queue.  Part of kernels_batch.hpp (included
// there after the engine's structures); batch_round_kernel runs it as one
// workgroup of kRoundThreads after the sweep.
#pragma once

namespace swe {

constexpr int kRoundThreads = 1024;

// ---- LTS bookkeeping of one block at the start of an iteration -------------
// All that setGhostLayer / receiveGhostLayer and checkAllGhostlayers decide
// at the start of an iteration follows from the blocks' times at the end of
// the previous one (BatchBlockState::total): which sides receive and the
// neighbour's time stamp they take (SWE_Block.hh:311-374), then the
// ghost-layer states, notifiedLastTimestep, the sub-step counter and whether
// the block is in sync (SWE_Block.hh:375-436).  The lines launch moves the
// data afterwards.
template <typename Desc, typename NeighbourTotal>
__device__ void lts_transitions(const Desc& d, BatchBlockState& s, const NeighbourTotal& neighbourTotal,
                                float duration) {
    const float tot = s.total;
    int recv = 0;
    for (int side = 0; side < 4; ++side) {
        if (d.bnd[side] != CONNECT_WITHIN_RANK || !lts_receivable(s, side)) continue;
        recv |= 1 << side;
        s.borderTs[side] = neighbourTotal(d.nbr[side]);
    }
    s.recv = recv;
    for (int side = 0; side < 4; ++side) {
        if (d.bnd[side] != CONNECT && d.bnd[side] != CONNECT_WITHIN_RANK) continue;
        const float bt = s.borderTs[side];
        if ((double)fabsf(bt - tot) <= DBL_EPSILON)
            s.gl[side] = GL_SYNC;
        else if (bt > tot)
            s.gl[side] = GL_INTER;
        else
            s.gl[side] = GL_UNVALID;
        if (bt >= duration) s.gl[side] = GL_DONE;
    }
    if ((double)fabsf(duration - tot) <= DBL_EPSILON || tot >= duration) s.notified = 1;
    if (lts_in_sync(s))
        for (int side = 0; side < 4; ++side)
            if (s.gl[side] != GL_DONE) s.gl[side] = (s.gl[side] == GL_INTER) ? GL_INTER : GL_NEXT;
    if (lts_in_sync(s) && s.stepSizeCounter < s.stepSize) s.stepSizeCounter++;
    s.active = lts_in_sync(s) ? 1 : 0;
}

// exclusive prefix sum over the workgroup's NT threads (full waves)
template <int NT>
__device__ int wg_exclusive_scan(int v, int& total) {
    __shared__ int sWave[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sWave[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
    for (int k = 0; k < NT / 64; ++k) {
        const int t = sWave[k];
        if (k < wave) before += t;
        all += t;
    }
    __syncthreads();
    total = all;
    return before + x - v;
}

// the queue entries of one block: its strip tiles when it computes, else the
// ghost-line steps of the sides in `lineMask` (fused strategy, LTS)
__device__ __forceinline__ int queue_entries(bool active, int ntiles, int lineMask) {
    return active ? ntiles : __popc(lineMask);
}
// one wave writes them
__device__ __forceinline__ void queue_write(int* items, int base, int b, bool active, int ntiles, int lineMask,
                                            int lane) {
    if (active) {
        for (int i = lane; i < ntiles; i += 64) items[base + i] = (b << kTileBits) | i;
    } else if (lane < 4 && ((lineMask >> lane) & 1)) {
        items[base + __popc(lineMask & ((1 << lane) - 1))] = (b << kTileBits) | (kLineTile + lane);
    }
}

// ---- ghost-line job of one (block, side) -------------------------------------
// From the block's state and its neighbour's after the round (buffer, sums
// set, pending flag and dt packed as line_flags / maxTimestep): global dt
// copies the neighbour's copy layer into the ghost line; LTS takes it into the
// side's receive line when the side receives, then copies (time stamps equal)
// or interpolates (the neighbour ahead: weight t_me / t_neighbour) the ghost
// line from the receive line (SWE_Block.hh:375-436).  A skipped iteration
// gets no job.
__device__ __forceinline__ int line_flags(const BatchBlockState& s) { return s.cur | (s.pend << 1) | (s.sidx << 2); }

__device__ __forceinline__ GhostJob line_job(const BatchBlockState& s, int side, int bnd, bool lts, bool skip,
                                            int nflags, float ndt) {
    GhostJob j{};
    if (skip || bnd != CONNECT_WITHIN_RANK) return j;
    int op = kLineReceive | (1 << kLineActShift);  // global dt: copy into the ghost line
    if (lts) {
        const bool receive = (s.recv >> side) & 1;
        const float bt = s.borderTs[side], tot = s.total;
        int act = 0;
        if ((double)fabsf(bt - tot) <= DBL_EPSILON) {
            act = 1;
        } else if (bt > tot) {
            act = 2;
            j.w = tot / bt;
        }
        if (!receive && !act) return j;
        op = (receive ? kLineReceive : 0) | kLineKeep | (act << kLineActShift);
    }
    j.op = op;
    j.srcBuf = nflags & 1;
    j.pend = (nflags >> 1) & 1;
    j.sidx = (nflags >> 2) & 1;
    j.ndt = ndt;
    j.dstBuf = s.cur;
    return j;
}

struct RoundDesc {
    float dx, dy;
    int bnd[4], nbr[4];
    int ntiles;
};

// ---- the round: per-block dt (SWE_DimensionalSplittingMpi.cpp:448-463), the
// global minimum (swe_mpi_overdecomp.cpp:297-307), the completion test
// (hasMaxLocalTimestep, SWE_Block.hh:355-358, swe_mpi_overdecomp.cpp:327-342)
// and resetStepSizeCounter (SWE_Block.hh:530-535); then every block's time,
// the next iteration's LTS bookkeeping, its ghost-line jobs and, with the
// fused strategy, its queue.  NT threads (tid 0..NT-1) handle MAXB blocks at
// most; each keeps the state of its blocks (blk = tid + r * NT) in registers
// from one load to one store, and what crosses blocks (neighbour times and
// buffers, queue offsets) goes through LDS: a short chain of dependent
// accesses.
template <int NT, int MAXB>
__device__ __forceinline__ void round_body(int tid, const BatchBlockDesc* D, BatchBlockState* st, BatchGlobals* G,
                                           const RoundIO& io) {
    constexpr int PER = (MAXB + NT - 1) / NT;
    if (tid == 0) G->deferCount = 0u;  // the deferred edges of this iteration are solved
    const int mode = G->mode;
    const float T = G->T;
    const int nb = G->nblocks;
    const bool fused = G->fused != 0;
    const float duration = G->duration;
    __shared__ unsigned int sminBits;
    __shared__ int allSynced, macroEnd, skipNext;
    __shared__ float sTotal[MAXB];
    __shared__ int sBase[MAXB];
    __shared__ int sTiles[MAXB];
    __shared__ int sMask[MAXB];
    __shared__ int sFlags[MAXB];
    __shared__ float sDt[MAXB];
    BatchBlockState S[PER];
    RoundDesc R[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int blk = tid + r * NT;
        if (blk >= nb) continue;
        S[r] = st[blk];
        const BatchBlockDesc& d = D[blk];
        R[r].dx = d.dx;
        R[r].dy = d.dy;
        for (int k = 0; k < 4; ++k) {
            R[r].bnd[k] = d.bnd[k];
            R[r].nbr[k] = d.nbr[k];
        }
        R[r].ntiles = d.ntiles;
    }
    if (tid == 0) {
        sminBits = __float_as_uint(FLT_MAX);
        allSynced = 1;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        if (tid + r * NT >= nb) continue;
        BatchBlockState& s = S[r];
        if (s.active) {
            float dt = timestep_from_max(__uint_as_float(s.smax), R[r].dx, R[r].dy);
            if (mode == BATCH_LTS) {
                if (s.stepSizeCounter <= 0) s.stepSize = lts_step_size(T, dt);
                dt = T / s.stepSize;
            }
            if (mode == BATCH_ENSEMBLE) s.t += dt;
            s.updates += 1;
            s.maxTimestep = dt;
            if (mode == BATCH_GLOBAL) atomicMin(&sminBits, __float_as_uint(dt));  // dt > 0: bits order as floats
        }
        s.smax = 0u;
        if (mode == BATCH_LTS) {
            const bool has = (s.stepSizeCounter >= s.stepSize && lts_in_sync(s)) || s.notified;
            if (!has) atomicAnd(&allSynced, 0);
        }
    }
    __syncthreads();
    if (tid == 0) {
        G->iterations += 1;
        G->itersUnit += 1;
        if (mode == BATCH_GLOBAL) G->gdt = __uint_as_float(sminBits);
        const bool unitDone = (mode != BATCH_LTS) || allSynced;
        if (unitDone) {
            const int u = G->done;
            if (u < G->logCap) {
                io.dtLog[u] = (mode == BATCH_GLOBAL) ? __uint_as_float(sminBits) : 0.f;
                io.iterLog[u] = G->itersUnit;
            }
            G->done = u + 1;
            G->itersUnit = 0;
            if (u + 1 >= G->target) G->stop = 1;
        }
        macroEnd = (mode == BATCH_LTS) && allSynced;
        skipNext = G->stop;
        io.hostStat[0] = G->stop;
        io.hostStat[1] = G->done;
        io.hostStat[2] = G->itersUnit;
    }
    __syncthreads();
    const float gdt = __uint_as_float(sminBits);
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int blk = tid + r * NT;
        if (blk >= nb) continue;
        BatchBlockState& s = S[r];
        if (s.active) {
            // every block of the global loop updates with the minimum (swe_mpi_overdecomp.cpp:305)
            if (mode == BATCH_GLOBAL) s.maxTimestep = gdt;
            // fused strategy: the update is applied by the block's next load, into the other buffer
            if (fused) {
                s.cur ^= 1;
                s.sidx ^= 1;
                s.pend = 1;
            }
        }
        if (macroEnd) {
            s.stepSizeCounter = 0;
            s.stepSize = 0;
            s.timestepCounter += 1;
        }
        if (mode == BATCH_LTS) sTotal[blk] = s.total = lts_total(T, s);
        sFlags[blk] = line_flags(s);
        sDt[blk] = s.maxTimestep;
    }
    if (mode != BATCH_ENSEMBLE) __syncthreads();  // every block's total, buffer and dt are in LDS
    if (mode == BATCH_LTS) {
#pragma unroll
        for (int r = 0; r < PER; ++r)
            if (tid + r * NT < nb) lts_transitions(R[r], S[r], [&](int k) { return sTotal[k]; }, duration);
    }
    int lineMask[PER];  // fused LTS: the sides of a waiting block whose ghost-line step is a queue item
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int blk = tid + r * NT;
        lineMask[r] = 0;
        if (blk >= nb) continue;
        st[blk] = S[r];
        if (mode == BATCH_ENSEMBLE) continue;
        for (int side = 0; side < 4; ++side) {
            const int m = R[r].nbr[side];
            const bool link = R[r].bnd[side] == CONNECT_WITHIN_RANK;
            const GhostJob j = line_job(S[r], side, R[r].bnd[side], mode == BATCH_LTS, skipNext != 0,
                                        link ? sFlags[m] : 0, link ? sDt[m] : 0.f);
            io.jobs[4 * blk + side] = j;
            if (j.op != 0 && !S[r].active && fused) lineMask[r] |= 1 << side;
        }
    }
    if (!fused) return;  // the store strategy's queue is built after its apply launch
    if (mode != BATCH_LTS) {  // every block computes every iteration: the queue stays as filled
        if (tid == 0) {
            G->count = G->allItems;
            G->head = 0;
            G->headApply = 0;
            G->skip = G->stop;
        }
        return;
    }
    // the next queue, block by block: the tiles of the blocks in sync, the
    // ghost-line steps of the others
    int mine = 0;
#pragma unroll
    for (int r = 0; r < PER; ++r)
        if (tid + r * NT < nb) mine += queue_entries(S[r].active != 0, R[r].ntiles, lineMask[r]);
    int total;
    int run = wg_exclusive_scan<NT>(mine, total);
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        const int blk = tid + r * NT;
        if (blk >= nb) continue;
        const bool a = S[r].active != 0;
        sBase[blk] = run;
        sTiles[blk] = a ? R[r].ntiles : 0;
        sMask[blk] = a ? 0 : lineMask[r];
        run += queue_entries(a, R[r].ntiles, lineMask[r]);
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    for (int b = wave; b < nb; b += NT / 64)
        queue_write(io.items, sBase[b], b, sTiles[b] > 0, sTiles[b], sMask[b], lane);
    if (tid == 0) {
        G->count = total;
        G->head = 0;
        G->headApply = 0;
        G->skip = G->stop;
    }
}

}  // namespace swe
