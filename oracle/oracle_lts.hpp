// oracle_lts.hpp -- CPU ORACLE (test infrastructure only).
//
// Multi-block driver of the overdecomposed SWE run (many blocks in one rank,
// neighbours connected in memory), with global time stepping or with the
// reference's multi-rate local time stepping (LTS), restated from:
//   SWE_Block.hh:311-562            LTS state machine (checkAllGhostlayers,
//                                   interpolate/copyGhostlayer,
//                                   getRoundTimestep, getTotalLocalTimestep,
//                                   hasMaxLocalTimestep, resetStepSizeCounter,
//                                   isReceivable, allGhostlayersInSync)
//   SWE_DimensionalSplittingMPIOverdecomp.cpp:208-235   in-rank bathymetry copy
//   SWE_DimensionalSplittingMPIOverdecomp.cpp:269-304   in-rank ghost copy
//   SWE_DimensionalSplittingMPIOverdecomp.cpp:487-606   fluxes / update
//   swe_mpi_overdecomp.cpp:263-343                      driver loop
// The reference's quirks are kept: a macro step performs stepSize+1 updates
// of T/stepSize (the counter only starts moving once stepSize is known), the
// ghost interpolation weight uses absolute local times, and the power-of-two
// divisor is derived from the absolute step size.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_block.hpp"

namespace oracle {

enum GhostlayerState { GL_DONE = 0, GL_UNVALID = 1, GL_INTER = 2, GL_SYNC = 3, GL_NEXT = 4 };

template <typename T>
struct LtsBlock {
    int nx = 0, ny = 0;
    T dx = 0, dy = 0;
    std::vector<T> h, hu, hv, b;
    std::vector<T> bufH, bufHu, bufHv;  // separate storage when LTS is on
    int bnd[4] = {OUTFLOW, OUTFLOW, OUTFLOW, OUTFLOW};
    int nb[4] = {-1, -1, -1, -1};
    int gl[4] = {GL_NEXT, GL_NEXT, GL_NEXT, GL_NEXT};
    T borderTs[4] = {0, 0, 0, 0};
    int stepSize = 0, stepSizeCounter = 0, timestepCounter = 0, iteration = 0;
    int maxDivisor = 1024;
    T maxTimestepLocal = 1, duration = 0, maxTimestep = 0;
    bool lts = false, notified = false;
    int64_t updates = 0;  // interior updates performed (for cell-update counts)

    int rows() const { return ny + 2; }
    T& at(std::vector<T>& a, int x, int y) { return a[(size_t)x * rows() + y]; }
    T* bh() { return lts ? bufH.data() : h.data(); }
    T* bhu() { return lts ? bufHu.data() : hu.data(); }
    T* bhv() { return lts ? bufHv.data() : hv.data(); }

    T total() const {
        return (T)(maxTimestepLocal * timestepCounter) +
               (stepSize > 0 ? ((T)(maxTimestepLocal * stepSizeCounter) / (T)stepSize) : T(0));
    }
    bool inSync() const {
        return gl[BND_LEFT] != GL_UNVALID && gl[BND_RIGHT] != GL_UNVALID && gl[BND_TOP] != GL_UNVALID &&
               gl[BND_BOTTOM] != GL_UNVALID && (stepSizeCounter <= stepSize);
    }
    bool hasMaxLocal() const { return (stepSizeCounter >= stepSize && inSync()) || notified; }
    bool receivable(int side) const {
        if (notified) return false;
        return gl[side] == GL_NEXT || gl[side] == GL_UNVALID || gl[side] == GL_SYNC;
    }
    void resetStepSizeCounter() {
        stepSizeCounter = 0;
        stepSize = 0;
        timestepCounter++;
    }
    // SWE_Block.hh:536-555, literally (float log, double log(2), pow, ceil)
    T roundTimestep(T ts) {
        if (stepSizeCounter <= 0) {
            if (ts > maxTimestepLocal) ts = maxTimestepLocal;
            if (ts < (T)maxTimestepLocal / maxDivisor) ts = (T)maxTimestepLocal / maxDivisor;
            int divisor = (int)std::pow(2, -(std::ceil(std::log(ts) / std::log(2))));
            if ((T)1.f / divisor > ts) divisor *= 2;
            stepSize = divisor;
            if (stepSize > maxDivisor) stepSize = maxDivisor;
        }
        return (T)maxTimestepLocal / stepSize;
    }
    void copyGhost(int side) {
        if (side == BND_LEFT || side == BND_RIGHT) {
            const int x = side == BND_LEFT ? 0 : nx + 1;
            for (int i = 1; i < ny + 1; i++) {
                at(h, x, i) = at(bufH, x, i);
                at(hu, x, i) = at(bufHu, x, i);
                at(hv, x, i) = at(bufHv, x, i);
            }
        } else {
            const int y = side == BND_BOTTOM ? 0 : ny + 1;
            for (int i = 1; i < nx + 1; i++) {
                at(h, i, y) = at(bufH, i, y);
                at(hu, i, y) = at(bufHu, i, y);
                at(hv, i, y) = at(bufHv, i, y);
            }
        }
    }
    T interp(T oldv, T newv, T remote) const { return (oldv + (newv - oldv) * (total() / remote)); }
    void interpolateGhost(int side, T remote) {
        if (side == BND_LEFT || side == BND_RIGHT) {
            const int x = side == BND_LEFT ? 0 : nx + 1;
            for (int i = 1; i < ny + 1; i++) {
                at(h, x, i) = interp(at(h, x, i), at(bufH, x, i), remote);
                at(hu, x, i) = interp(at(hu, x, i), at(bufHu, x, i), remote);
                at(hv, x, i) = interp(at(hv, x, i), at(bufHv, x, i), remote);
            }
        } else {
            const int y = side == BND_BOTTOM ? 0 : ny + 1;
            for (int i = 1; i < nx + 1; i++) {
                at(h, i, y) = interp(at(h, i, y), at(bufH, i, y), remote);
                at(hu, i, y) = interp(at(hu, i, y), at(bufHu, i, y), remote);
                at(hv, i, y) = interp(at(hv, i, y), at(bufHv, i, y), remote);
            }
        }
    }
    // SWE_Block.hh:375-436
    void checkAllGhostlayers() {
        if (!lts) return;
        for (int border = BND_LEFT; border <= BND_TOP; border++) {
            if (bnd[border] == CONNECT || bnd[border] == CONNECT_WITHIN_RANK) {
                if (std::fabs(borderTs[border] - total()) <= std::numeric_limits<double>::epsilon()) {
                    copyGhost(border);
                    gl[border] = GL_SYNC;
                } else if (borderTs[border] > total()) {
                    interpolateGhost(border, borderTs[border]);
                    gl[border] = GL_INTER;
                } else {
                    gl[border] = GL_UNVALID;
                }
                if (borderTs[border] >= duration) gl[border] = GL_DONE;
            }
        }
        if (std::fabs(duration - total()) <= std::numeric_limits<double>::epsilon() || total() >= duration)
            notified = true;
        if (inSync()) {
            for (int s = 0; s < 4; s++)
                if (gl[s] != GL_DONE) gl[s] = (gl[s] == GL_INTER) ? GL_INTER : GL_NEXT;
        }
        if (inSync()) {
            if (stepSizeCounter < stepSize) stepSizeCounter++;
        }
        iteration++;
    }
};

template <typename T>
struct MultiBlock {
    std::vector<LtsBlock<T>> blocks;
    int solver = SOLVER_AUGRIE;
    int nthreads = 1;
    std::vector<NetUpdates<T>> nus;

    void prepare() {
        nus.clear();
        for (auto& bl : blocks) nus.emplace_back(bl.nx, bl.ny);
    }

    // SWE_DimensionalSplittingMPIOverdecomp.cpp:208-235
    void sendBathymetry() {
        for (auto& me : blocks) {
            if (me.bnd[BND_RIGHT] == CONNECT_WITHIN_RANK) {
                LtsBlock<T>& r = blocks[me.nb[BND_RIGHT]];
                for (int i = 1; i < me.ny + 1; i++) me.at(me.b, me.nx + 1, i) = r.at(r.b, 1, i);
            }
            if (me.bnd[BND_LEFT] == CONNECT_WITHIN_RANK) {
                LtsBlock<T>& l = blocks[me.nb[BND_LEFT]];
                for (int i = 1; i < me.ny + 1; i++) me.at(me.b, 0, i) = l.at(l.b, l.nx, i);
            }
            if (me.bnd[BND_TOP] == CONNECT_WITHIN_RANK) {
                LtsBlock<T>& t = blocks[me.nb[BND_TOP]];
                for (int i = 1; i < me.nx + 1; i++) me.at(me.b, i, me.ny + 1) = t.at(t.b, i, 1);
            }
            if (me.bnd[BND_BOTTOM] == CONNECT_WITHIN_RANK) {
                LtsBlock<T>& bo = blocks[me.nb[BND_BOTTOM]];
                for (int i = 1; i < me.nx + 1; i++) me.at(me.b, i, 0) = bo.at(bo.b, i, bo.ny);
            }
        }
    }

    // SWE_DimensionalSplittingMPIOverdecomp.cpp:269-304
    void setGhostLayer(LtsBlock<T>& me) {
        apply_boundary_conditions<T>(me.nx, me.ny, me.bnd, me.h.data(), me.hu.data(), me.hv.data());
        const int rows = me.rows();
        if (me.bnd[BND_RIGHT] == CONNECT_WITHIN_RANK && me.receivable(BND_RIGHT)) {
            LtsBlock<T>& r = blocks[me.nb[BND_RIGHT]];
            me.borderTs[BND_RIGHT] = r.total();
            for (int i = 1; i < me.ny + 1; i++) {
                me.bh()[(size_t)(me.nx + 1) * rows + i] = r.at(r.h, 1, i);
                me.bhu()[(size_t)(me.nx + 1) * rows + i] = r.at(r.hu, 1, i);
                me.bhv()[(size_t)(me.nx + 1) * rows + i] = r.at(r.hv, 1, i);
            }
        }
        if (me.bnd[BND_LEFT] == CONNECT_WITHIN_RANK && me.receivable(BND_LEFT)) {
            LtsBlock<T>& l = blocks[me.nb[BND_LEFT]];
            me.borderTs[BND_LEFT] = l.total();
            for (int i = 1; i < me.ny + 1; i++) {
                me.bh()[i] = l.at(l.h, l.nx, i);
                me.bhu()[i] = l.at(l.hu, l.nx, i);
                me.bhv()[i] = l.at(l.hv, l.nx, i);
            }
        }
        if (me.bnd[BND_TOP] == CONNECT_WITHIN_RANK && me.receivable(BND_TOP)) {
            LtsBlock<T>& t = blocks[me.nb[BND_TOP]];
            me.borderTs[BND_TOP] = t.total();
            for (int i = 1; i < me.nx + 1; i++) {
                me.bh()[(size_t)i * rows + me.ny + 1] = t.at(t.h, i, 1);
                me.bhu()[(size_t)i * rows + me.ny + 1] = t.at(t.hu, i, 1);
                me.bhv()[(size_t)i * rows + me.ny + 1] = t.at(t.hv, i, 1);
            }
        }
        if (me.bnd[BND_BOTTOM] == CONNECT_WITHIN_RANK && me.receivable(BND_BOTTOM)) {
            LtsBlock<T>& bo = blocks[me.nb[BND_BOTTOM]];
            me.borderTs[BND_BOTTOM] = bo.total();
            for (int i = 1; i < me.nx + 1; i++) {
                me.bh()[(size_t)i * rows + 0] = bo.at(bo.h, i, bo.ny);
                me.bhu()[(size_t)i * rows + 0] = bo.at(bo.hu, i, bo.ny);
                me.bhv()[(size_t)i * rows + 0] = bo.at(bo.hv, i, bo.ny);
            }
        }
    }

    void computeFluxes(size_t k) {
        LtsBlock<T>& me = blocks[k];
        if (!me.inSync()) return;
        T s = fluxes_u<T>(solver, me.nx, me.ny, me.h.data(), me.hu.data(), me.hv.data(), me.b.data(), nus[k],
                          nthreads);
        me.maxTimestep = timestep_u<T>(s, me.dx, me.dy);
    }

    void updateUnknowns(size_t k) {
        LtsBlock<T>& me = blocks[k];
        if (!me.inSync()) return;
        update_u<T>(me.nx, me.ny, me.dx, me.dy, me.maxTimestep, me.h.data(), me.hu.data(), me.hv.data(), nus[k],
                    nthreads);
        me.updates++;
    }

    // One pass of the driver's inner do-loop (swe_mpi_overdecomp.cpp:266-336).
    // Returns the global dt in global mode.
    T iterate(bool lts) {
        for (auto& bl : blocks) setGhostLayer(bl);
        for (auto& bl : blocks) bl.checkAllGhostlayers();
        for (size_t k = 0; k < blocks.size(); k++) computeFluxes(k);
        T dt = T(0);
        if (!lts) {
            dt = std::numeric_limits<T>::max();
            for (auto& bl : blocks) dt = std::min(dt, bl.maxTimestep);
            for (auto& bl : blocks) bl.maxTimestep = dt;
        } else {
            for (auto& bl : blocks)
                if (bl.inSync()) bl.maxTimestep = bl.roundTimestep(bl.maxTimestep);
        }
        for (size_t k = 0; k < blocks.size(); k++) updateUnknowns(k);
        return dt;
    }

    // One macro step of LTS (or one global step); returns the number of
    // driver iterations, or -1 when the iteration cap is hit.
    int macroStep(bool lts, T* dtOut, int maxIter = 1 << 20) {
        int it = 0;
        bool synced = true;
        do {
            T dt = iterate(lts);
            if (dtOut) *dtOut = dt;
            it++;
            if (lts) {
                synced = true;
                for (auto& bl : blocks)
                    if (!bl.hasMaxLocal()) synced = false;
            }
            if (it >= maxIter) return -1;
        } while (lts && !synced);
        if (lts)
            for (auto& bl : blocks) bl.resetStepSizeCounter();
        return it;
    }
};

}  // namespace oracle
