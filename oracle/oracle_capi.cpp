// oracle_capi.cpp -- CPU ORACLE (test infrastructure only).
//
// C entry points of the oracle for the Python tests (ctypes) and for the
// bench's CPU-baseline leg.  Nothing in the product links or loads this.
// Every function takes caller-owned arrays in the reference's Float2D layout
// (a[x][y] = raw[x*(ny+2)+y]) and returns 0 on success.
#include <cstdint>
#include <cstring>

#include "oracle_block.hpp"
#include "oracle_lts.hpp"

using namespace oracle;

namespace {

template <typename T>
int solver_batch(int solver, int64_t n, const T* hL, const T* hR, const T* huL, const T* huR, const T* bL,
                 const T* bR, T* hNetL, T* hNetR, T* huNetL, T* huNetR, T* maxSpeed) {
    for (int64_t k = 0; k < n; k++)
        solve_edge<T>(solver, hL[k], hR[k], huL[k], huR[k], bL[k], bR[k], hNetL[k], hNetR[k], huNetL[k],
                      huNetR[k], maxSpeed[k]);
    return 0;
}

template <typename T>
int middle_batch(int64_t n, const T* hL, const T* hR, const T* uL, const T* uR, T* hm, T* s1m, T* s2m,
                 int32_t* kind, int32_t* iters) {
    for (int64_t k = 0; k < n; k++) {
        int it = 0;
        int kd = middle_state<T>(hL[k], hR[k], uL[k], uR[k], hm[k], s1m[k], s2m[k], &it);
        if (kind) kind[k] = kd;
        if (iters) iters[k] = it;
    }
    return 0;
}

// nsteps of flavour U with global time stepping on one block:
// setGhostLayer -> computeNumericalFluxes -> getMaxTimestep -> updateUnknowns.
// dt_in (optional) forces the step sizes; dt_out (optional) records them.
template <typename T>
int run_u(int nx, int ny, T dx, T dy, const int* bnd, int solver, int nsteps, int nthreads, T* h, T* hu,
          T* hv, const T* b, const T* dt_in, T* dt_out) {
    NetUpdates<T> nu(nx, ny);
    for (int s = 0; s < nsteps; s++) {
        apply_boundary_conditions<T>(nx, ny, bnd, h, hu, hv);
        T smax = fluxes_u<T>(solver, nx, ny, h, hu, hv, b, nu, nthreads);
        T dt = dt_in ? dt_in[s] : timestep_u<T>(smax, dx, dy);
        if (dt_out) dt_out[s] = dt;
        update_u<T>(nx, ny, dx, dy, dt, h, hu, hv, nu, nthreads);
    }
    return 0;
}

template <typename T>
int run_s(int nx, int ny, T dx, const int* bnd, int solver, int nsteps, int nthreads, T* h, T* hu, T* hv,
          const T* b, T* dt_out) {
    SplitScratch<T> sc(nx, ny);
    for (int s = 0; s < nsteps; s++) {
        apply_boundary_conditions<T>(nx, ny, bnd, h, hu, hv);
        T dt = fluxes_s<T>(solver, nx, ny, dx, h, hu, hv, b, sc, nullptr, nthreads);
        if (dt_out) dt_out[s] = dt;
        update_s<T>(nx, ny, dx, dt, h, hu, hv, sc, nthreads);
    }
    return 0;
}

template <typename T>
int fluxes_u_out(int nx, int ny, int solver, const T* h, const T* hu, const T* hv, const T* b, T* L, T* R,
                 T* huL, T* huR, T* B, T* A, T* hvB, T* hvA, T* maxSpeed) {
    NetUpdates<T> nu(nx, ny);
    *maxSpeed = fluxes_u<T>(solver, nx, ny, h, hu, hv, b, nu, 1);
    const size_t a = (size_t)(nx + 2) * (ny + 2), c = (size_t)(nx + 1) * (ny + 2);
    std::memcpy(L, nu.hL.data(), a * sizeof(T));
    std::memcpy(R, nu.hR.data(), a * sizeof(T));
    std::memcpy(huL, nu.huL.data(), a * sizeof(T));
    std::memcpy(huR, nu.huR.data(), a * sizeof(T));
    std::memcpy(B, nu.hB.data(), c * sizeof(T));
    std::memcpy(A, nu.hA.data(), c * sizeof(T));
    std::memcpy(hvB, nu.hvB.data(), c * sizeof(T));
    std::memcpy(hvA, nu.hvA.data(), c * sizeof(T));
    return 0;
}

template <typename T>
int multiblock_run(int nblocks, const int* nx, const int* ny, T dx, T dy, const int* bnd, const int* nbr,
                   int solver, int lts, T maxTimestepLocal, T duration, int nmacro, int nthreads, T* h, T* hu,
                   T* hv, T* b, int64_t* updates_out, int32_t* iters_out, T* dt_out) {
    MultiBlock<T> mb;
    mb.solver = solver;
    mb.nthreads = nthreads;
    mb.blocks.resize(nblocks);
    size_t off = 0;
    std::vector<size_t> offs(nblocks);
    for (int k = 0; k < nblocks; k++) {
        LtsBlock<T>& bl = mb.blocks[k];
        bl.nx = nx[k];
        bl.ny = ny[k];
        bl.dx = dx;
        bl.dy = dy;
        const size_t n = (size_t)(nx[k] + 2) * (ny[k] + 2);
        offs[k] = off;
        bl.h.assign(h + off, h + off + n);
        bl.hu.assign(hu + off, hu + off + n);
        bl.hv.assign(hv + off, hv + off + n);
        bl.b.assign(b + off, b + off + n);
        bl.lts = lts != 0;
        if (bl.lts) {
            bl.bufH.assign(n, T(0));
            bl.bufHu.assign(n, T(0));
            bl.bufHv.assign(n, T(0));
        }
        for (int s = 0; s < 4; s++) {
            bl.bnd[s] = bnd[4 * k + s];
            bl.nb[s] = nbr[4 * k + s];
        }
        bl.maxTimestepLocal = maxTimestepLocal;
        bl.duration = duration;
        off += n;
    }
    mb.prepare();
    mb.sendBathymetry();
    int rc = 0;
    for (int m = 0; m < nmacro; m++) {
        T dt = 0;
        int it = mb.macroStep(lts != 0, &dt);
        if (iters_out) iters_out[m] = it;
        if (dt_out) dt_out[m] = dt;
        if (it < 0) {
            rc = -1;
            break;
        }
    }
    for (int k = 0; k < nblocks; k++) {
        LtsBlock<T>& bl = mb.blocks[k];
        std::copy(bl.h.begin(), bl.h.end(), h + offs[k]);
        std::copy(bl.hu.begin(), bl.hu.end(), hu + offs[k]);
        std::copy(bl.hv.begin(), bl.hv.end(), hv + offs[k]);
        std::copy(bl.b.begin(), bl.b.end(), b + offs[k]);
        if (updates_out) updates_out[k] = bl.updates;
    }
    return rc;
}

// Driver iterations of the overdecomposed LTS loop one at a time (the inner
// do-loop of swe_mpi_overdecomp.cpp:266-336, with resetStepSizeCounter when
// every block has its macro step), for tests that look inside a macro step.
// info receives per block, after the last iteration, kTraceInfo floats:
// total, stepSize, stepSizeCounter, timestepCounter, notified, gl[4],
// borderTs[4], updates.
constexpr int kTraceInfo = 14;

template <typename T>
int lts_trace(int nblocks, const int* nx, const int* ny, T dx, T dy, const int* bnd, const int* nbr, int solver,
              T maxTimestepLocal, T duration, int niter, T* h, T* hu, T* hv, T* b, T* info, int32_t* macroEnds) {
    MultiBlock<T> mb;
    mb.solver = solver;
    mb.blocks.resize(nblocks);
    size_t off = 0;
    std::vector<size_t> offs(nblocks);
    for (int k = 0; k < nblocks; k++) {
        LtsBlock<T>& bl = mb.blocks[k];
        bl.nx = nx[k];
        bl.ny = ny[k];
        bl.dx = dx;
        bl.dy = dy;
        const size_t n = (size_t)(nx[k] + 2) * (ny[k] + 2);
        offs[k] = off;
        bl.h.assign(h + off, h + off + n);
        bl.hu.assign(hu + off, hu + off + n);
        bl.hv.assign(hv + off, hv + off + n);
        bl.b.assign(b + off, b + off + n);
        bl.lts = true;
        bl.bufH.assign(n, T(0));
        bl.bufHu.assign(n, T(0));
        bl.bufHv.assign(n, T(0));
        for (int s = 0; s < 4; s++) {
            bl.bnd[s] = bnd[4 * k + s];
            bl.nb[s] = nbr[4 * k + s];
        }
        bl.maxTimestepLocal = maxTimestepLocal;
        bl.duration = duration;
        off += n;
    }
    mb.prepare();
    mb.sendBathymetry();
    int macro = 0;
    for (int it = 0; it < niter; it++) {
        mb.iterate(true);
        bool synced = true;
        for (auto& bl : mb.blocks)
            if (!bl.hasMaxLocal()) synced = false;
        if (synced) {
            for (auto& bl : mb.blocks) bl.resetStepSizeCounter();
            if (macroEnds) macroEnds[macro] = it + 1;
            macro++;
        }
    }
    for (int k = 0; k < nblocks; k++) {
        LtsBlock<T>& bl = mb.blocks[k];
        std::copy(bl.h.begin(), bl.h.end(), h + offs[k]);
        std::copy(bl.hu.begin(), bl.hu.end(), hu + offs[k]);
        std::copy(bl.hv.begin(), bl.hv.end(), hv + offs[k]);
        T* o = info + (size_t)k * kTraceInfo;
        o[0] = bl.total();
        o[1] = (T)bl.stepSize;
        o[2] = (T)bl.stepSizeCounter;
        o[3] = (T)bl.timestepCounter;
        o[4] = bl.notified ? T(1) : T(0);
        for (int s = 0; s < 4; s++) {
            o[5 + s] = (T)bl.gl[s];
            o[9 + s] = bl.borderTs[s];
        }
        o[13] = (T)bl.updates;
    }
    return macro;
}

}  // namespace

extern "C" {

int oracle_lts_trace_f32(int nblocks, const int* nx, const int* ny, float dx, float dy, const int* bnd,
                         const int* nbr, int solver, float maxTimestepLocal, float duration, int niter, float* h,
                         float* hu, float* hv, float* b, float* info, int32_t* macroEnds) {
    return lts_trace<float>(nblocks, nx, ny, dx, dy, bnd, nbr, solver, maxTimestepLocal, duration, niter, h, hu, hv,
                            b, info, macroEnds);
}

int oracle_solver_f32(int solver, int64_t n, const float* hL, const float* hR, const float* huL,
                      const float* huR, const float* bL, const float* bR, float* hNetL, float* hNetR,
                      float* huNetL, float* huNetR, float* maxSpeed) {
    return solver_batch<float>(solver, n, hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
}
int oracle_solver_f64(int solver, int64_t n, const double* hL, const double* hR, const double* huL,
                      const double* huR, const double* bL, const double* bR, double* hNetL, double* hNetR,
                      double* huNetL, double* huNetR, double* maxSpeed) {
    return solver_batch<double>(solver, n, hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
}
int oracle_middle_state_f32(int64_t n, const float* hL, const float* hR, const float* uL, const float* uR,
                            float* hm, float* s1m, float* s2m, int32_t* kind, int32_t* iters) {
    return middle_batch<float>(n, hL, hR, uL, uR, hm, s1m, s2m, kind, iters);
}
int oracle_middle_state_f64(int64_t n, const double* hL, const double* hR, const double* uL, const double* uR,
                            double* hm, double* s1m, double* s2m, int32_t* kind, int32_t* iters) {
    return middle_batch<double>(n, hL, hR, uL, uR, hm, s1m, s2m, kind, iters);
}
int oracle_apply_bc_f32(int nx, int ny, const int* bnd, float* h, float* hu, float* hv) {
    apply_boundary_conditions<float>(nx, ny, bnd, h, hu, hv);
    return 0;
}
int oracle_apply_bc_bathymetry_f32(int nx, int ny, const int* bnd, float* b) {
    apply_boundary_bathymetry<float>(nx, ny, bnd, b);
    return 0;
}
int oracle_run_u_f32(int nx, int ny, float dx, float dy, const int* bnd, int solver, int nsteps, int nthreads,
                     float* h, float* hu, float* hv, const float* b, const float* dt_in, float* dt_out) {
    return run_u<float>(nx, ny, dx, dy, bnd, solver, nsteps, nthreads, h, hu, hv, b, dt_in, dt_out);
}
int oracle_run_u_f64(int nx, int ny, double dx, double dy, const int* bnd, int solver, int nsteps,
                     int nthreads, double* h, double* hu, double* hv, const double* b, const double* dt_in,
                     double* dt_out) {
    return run_u<double>(nx, ny, dx, dy, bnd, solver, nsteps, nthreads, h, hu, hv, b, dt_in, dt_out);
}
int oracle_run_s_f32(int nx, int ny, float dx, const int* bnd, int solver, int nsteps, int nthreads, float* h,
                     float* hu, float* hv, const float* b, float* dt_out) {
    return run_s<float>(nx, ny, dx, bnd, solver, nsteps, nthreads, h, hu, hv, b, dt_out);
}
int oracle_run_s_f64(int nx, int ny, double dx, const int* bnd, int solver, int nsteps, int nthreads,
                     double* h, double* hu, double* hv, const double* b, double* dt_out) {
    return run_s<double>(nx, ny, dx, bnd, solver, nsteps, nthreads, h, hu, hv, b, dt_out);
}
int oracle_fluxes_u_f32(int nx, int ny, int solver, const float* h, const float* hu, const float* hv,
                        const float* b, float* L, float* R, float* huL, float* huR, float* B, float* A,
                        float* hvB, float* hvA, float* maxSpeed) {
    return fluxes_u_out<float>(nx, ny, solver, h, hu, hv, b, L, R, huL, huR, B, A, hvB, hvA, maxSpeed);
}
float oracle_compute_max_timestep_f32(int nx, int ny, float dx, float dy, const float* h, const float* hu,
                                      const float* hv, float dryTol, float cfl) {
    return compute_max_timestep<float>(nx, ny, dx, dy, h, hu, hv, dryTol, cfl);
}
float oracle_timestep_u_f32(float maxSpeed, float dx, float dy) { return timestep_u<float>(maxSpeed, dx, dy); }
int oracle_multiblock_run_f32(int nblocks, const int* nx, const int* ny, float dx, float dy, const int* bnd,
                              const int* nbr, int solver, int lts, float maxTimestepLocal, float duration,
                              int nmacro, int nthreads, float* h, float* hu, float* hv, float* b,
                              int64_t* updates_out, int32_t* iters_out, float* dt_out) {
    return multiblock_run<float>(nblocks, nx, ny, dx, dy, bnd, nbr, solver, lts, maxTimestepLocal, duration,
                                 nmacro, nthreads, h, hu, hv, b, updates_out, iters_out, dt_out);
}
int oracle_multiblock_run_f64(int nblocks, const int* nx, const int* ny, double dx, double dy, const int* bnd,
                              const int* nbr, int solver, int lts, double maxTimestepLocal, double duration,
                              int nmacro, int nthreads, double* h, double* hu, double* hv, double* b,
                              int64_t* updates_out, int32_t* iters_out, double* dt_out) {
    return multiblock_run<double>(nblocks, nx, ny, dx, dy, bnd, nbr, solver, lts, maxTimestepLocal, duration,
                                  nmacro, nthreads, h, hu, hv, b, updates_out, iters_out, dt_out);
}
// getRoundTimestep of a fresh block (stepSizeCounter == 0), for checking the
// product's divisor computation against the literal formula.
int oracle_round_timestep_f32(float maxTimestepLocal, float ts, float* rounded) {
    LtsBlock<float> bl;
    bl.maxTimestepLocal = maxTimestepLocal;
    *rounded = bl.roundTimestep(ts);
    return bl.stepSize;
}

}  // extern "C"
