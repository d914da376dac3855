// oracle_block.hpp -- CPU ORACLE (test infrastructure only).
//
// Block-level loops of the SWE dimensional-splitting step, restated on the
// CPU from the reference's C++ text, templated on the scalar type.
//
// Layout (reference src/tools/Float2D.hh:74-80): every field is an
// (nx+2) x (ny+2) array with a[x][y] = raw[x*(ny+2) + y]; the interior is
// x, y in [1, n], ghost cells sit at 0 and n+1.
//
//   apply_boundary_conditions   SWE_Block.hh:757-911
//   apply_boundary_bathymetry   SWE_Block.hh:724-749
//   fluxes_u / update_u         SWE_DimensionalSplittingMpi.cpp:383-518
//                               (flavour U: every parallel block)
//   fluxes_s / update_s         SWE_DimensionalSplitting.cpp:190-286
//                               (flavour S: swe_simple)
//   compute_max_timestep        SWE_Block.hh:693-717
#pragma once

#include <cfloat>
#include <cstddef>
#include <limits>
#include <vector>

#include "oracle_solvers.hpp"

namespace oracle {

enum BoundaryType { OUTFLOW = 0, WALL = 1, INFLOW = 2, CONNECT = 3, CONNECT_WITHIN_RANK = 4, PASSIVE = 5 };
enum Boundary { BND_LEFT = 0, BND_RIGHT = 1, BND_BOTTOM = 2, BND_TOP = 3 };

template <typename T>
struct Field {
    T* p;
    int rows;  // ny + 2
    T* operator[](int x) { return p + (size_t)x * rows; }
    const T* operator[](int x) const { return p + (size_t)x * rows; }
};

// Net-update scratch, sized like the reference's members
// (SWE_DimensionalSplittingMpi.cpp:80-91).
template <typename T>
struct NetUpdates {
    int nx, ny;
    std::vector<T> hL, hR, huL, huR;  // (nx+2) x (ny+2)
    std::vector<T> hB, hA, hvB, hvA;  // (nx+1) x (ny+2)
    NetUpdates(int nx_, int ny_) : nx(nx_), ny(ny_) {
        const size_t a = (size_t)(nx + 2) * (ny + 2);
        const size_t c = (size_t)(nx + 1) * (ny + 2);
        hL.assign(a, T(0));
        hR.assign(a, T(0));
        huL.assign(a, T(0));
        huR.assign(a, T(0));
        hB.assign(c, T(0));
        hA.assign(c, T(0));
        hvB.assign(c, T(0));
        hvA.assign(c, T(0));
    }
    Field<T> f(std::vector<T>& v) { return Field<T>{v.data(), ny + 2}; }
};

template <typename T>
void apply_boundary_conditions(int nx, int ny, const int* bnd, T* hp, T* hup, T* hvp) {
    Field<T> h{hp, ny + 2}, hu{hup, ny + 2}, hv{hvp, ny + 2};
    // left
    if (bnd[BND_LEFT] == WALL) {
        for (int j = 1; j <= ny; j++) {
            h[0][j] = h[1][j];
            hu[0][j] = -hu[1][j];
            hv[0][j] = hv[1][j];
        }
    } else if (bnd[BND_LEFT] == OUTFLOW) {
        for (int j = 1; j <= ny; j++) {
            h[0][j] = h[1][j];
            hu[0][j] = hu[1][j];
            hv[0][j] = hv[1][j];
        }
    }
    // right
    if (bnd[BND_RIGHT] == WALL) {
        for (int j = 1; j <= ny; j++) {
            h[nx + 1][j] = h[nx][j];
            hu[nx + 1][j] = -hu[nx][j];
            hv[nx + 1][j] = hv[nx][j];
        }
    } else if (bnd[BND_RIGHT] == OUTFLOW) {
        for (int j = 1; j <= ny; j++) {
            h[nx + 1][j] = h[nx][j];
            hu[nx + 1][j] = hu[nx][j];
            hv[nx + 1][j] = hv[nx][j];
        }
    }
    // bottom
    if (bnd[BND_BOTTOM] == WALL) {
        for (int i = 1; i <= nx; i++) {
            h[i][0] = h[i][1];
            hu[i][0] = hu[i][1];
            hv[i][0] = -hv[i][1];
        }
    } else if (bnd[BND_BOTTOM] == OUTFLOW) {
        for (int i = 1; i <= nx; i++) {
            h[i][0] = h[i][1];
            hu[i][0] = hu[i][1];
            hv[i][0] = hv[i][1];
        }
    }
    // top
    if (bnd[BND_TOP] == WALL) {
        for (int i = 1; i <= nx; i++) {
            h[i][ny + 1] = h[i][ny];
            hu[i][ny + 1] = hu[i][ny];
            hv[i][ny + 1] = -hv[i][ny];
        }
    } else if (bnd[BND_TOP] == OUTFLOW) {
        for (int i = 1; i <= nx; i++) {
            h[i][ny + 1] = h[i][ny];
            hu[i][ny + 1] = hu[i][ny];
            hv[i][ny + 1] = hv[i][ny];
        }
    }
    // corners always take the diagonal interior cell
    const int cx[4] = {0, 0, nx + 1, nx + 1};
    const int cy[4] = {0, ny + 1, 0, ny + 1};
    const int sx[4] = {1, 1, nx, nx};
    const int sy[4] = {1, ny, 1, ny};
    for (int k = 0; k < 4; k++) {
        h[cx[k]][cy[k]] = h[sx[k]][sy[k]];
        hu[cx[k]][cy[k]] = hu[sx[k]][sy[k]];
        hv[cx[k]][cy[k]] = hv[sx[k]][sy[k]];
    }
}

template <typename T>
void apply_boundary_bathymetry(int nx, int ny, const int* bnd, T* bp) {
    Field<T> b{bp, ny + 2};
    auto ow = [](int t) { return t == OUTFLOW || t == WALL; };
    if (ow(bnd[BND_LEFT]))
        for (int j = 0; j < ny + 2; j++) b[0][j] = b[1][j];
    if (ow(bnd[BND_RIGHT]))
        for (int j = 0; j < ny + 2; j++) b[nx + 1][j] = b[nx][j];
    if (ow(bnd[BND_BOTTOM]))
        for (int i = 0; i <= nx + 1; i++) b[i][0] = b[i][1];
    if (ow(bnd[BND_TOP]))
        for (int i = 0; i <= nx + 1; i++) b[i][ny + 1] = b[i][ny];
    b[0][0] = b[1][1];
    b[0][ny + 1] = b[1][ny];
    b[nx + 1][0] = b[nx][1];
    b[nx + 1][ny + 1] = b[nx][ny];
}

// Flavour U fluxes: every vertical and horizontal edge computed from the
// same state.  Returns the maximum edge speed.
template <typename T>
T fluxes_u(int solver, int nx, int ny, const T* hp, const T* hup, const T* hvp, const T* bp,
           NetUpdates<T>& nu, int nthreads) {
    const int rows = ny + 2;
    Field<T> h{const_cast<T*>(hp), rows}, hu{const_cast<T*>(hup), rows},
        hv{const_cast<T*>(hvp), rows}, b{const_cast<T*>(bp), rows};
    Field<T> L = nu.f(nu.hL), R = nu.f(nu.hR), huL = nu.f(nu.huL), huR = nu.f(nu.huR);
    Field<T> B = nu.f(nu.hB), A = nu.f(nu.hA), hvB = nu.f(nu.hvB), hvA = nu.f(nu.hvA);
    T maxWaveSpeed = T(0);
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads) reduction(max : maxWaveSpeed) schedule(static)
    for (int i = 1; i < nx + 2; i++) {
        for (int j = 1; j < ny + 1; j++) {
            T s;
            solve_edge<T>(solver, h[i - 1][j], h[i][j], hu[i - 1][j], hu[i][j], b[i - 1][j], b[i][j],
                          L[i - 1][j - 1], R[i - 1][j - 1], huL[i - 1][j - 1], huR[i - 1][j - 1], s);
            maxWaveSpeed = std::max(maxWaveSpeed, s);
        }
    }
#pragma omp parallel for num_threads(nthreads) reduction(max : maxWaveSpeed) schedule(static)
    for (int i = 1; i < nx + 1; i++) {
        for (int j = 1; j < ny + 2; j++) {
            T s;
            solve_edge<T>(solver, h[i][j - 1], h[i][j], hv[i][j - 1], hv[i][j], b[i][j - 1], b[i][j],
                          B[i - 1][j - 1], A[i - 1][j - 1], hvB[i - 1][j - 1], hvA[i - 1][j - 1], s);
            maxWaveSpeed = std::max(maxWaveSpeed, s);
        }
    }
    return maxWaveSpeed;
}

// dt rule of flavour U (SWE_DimensionalSplittingMpi.cpp:448-456).
template <typename T>
T timestep_u(T maxWaveSpeed, T dx, T dy) {
    if (maxWaveSpeed > T(0.00001)) {
        T dt = std::min(dx / maxWaveSpeed, dy / maxWaveSpeed);
        dt *= T(0.4);
        return dt;
    }
    return std::numeric_limits<T>::max();
}

// Flavour U update with the dry clamp (SWE_DimensionalSplittingMpi.cpp:484-518).
template <typename T>
void update_u(int nx, int ny, T dx, T dy, T dt, T* hp, T* hup, T* hvp, NetUpdates<T>& nu,
              int nthreads) {
    const int rows = ny + 2;
    Field<T> h{hp, rows}, hu{hup, rows}, hv{hvp, rows};
    Field<T> L = nu.f(nu.hL), R = nu.f(nu.hR), huL = nu.f(nu.huL), huR = nu.f(nu.huR);
    Field<T> B = nu.f(nu.hB), A = nu.f(nu.hA), hvB = nu.f(nu.hvB), hvA = nu.f(nu.hvA);
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 1; i < nx + 1; i++) {
        for (int j = 1; j < ny + 1; j++) {
            h[i][j] -= dt / dx * (R[i - 1][j - 1] + L[i][j - 1]) + dt / dy * (A[i - 1][j - 1] + B[i - 1][j]);
            hu[i][j] -= dt / dx * (huR[i - 1][j - 1] + huL[i][j - 1]);
            hv[i][j] -= dt / dy * (hvA[i - 1][j - 1] + hvB[i - 1][j]);
            if (h[i][j] < T(0)) {
                h[i][j] = hu[i][j] = hv[i][j] = T(0);
            } else if (h[i][j] < T(0.1)) {
                hu[i][j] = hv[i][j] = T(0);
            }
        }
    }
}

// Flavour S (swe_simple): x-sweep over all rows incl. ghosts, dt from the
// x-speeds only, intermediate Q*, y-sweep on the OLD h and hv, update with
// dx used for the y-part as the reference does
// (SWE_DimensionalSplitting.cpp:190-286).
//
// dt rule: 0.4 * dx / (the maximum edge speed of the x-sweep).  The
// reference passes its OpenMP reduction variable itself as the solver's
// speed output (SWE_DimensionalSplitting.cpp:111-136), and the solver
// overwrites that argument with each edge's speed instead of raising it, so
// each thread's copy ends as the speed of the last edge that thread solved,
// and dt depends on the thread count and the loop partition.  This
// restatement (and the GPU) takes the true maximum over all x-edges, the
// value the reduction is written to compute (with one thread the
// reference's dt would come from the last x-edge alone);
// tests/test_oracle_scheme.py pins it with a known-answer case.
template <typename T>
struct SplitScratch {
    int nx, ny;
    std::vector<T> hL, hR, huL, huR;  // (nx+2) x (ny+2)
    std::vector<T> hB, hA, hvB, hvA;  // (nx+1) x (ny+2)
    std::vector<T> hStar, huStar;     // (nx+1) x (ny+2)
    SplitScratch(int nx_, int ny_) : nx(nx_), ny(ny_) {
        const size_t a = (size_t)(nx + 2) * (ny + 2);
        const size_t c = (size_t)(nx + 1) * (ny + 2);
        hL.assign(a, T(0));
        hR.assign(a, T(0));
        huL.assign(a, T(0));
        huR.assign(a, T(0));
        hB.assign(c, T(0));
        hA.assign(c, T(0));
        hvB.assign(c, T(0));
        hvA.assign(c, T(0));
        hStar.assign(c, T(0));
        huStar.assign(c, T(0));
    }
};

template <typename T>
T fluxes_s(int solver, int nx, int ny, T dx, const T* hp, const T* hup, const T* hvp, const T* bp,
           SplitScratch<T>& s, T* maxVertical, int nthreads) {
    const int rows = ny + 2;
    Field<T> h{const_cast<T*>(hp), rows}, hu{const_cast<T*>(hup), rows},
        hv{const_cast<T*>(hvp), rows}, b{const_cast<T*>(bp), rows};
    Field<T> L{s.hL.data(), rows}, R{s.hR.data(), rows}, huL{s.huL.data(), rows}, huR{s.huR.data(), rows};
    Field<T> B{s.hB.data(), rows}, A{s.hA.data(), rows}, hvB{s.hvB.data(), rows}, hvA{s.hvA.data(), rows};
    Field<T> hS{s.hStar.data(), rows}, huS{s.huStar.data(), rows};
    T maxH = T(0), maxV = T(0);
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads) reduction(max : maxH) schedule(static)
    for (int x = 0; x < nx + 1; x++) {
        for (int y = 0; y < ny + 2; y++) {
            T sp;
            solve_edge<T>(solver, h[x][y], h[x + 1][y], hu[x][y], hu[x + 1][y], b[x][y], b[x + 1][y],
                          L[x][y], R[x + 1][y], huL[x][y], huR[x + 1][y], sp);
            maxH = std::max(maxH, sp);
        }
    }
    const T dt = T(0.4) * (dx / maxH);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int x = 1; x < nx + 1; x++) {
        for (int y = 0; y < ny + 2; y++) {
            hS[x][y] = h[x][y] - (dt / dx) * (L[x][y] + R[x][y]);
            huS[x][y] = hu[x][y] - (dt / dx) * (huL[x][y] + huR[x][y]);
        }
    }
#pragma omp parallel for num_threads(nthreads) reduction(max : maxV) schedule(static)
    for (int x = 1; x < nx + 1; x++) {
        for (int y = 0; y < ny + 1; y++) {
            T sp;
            solve_edge<T>(solver, h[x][y], h[x][y + 1], hv[x][y], hv[x][y + 1], b[x][y], b[x][y + 1],
                          B[x][y], A[x][y + 1], hvB[x][y], hvA[x][y + 1], sp);
            maxV = std::max(maxV, sp);
        }
    }
    if (maxVertical) *maxVertical = maxV;
    return dt;
}

template <typename T>
void update_s(int nx, int ny, T dx, T dt, T* hp, T* hup, T* hvp, SplitScratch<T>& s, int nthreads) {
    const int rows = ny + 2;
    Field<T> h{hp, rows}, hu{hup, rows}, hv{hvp, rows};
    Field<T> B{s.hB.data(), rows}, A{s.hA.data(), rows}, hvB{s.hvB.data(), rows}, hvA{s.hvA.data(), rows};
    Field<T> hS{s.hStar.data(), rows}, huS{s.huStar.data(), rows};
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int x = 1; x < nx + 1; x++) {
        for (int y = 1; y < ny + 1; y++) {
            h[x][y] = hS[x][y] - (dt / dx) * (B[x][y] + A[x][y]);
            hu[x][y] = huS[x][y];
            hv[x][y] = hv[x][y] - (dt / dx) * (hvB[x][y] + hvA[x][y]);
        }
    }
}

// Cell-based CFL estimate used to initialise local time stepping
// (SWE_Block.hh:693-717).
template <typename T>
T compute_max_timestep(int nx, int ny, T dx, T dy, const T* hp, const T* hup, const T* hvp, T dryTol,
                       T cfl) {
    const int rows = ny + 2;
    Field<T> h{const_cast<T*>(hp), rows}, hu{const_cast<T*>(hup), rows}, hv{const_cast<T*>(hvp), rows};
    T maxSpeed = T(0);
    for (int i = 1; i < nx + 1; i++) {
        for (int j = 1; j < ny + 1; j++) {
            if (h[i][j] > dryTol) {
                T momentum = std::max(std::fabs(hu[i][j]), std::fabs(hv[i][j]));
                T particleVelocity = momentum / h[i][j];
                T waveSpeed = particleVelocity + std::sqrt(T(9.81) * h[i][j]);
                maxSpeed = std::max(maxSpeed, waveSpeed);
            }
        }
    }
    T dt = std::min(dx, dy) / maxSpeed;
    dt *= cfl;
    return dt;
}

}  // namespace oracle
