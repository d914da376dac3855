// oracle_solvers.hpp -- CPU ORACLE (test infrastructure only).
//
// This file is the checker, never the product: only tests/, the smoke check
// and the bench's cpu_baseline leg may load what is built from it.
//
// Per-edge Riemann solvers of the SWE dimensional-splitting step, restated on
// the CPU, templated on the scalar type so that the same code gives the FP32
// oracle and the FP64 reference run.
//
// The reference includes these solvers from the git submodule
// TUM-I5/swe_solvers (reference .gitmodules:1-3), which is not vendored:
// submodules/swe_solvers/ is empty and its pinned commit is unknown.  The call
// interface is taken from the reference's call sites
//   solver.computeNetUpdates(hL, hR, huL, huR, bL, bR,
//                            hNetL, hNetR, huNetL, huNetR, maxEdgeSpeed)
// (src/blocks/SWE_DimensionalSplittingMpi.cpp:403-410, 432-439;
//  src/blocks/SWE_DimensionalSplitting.cpp:122-130), and the arithmetic is
// restated from the published methods:
//   * f-wave:  Bale, LeVeque, Mitran, Rossmanith, SIAM J. Sci. Comput. 24
//     (2002) 955-978, with Einfeldt (HLLE) speeds and the bathymetry source
//     folded into the flux jump, dry neighbours treated as reflecting walls;
//   * augmented Riemann solver:  D. L. George, J. Comput. Phys. 227 (2008)
//     3089-3113 ("Augmented Riemann solvers for the shallow water equations
//     over variable topography with steady states and inundation"), with the
//     wet/dry wall and inundation treatment of that paper and a Newton
//     middle-state computation for the shock cases.
//
// Stated constants (the un-vendored submodule's own values cannot be read):
//   g = 9.81, dryTol = 0.01, zeroTol = 1e-7, newtonTol = 1e-6 (relative),
//   maxNewton = 10, criticalTol = 1e-6, hybridTol = 1.0 (Hybrid dispatch:
//   f-wave when both depths >= hybridTol, augmented solver otherwise).
//
// Each function is pinned by the oracle tests (tests/test_oracle_solvers.py,
// tests/test_oracle_scheme.py): conservation of the wave decomposition, lake
// at rest, mirror symmetry, exact Riemann middle states (FP64 bisection),
// closed forms, dam-break convergence to the Stoker and Ritter solutions.
#pragma once

#include <algorithm>
#include <cmath>

namespace oracle {

enum SolverKind { SOLVER_HLLE = 0, SOLVER_FWAVE = 1, SOLVER_AUGRIE = 2, SOLVER_HYBRID = 3 };
// the augmented solver's decomposition without the still-edge path (for the
// tests that pin that path against it)
constexpr int SOLVER_AUGRIE_FULL = 12;

template <typename T>
struct SolverConsts {
    static constexpr T g = T(9.81);
    // sqrt(g): a cell's celerity is sqrtG * sqrt(h); in single precision the
    // correctly rounded square root of 9.81f, as the kernels use it
    static constexpr T sqrtG = sizeof(T) == sizeof(float) ? T(0x1.90e864p+1f) : T(0x1.90e8639c65e42p+1);
    static constexpr T dryTol = T(0.01);
    static constexpr T zeroTol = T(1e-7);
    static constexpr T newtonTol = T(1e-6);
    static constexpr int maxNewton = 10;
    static constexpr T criticalTol = T(1e-6);
    static constexpr T hybridTol = T(1.0);
    // still edges of the augmented solver (augrie_still): both depths at
    // least stillTol and within a factor stillRatio of each other
    static constexpr T stillTol = T(0.1);
    static constexpr T stillRatio = T(2);
};

// A cell's velocity and celerity.  The kernels compute them once per loaded
// cell as hu * (1/h) and sqrt(g) * sqrt(h) (the reciprocal and the root are
// shared by the cell's edges); the oracle follows that rounding.  Built with
// ORACLE_CANONICAL=1 it uses the textbook forms hu / h and sqrt(g h) instead,
// so that a test can bound what the kernels' rounding choice changes.
#ifndef ORACLE_CANONICAL
#define ORACLE_CANONICAL 0
#endif
template <typename T>
inline T cell_velocity(T hu, T h) {
    return ORACLE_CANONICAL ? hu / h : hu * (T(1) / h);
}
template <typename T>
inline T cell_celerity(T h) {
    return ORACLE_CANONICAL ? std::sqrt(SolverConsts<T>::g * h) : SolverConsts<T>::sqrtG * std::sqrt(h);
}

// ---------------------------------------------------------------------------
// Still edge of the augmented solver: both cells wet with at least stillTol
// of water (cells shallower than 0.1 m have their momentum zeroed by the dry
// clamp of the update, so zero momentum there says nothing), zero normal
// momentum on both sides, the same free surface h + b in the working
// precision, and depths within a factor stillRatio.  Such an edge is a lake
// at rest: its exact Riemann solution has no flux, so the solver returns zero
// net updates (+0) instead of the round-off of the full decomposition, and
// still water stays exactly still.
//
// Its speed is the one the full solver reports.  At rest uL = uR = 0 and the
// celerity-weighted average uHat is 0, so the Einfeldt speeds are
// sE1 = min(-cL, -cHat) and sE2 = max(cR, cHat).  The middle state lies
// between the two depths (the two-shock case with hL = hR returns hm = hL;
// otherwise it is a shock towards the shallower side and a rarefaction
// towards the deeper one, and the Newton iterates rise from hMin to the root
// below hMax), so with hMax <= 2 hMin, i.e. cMax <= sqrt(2) cMin:
//   hL >= hR: s2m = 2 cL - cm >= cL > 0 > sE1 and
//             s1m = 2 cL - 3 cm <= (2 sqrt(2) - 3) cR < 0 < sE2,
//   hL <  hR: s1m = cm - 2 cR <= -cR < 0 < sE2 and
//             s2m = 3 cm - 2 cR >= (3 - 2 sqrt(2)) cL > 0 > sE1,
// hence l1 = sE1, l3 = sE2 and the speed is max(|sE1|, |sE2|) with no middle
// state.  tests/test_oracle_solvers.py checks the speed against the full
// solver bit for bit, and the full solver's updates against zero to within
// its round-off, on random still edges.
// ---------------------------------------------------------------------------
template <typename T>
inline bool augrie_still(T hL, T hR, T huL, T huR, T bL, T bR) {
    typedef SolverConsts<T> C;
    return hL >= C::stillTol && hR >= C::stillTol && huL == T(0) && huR == T(0) && hL + bL == hR + bR &&
           hL <= C::stillRatio * hR && hR <= C::stillRatio * hL;
}

template <typename T>
inline T augrie_still_speed(T hL, T hR) {
    typedef SolverConsts<T> C;
    const T cL = cell_celerity(hL);
    const T cR = cell_celerity(hR);
    const T cHat = std::sqrt(C::g * T(0.5) * (hR + hL));
    const T sE1 = std::min(-cL, -cHat);
    const T sE2 = std::max(cR, cHat);
    return std::max(std::fabs(sE1), std::fabs(sE2));
}

// Add one f-wave (its h and hu components) to the side its speed points to;
// a wave of (near) zero speed is shared equally.
template <typename T>
inline void accumulate_wave(T speed, T zh, T zhu, T& hNetL, T& hNetR, T& huNetL, T& huNetR) {
    const T zt = SolverConsts<T>::zeroTol;
    if (speed < -zt) {
        hNetL += zh;
        huNetL += zhu;
    } else if (speed > zt) {
        hNetR += zh;
        huNetR += zhu;
    } else {
        hNetL += T(0.5) * zh;
        huNetL += T(0.5) * zhu;
        hNetR += T(0.5) * zh;
        huNetR += T(0.5) * zhu;
    }
}

// ---------------------------------------------------------------------------
// f-wave solver (Bale et al. 2002) with Einfeldt speeds.
// ---------------------------------------------------------------------------
template <typename T>
void fwave(T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR, T& huNetL, T& huNetR,
           T& maxSpeed) {
    typedef SolverConsts<T> C;
    hNetL = hNetR = huNetL = huNetR = T(0);
    maxSpeed = T(0);
    const bool dryL = hL < C::dryTol;
    const bool dryR = hR < C::dryTol;
    if (dryL && dryR) return;
    // a dry neighbour acts as a reflecting wall: mirror the wet state
    if (dryL) {
        hL = hR;
        huL = -huR;
        bL = bR;
    } else if (dryR) {
        hR = hL;
        huR = -huL;
        bR = bL;
    }
    // a cell's velocity and celerity from 1/h and sqrt(h)
    const T uL = cell_velocity(huL, hL);
    const T uR = cell_velocity(huR, hR);
    const T sqhL = std::sqrt(hL);
    const T sqhR = std::sqrt(hR);
    const T cL = cell_celerity(hL);
    const T cR = cell_celerity(hR);
    // Roe averages
    const T hRoe = T(0.5) * (hL + hR);
    const T uRoe = (uL * sqhL + uR * sqhR) / (sqhL + sqhR);
    const T cRoe = std::sqrt(C::g * hRoe);
    // Einfeldt speeds
    const T s1 = std::min(uL - cL, uRoe - cRoe);
    const T s2 = std::max(uR + cR, uRoe + cRoe);
    // flux jump with the bathymetry source folded in
    const T df1 = huR - huL;
    const T fR = huR * uR + T(0.5) * C::g * hR * hR;
    const T fL = huL * uL + T(0.5) * C::g * hL * hL;
    const T df2 = (fR - fL) + T(0.5) * C::g * (hL + hR) * (bR - bL);
    // eigen-coefficients of R = [[1,1],[s1,s2]]
    const T invDs = T(1) / (s2 - s1);
    const T beta1 = (s2 * df1 - df2) * invDs;
    const T beta2 = (df2 - s1 * df1) * invDs;
    accumulate_wave(s1, beta1, beta1 * s1, hNetL, hNetR, huNetL, huNetR);
    accumulate_wave(s2, beta2, beta2 * s2, hNetL, hNetR, huNetL, huNetR);
    // the dry side receives nothing
    if (dryL) {
        hNetL = T(0);
        huNetL = T(0);
    } else if (dryR) {
        hNetR = T(0);
        huNetR = T(0);
    }
    maxSpeed = std::max(std::fabs(s1), std::fabs(s2));
}

// ---------------------------------------------------------------------------
// Middle state of the Riemann problem (George 2008, sec. 3; Toro's exact
// solver for the shallow water equations).  Returns the middle depth hm and
// the speeds s1m (1-wave tail/shock at the middle state) and s2m.
// kind: 0 dry side, 1 two rarefactions, 2 two shocks, 3 shock+rarefaction.
//
// The depth solves phi(h) = du + f(h; hMin) + f(h; hMax) = 0 with the wave
// curves f(h; k) = (h - k) gk(h), gk(h) = sqrt(g/2 (1/h + 1/k)) (shock,
// h > k) and f(h; k) = 2 (sqrt(g h) - sqrt(g k)) (rarefaction).
//  * two rarefactions: closed form;
//  * two shocks: Newton from h = hMax, both curves shocks;
//  * shock + rarefaction: Newton from h = hMin, shock towards hMin.
// phi is increasing and concave on each search interval, so Newton from the
// left end increases monotonically to the root.  The derivative's two
// quotients share one division:
//   (h-hMin)/gMin + (h-hMax)/gMax = ((h-hMin) gMax + (h-hMax) gMin) / (gMin gMax) = num / den,
// and the step phi / dphi is taken with both sides scaled by den, so an
// iteration divides twice (1/h and the step).
// After convergence (|step| <= newtonTol h) the middle velocities use the
// last iteration's gk.
// ---------------------------------------------------------------------------
template <typename T>
int middle_state(T hL, T hR, T uL, T uR, T& hm, T& s1m, T& s2m, int* iters = nullptr) {
    typedef SolverConsts<T> C;
    const T hMin = std::min(hL, hR);
    const T hMax = std::max(hL, hR);
    const T du = uR - uL;
    const T cL = cell_celerity(hL);
    const T cR = cell_celerity(hR);
    int it = 0;
    int kind;
    if (hMin < C::dryTol) {
        hm = T(0);
        s1m = uR + uL - T(2) * cR + T(2) * cL;
        s2m = s1m;
        kind = 0;
    } else {
        const T cMin = cell_celerity(hMin);
        const T cMax = cell_celerity(hMax);
        const T invMin = T(1) / hMin;
        const T invMax = T(1) / hMax;
        const T halfG = T(0.5) * C::g;
        const T quarterG = T(0.25) * C::g;
        const T fMin = du + T(2) * (cMin - cMax);
        const T fMax = du + (hMax - hMin) * std::sqrt(halfG * (invMax + invMin));
        if (fMin > T(0)) {
            // two rarefactions: closed form
            const T a = std::max(T(0), -du + T(2) * (cL + cR));
            hm = a * a / (T(16) * C::g);
            const T cm = std::sqrt(C::g * hm);
            s1m = uL + T(2) * cL - T(3) * cm;
            s2m = uR - T(2) * cR + T(3) * cm;
            kind = 1;
        } else if (fMax <= T(0)) {
            // two shocks
            T h = hMax, gMin = 0, gMax = 0;
            for (it = 0; it < C::maxNewton; ++it) {
                const T invH = T(1) / h;
                gMin = std::sqrt(halfG * (invH + invMin));
                gMax = std::sqrt(halfG * (invH + invMax));
                const T phi = (du + (h - hMin) * gMin) + (h - hMax) * gMax;
                const T num = (h - hMin) * gMax + (h - hMax) * gMin;
                const T den = gMin * gMax;
                // phi / dphi with dphi = (gMin + gMax) - quarterG invH^2 num / den,
                // numerator and denominator scaled by den (one division)
                const T step = (phi * den) / ((gMin + gMax) * den - quarterG * invH * invH * num);
                h = h - step;
                if (std::fabs(step) <= C::newtonTol * h) {
                    ++it;
                    break;
                }
            }
            hm = h;
            const bool leftIsMin = !(hR < hL);
            const T gL = leftIsMin ? gMin : gMax;
            const T gR = leftIsMin ? gMax : gMin;
            const T u1m = uL - (hm - hL) * gL;
            const T u2m = uR + (hm - hR) * gR;
            const T cm = std::sqrt(C::g * hm);
            s1m = u1m - cm;
            s2m = u2m + cm;
            kind = 2;
        } else {
            // shock towards hMin, rarefaction towards hMax
            T h = hMin;
            for (it = 0; it < C::maxNewton; ++it) {
                const T invH = T(1) / h;
                const T gMin = std::sqrt(halfG * (invH + invMin));
                const T ch = std::sqrt(C::g * h);
                const T phi = (du + (h - hMin) * gMin) + T(2) * (ch - cMax);
                const T num = h - hMin;
                const T den = gMin;
                const T step = (phi * den) / ((gMin + ch * invH) * den - quarterG * invH * invH * num);
                h = h - step;
                if (std::fabs(step) <= C::newtonTol * h) {
                    ++it;
                    break;
                }
            }
            hm = h;
            const T cm = std::sqrt(C::g * hm);
            if (hL > hR) {
                s1m = uL + T(2) * cL - T(3) * cm;
                s2m = uL + T(2) * cL - cm;
            } else {
                s1m = uR - T(2) * cR + cm;
                s2m = uR - T(2) * cR + T(3) * cm;
            }
            kind = 3;
        }
    }
    if (iters) *iters = it;
    return kind;
}

// ---------------------------------------------------------------------------
// Augmented Riemann solver (George 2008): three f-waves (two acoustic waves
// at modified Einfeldt speeds plus a middle wave carrying the momentum-flux
// deficit), the steady-state wave subtracted for well-balancing, wet/dry
// walls and inundation.
// ---------------------------------------------------------------------------
// augrie_full is the decomposition alone; augrie (below) takes still edges
// out first.
template <typename T>
void augrie_full(T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR, T& huNetL, T& huNetR,
                 T& maxSpeed) {
    typedef SolverConsts<T> C;
    hNetL = hNetR = huNetL = huNetR = T(0);
    maxSpeed = T(0);
    if (hL < C::dryTol && hR < C::dryTol) return;

    T uL, uR;
    if (hL < C::dryTol) {
        hL = T(0);
        huL = T(0);
        uL = T(0);
    } else {
        uL = cell_velocity(huL, hL);
    }
    if (hR < C::dryTol) {
        hR = T(0);
        huR = T(0);
        uR = T(0);
    } else {
        uR = cell_velocity(huR, hR);
    }

    // wet/dry interface: wall (the dry cell's bathymetry is above anything the
    // wet state can reach) or inundation (limit the step to the water surface)
    bool wallL = false, wallR = false;
    if (hR < C::dryTol) {
        T hs, s1, s2;
        middle_state(hL, hL, uL, -uL, hs, s1, s2);
        const T hTest = std::max(hL, hs);
        if (hTest + bL < bR) {
            wallR = true;
            hR = hL;
            huR = -huL;
            uR = -uL;
            bR = bL;
        } else if (hL + bL < bR) {
            bR = hL + bL;
        }
    } else if (hL < C::dryTol) {
        T hs, s1, s2;
        middle_state(hR, hR, -uR, uR, hs, s1, s2);
        const T hTest = std::max(hR, hs);
        if (hTest + bR < bL) {
            wallL = true;
            hL = hR;
            huL = -huR;
            uL = -uR;
            bL = bR;
        } else if (hR + bR < bL) {
            bL = hR + bR;
        }
    }

    // Einfeldt speeds from the Roe average weighted by the celerities
    const T cL = cell_celerity(hL);
    const T cR = cell_celerity(hR);
    const T uHat = (cL * uL + cR * uR) / (cL + cR);
    const T cHat = std::sqrt(C::g * T(0.5) * (hR + hL));
    const T sE1 = std::min(uL - cL, uHat - cHat);
    const T sE2 = std::max(uR + cR, uHat + cHat);

    T hm, s1m, s2m;
    middle_state(hL, hR, uL, uR, hm, s1m, s2m);

    // modified Einfeldt speeds; middle wave halfway between
    const T l1 = std::min(sE1, s2m);
    const T l3 = std::max(sE2, s1m);
    const T l2 = T(0.5) * (l1 + l3);
    const T dl = l3 - l1;
    const T invDl = T(1) / dl;

    const T hStarHLL = std::max((huL - huR + l3 * hR - l1 * hL) * invDl, T(0));

    // steady-state wave (one pass)
    const T dh = hR - hL;
    const T dhu = huR - huL;
    const T phiL = huL * uL + T(0.5) * C::g * hL * hL;
    const T phiR = huR * uR + T(0.5) * C::g * hR * hR;
    const T dphi = phiR - phiL;
    const T db = bR - bL;

    const T hBar = std::max(T(0.5) * (hL + hR), T(0));
    const T uSum = uL + uR;
    const T s1s2bar = T(0.25) * uSum * uSum - C::g * hBar;
    const T s1s2tilde = std::max(T(0), uL * uR) - C::g * hBar;
    const T ct = C::criticalTol;

    bool sonic = false;
    if (std::fabs(s1s2bar) <= ct) sonic = true;
    if (s1s2bar * s1s2tilde <= ct) sonic = true;
    if (s1s2bar * l1 * l3 <= ct) sonic = true;
    if (std::min(std::fabs(l1), std::fabs(l3)) < ct) sonic = true;
    if (l1 < T(0) && s1m > T(0)) sonic = true;
    if (l3 > T(0) && s2m < T(0)) sonic = true;
    if ((uL + cL) * (uR + cR) < T(0)) sonic = true;
    if ((uL - cL) * (uR - cR) < T(0)) sonic = true;

    // the steady-state jumps share one reciprocal of s1s2bar; the depth
    // bounds hStarHLL dl / l3 and hStarHLL dl / l1 share one of l1 l3
    const T invBar = T(1) / s1s2bar;
    T ddh = sonic ? -db : db * C::g * hBar * invBar;
    const T hsd = hStarHLL * dl;
    if (l1 < -ct && l3 > ct) {
        const T inv13 = T(1) / (l1 * l3);
        ddh = std::min(ddh, hsd * l1 * inv13);
        ddh = std::max(ddh, hsd * l3 * inv13);
    } else if (l1 >= ct) {
        const T inv13 = T(1) / (l1 * l3);
        ddh = std::min(ddh, hsd * l3 * inv13);
        ddh = std::max(ddh, -hL);
    } else if (l3 <= -ct) {
        const T inv13 = T(1) / (l1 * l3);
        ddh = std::min(ddh, hR);
        ddh = std::max(ddh, hsd * l1 * inv13);
    }

    T ddphi = sonic ? -C::g * hBar * db : -db * C::g * hBar * s1s2tilde * invBar;
    const T pL = -hL * db;
    const T pR = -hR * db;
    ddphi = std::min(ddphi, C::g * std::max(pL, pR));
    ddphi = std::max(ddphi, C::g * std::min(pL, pR));

    const T del1 = dh - ddh;
    const T del2 = dhu;
    const T del3 = dphi - ddphi;

    // eigenvectors r1 = (1, l1, l1^2), r2 = (0, 0, 1), r3 = (1, l3, l3^2)
    const T l1sq = l1 * l1;
    const T l3sq = l3 * l3;
    const T beta1 = (l3 * del1 - del2) * invDl;
    const T beta3 = (del2 - l1 * del1) * invDl;
    const T beta2 = del3 - (beta1 * l1sq + beta3 * l3sq);

    T z1h = beta1 * l1, z1hu = beta1 * l1sq, s1 = l1;
    T z2h = T(0), z2hu = beta2, s2 = l2;
    T z3h = beta3 * l3, z3hu = beta3 * l3sq, s3 = l3;
    if (wallR) {
        z2h = z2hu = s2 = T(0);
        z3h = z3hu = s3 = T(0);
    }
    if (wallL) {
        z1h = z1hu = s1 = T(0);
        z2h = z2hu = s2 = T(0);
    }
    accumulate_wave(s1, z1h, z1hu, hNetL, hNetR, huNetL, huNetR);
    accumulate_wave(s2, z2h, z2hu, hNetL, hNetR, huNetL, huNetR);
    accumulate_wave(s3, z3h, z3hu, hNetL, hNetR, huNetL, huNetR);
    maxSpeed = std::max(std::fabs(s1), std::fabs(s3));
}

template <typename T>
void augrie(T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR, T& huNetL, T& huNetR,
            T& maxSpeed) {
    if (augrie_still(hL, hR, huL, huR, bL, bR)) {
        hNetL = hNetR = huNetL = huNetR = T(0);
        maxSpeed = augrie_still_speed(hL, hR);
        return;
    }
    augrie_full(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
}

// ---------------------------------------------------------------------------
// HLLE solver in f-wave form (Einfeldt 1988): two waves at the Einfeldt
// speeds, the HLL middle state splitting the flux jump, with the HLL
// diffusion in surface-gradient form (it acts on the jump of h + b, so a lake
// at rest stays at rest).  Same dry handling as the f-wave solver.
// ---------------------------------------------------------------------------
template <typename T>
void hlle(T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR, T& huNetL, T& huNetR,
          T& maxSpeed) {
    typedef SolverConsts<T> C;
    hNetL = hNetR = huNetL = huNetR = T(0);
    maxSpeed = T(0);
    const bool dryL = hL < C::dryTol;
    const bool dryR = hR < C::dryTol;
    if (dryL && dryR) return;
    if (dryL) {
        hL = hR;
        huL = -huR;
        bL = bR;
    } else if (dryR) {
        hR = hL;
        huR = -huL;
        bR = bL;
    }
    const T uL = cell_velocity(huL, hL);
    const T uR = cell_velocity(huR, hR);
    const T sqhL = std::sqrt(hL);
    const T sqhR = std::sqrt(hR);
    const T cL = cell_celerity(hL);
    const T cR = cell_celerity(hR);
    const T uRoe = (uL * sqhL + uR * sqhR) / (sqhL + sqhR);
    const T cRoe = std::sqrt(C::g * T(0.5) * (hL + hR));
    // HLLE speeds bounded by zero so that each side gets one wave
    const T sL = std::min(std::min(uL - cL, uRoe - cRoe), T(0));
    const T sR = std::max(std::max(uR + cR, uRoe + cRoe), T(0));
    const T df1 = huR - huL;
    const T fR = huR * uR + T(0.5) * C::g * hR * hR;
    const T fL = huL * uL + T(0.5) * C::g * hL * hL;
    const T df2 = (fR - fL) + T(0.5) * C::g * (hL + hR) * (bR - bL);
    // surface-gradient form (Zhou et al. 2001): the HLL diffusion acts on the
    // jump of the water surface h + b, so a lake at rest gives no update
    const T dq1 = (hR + bR) - (hL + bL);
    const T dq2 = huR - huL;
    const T invDs = T(1) / (sR - sL);
    // HLL flux F* minus the left flux, in f-wave form:
    //   A- dQ = sL (Q* - QL),  A+ dQ = sR (QR - Q*),  sum = dF
    // with Q* chosen so that the sum is the (source-balanced) flux jump.
    const T q1 = (sR * dq1 - df1) * invDs;  // Q* - QL (h)
    const T q2 = (sR * dq2 - df2) * invDs;  // Q* - QL (hu)
    hNetL = sL * q1;
    huNetL = sL * q2;
    hNetR = df1 - hNetL;
    huNetR = df2 - huNetL;
    if (dryL) {
        hNetL = T(0);
        huNetL = T(0);
    } else if (dryR) {
        hNetR = T(0);
        huNetR = T(0);
    }
    maxSpeed = std::max(std::fabs(sL), std::fabs(sR));
}

// Hybrid: f-wave for edges that are well wet on both sides, augmented
// solver otherwise (near-dry cells, wet/dry fronts).
template <typename T>
void hybrid(T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR, T& huNetL, T& huNetR,
            T& maxSpeed) {
    typedef SolverConsts<T> C;
    if (hL >= C::hybridTol && hR >= C::hybridTol)
        fwave(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
    else
        augrie(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
}

template <typename T>
inline void solve_edge(int solver, T hL, T hR, T huL, T huR, T bL, T bR, T& hNetL, T& hNetR,
                       T& huNetL, T& huNetR, T& maxSpeed) {
    switch (solver) {
        case SOLVER_HLLE:
            hlle(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
            break;
        case SOLVER_FWAVE:
            fwave(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
            break;
        case SOLVER_AUGRIE:
            augrie(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
            break;
        case SOLVER_AUGRIE_FULL:
            augrie_full(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
            break;
        default:
            hybrid(hL, hR, huL, huR, bL, bR, hNetL, hNetR, huNetL, huNetR, maxSpeed);
            break;
    }
}

}  // namespace oracle
