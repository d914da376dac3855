// solvers_dev.hpp -- per-edge Riemann solvers as CDNA4 device code.
//
// Interface of the reference's solver call sites
//   solver.computeNetUpdates(hL, hR, huL, huR, bL, bR,
//                            hNetL, hNetR, huNetL, huNetR, maxEdgeSpeed)
// (src/blocks/SWE_DimensionalSplittingMpi.cpp:403-410), whose bodies live in
// the un-vendored TUM-I5/swe_solvers submodule.  The arithmetic here is this
// library's own statement of the published methods:
//   f-wave   -- Bale, LeVeque, Mitran, Rossmanith 2002, Einfeldt speeds;
//   AugRie   -- George 2008 (J. Comput. Phys. 227), three f-waves with the
//               steady-state wave removed, wet/dry walls and inundation,
//               Newton middle state for shock cases;
//   HLLE     -- Einfeldt 1988 in fluctuation form;
//   Hybrid   -- f-wave when both depths >= kHybridTol, AugRie otherwise.
//
// Written for wave64 execution: the wet/dry and wave-direction decisions are
// predicates folded into selects; only the expensive, spatially clustered
// paths (wet/dry wall test, Newton loops) are real branches, which stay
// uniform inside most waves.  Every floating-point expression keeps one fixed
// evaluation order (no contraction in parity builds), so that results are
// reproducible against an IEEE CPU evaluation of the same formulas.
#pragma once

#include <hip/hip_runtime.h>

#include "swe_common.h"

namespace swe {
namespace dev {

// std::min / std::max semantics (first operand wins ties), not fminf/fmaxf.
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

struct EdgeOut {
    float hL, hR, huL, huR, speed;
};

// Add one f-wave to the side its speed points to; a wave of (near) zero speed
// is shared equally.  Expressed as selects of the three branch results.
__device__ __forceinline__ void acc_wave(float s, float zh, float zhu, EdgeOut& o) {
    const bool left = s < -kZeroTol;
    const bool right = s > kZeroTol;
    const float hzh = 0.5f * zh;
    const float hzhu = 0.5f * zhu;
    const float hL = left ? o.hL + zh : (right ? o.hL : o.hL + hzh);
    const float huL = left ? o.huL + zhu : (right ? o.huL : o.huL + hzhu);
    const float hR = right ? o.hR + zh : (left ? o.hR : o.hR + hzh);
    const float huR = right ? o.huR + zhu : (left ? o.huR : o.huR + hzhu);
    o.hL = hL;
    o.huL = huL;
    o.hR = hR;
    o.huR = huR;
}

// ---------------------------------------------------------------------------
// f-wave
// ---------------------------------------------------------------------------
__device__ __forceinline__ EdgeOut fwave(float hL, float hR, float huL, float huR, float bL, float bR) {
    const bool dryL = hL < kDryTol;
    const bool dryR = hR < kDryTol;
    const bool mirrorR = !dryL && dryR;
    // a dry neighbour is a reflecting wall: mirror the wet state
    const float hl = dryL ? hR : hL;
    const float hul = dryL ? -huR : huL;
    const float bl = dryL ? bR : bL;
    const float hr = mirrorR ? hL : hR;
    const float hur = mirrorR ? -huL : huR;
    const float br = mirrorR ? bL : bR;

    const float uL = hul / hl;
    const float uR = hur / hr;
    const float sqhL = sqrtf(hl);
    const float sqhR = sqrtf(hr);
    const float cL = sqrtf(kGravity * hl);
    const float cR = sqrtf(kGravity * hr);
    const float hRoe = 0.5f * (hl + hr);
    const float uRoe = (uL * sqhL + uR * sqhR) / (sqhL + sqhR);
    const float cRoe = sqrtf(kGravity * hRoe);
    const float s1 = smin(uL - cL, uRoe - cRoe);
    const float s2 = smax(uR + cR, uRoe + cRoe);

    const float df1 = hur - hul;
    const float fR = hur * uR + 0.5f * kGravity * hr * hr;
    const float fL = hul * uL + 0.5f * kGravity * hl * hl;
    const float df2 = (fR - fL) + 0.5f * kGravity * (hl + hr) * (br - bl);
    const float ds = s2 - s1;
    const float beta1 = (s2 * df1 - df2) / ds;
    const float beta2 = (df2 - s1 * df1) / ds;

    EdgeOut o{0.f, 0.f, 0.f, 0.f, 0.f};
    acc_wave(s1, beta1, beta1 * s1, o);
    acc_wave(s2, beta2, beta2 * s2, o);
    // the dry side receives nothing
    o.hL = dryL ? 0.f : o.hL;
    o.huL = dryL ? 0.f : o.huL;
    o.hR = mirrorR ? 0.f : o.hR;
    o.huR = mirrorR ? 0.f : o.huR;
    o.speed = smax(fabsf(s1), fabsf(s2));
    if (dryL && dryR) o = EdgeOut{0.f, 0.f, 0.f, 0.f, 0.f};
    return o;
}

// ---------------------------------------------------------------------------
// Middle state of the Riemann problem: closed form for two rarefactions,
// Newton (monotone from the left: the wave curves are increasing and concave)
// for two shocks and for shock + rarefaction.
// ---------------------------------------------------------------------------
struct Middle {
    float hm, s1m, s2m;
};

__device__ __forceinline__ Middle middle_state(float hL, float hR, float uL, float uR) {
    const float hMin = smin(hL, hR);
    const float hMax = smax(hL, hR);
    const float du = uR - uL;
    const float cL = sqrtf(kGravity * hL);
    const float cR = sqrtf(kGravity * hR);
    Middle m;
    if (hMin < kDryTol) {
        m.hm = 0.f;
        m.s1m = uR + uL - 2.f * cR + 2.f * cL;
        m.s2m = m.s1m;
        return m;
    }
    const float cMin = sqrtf(kGravity * hMin);
    const float cMax = sqrtf(kGravity * hMax);
    const float fMin = du + 2.f * (cMin - cMax);
    const float fMax = du + (hMax - hMin) * sqrtf(0.5f * kGravity * (hMax + hMin) / (hMax * hMin));
    if (fMin > 0.f) {
        const float a = smax(0.f, -du + 2.f * (cL + cR));
        m.hm = a * a / (16.f * kGravity);
        const float cm = sqrtf(kGravity * m.hm);
        m.s1m = uL + 2.f * cL - 3.f * cm;
        m.s2m = uR - 2.f * cR + 3.f * cm;
        return m;
    }
    const bool shocks = fMax <= 0.f;
    float h = shocks ? hMax : hMin;
    const float invL = 1.f / hL;
    const float invR = 1.f / hR;
    const float invMin = 1.f / hMin;
    for (int it = 0; it < kMaxNewton; ++it) {
        const float invH = 1.f / h;
        float step;
        if (shocks) {
            const float gL = sqrtf(0.5f * kGravity * (invH + invL));
            const float gR = sqrtf(0.5f * kGravity * (invH + invR));
            const float phi = du + (h - hL) * gL + (h - hR) * gR;
            const float dphi =
                gL - kGravity * (h - hL) / (4.f * h * h * gL) + gR - kGravity * (h - hR) / (4.f * h * h * gR);
            step = phi / dphi;
        } else {
            const float gm = sqrtf(0.5f * kGravity * (invH + invMin));
            const float ch = sqrtf(kGravity * h);
            const float phi = du + 2.f * (ch - cMax) + (h - hMin) * gm;
            const float dphi = ch / h + gm - kGravity * (h - hMin) / (4.f * h * h * gm);
            step = phi / dphi;
        }
        h = h - step;
        if (fabsf(step) <= kNewtonTol * h) break;
    }
    m.hm = h;
    const float cm = sqrtf(kGravity * h);
    if (shocks) {
        const float gmL = sqrtf(0.5f * kGravity * (1.f / h + invL));
        const float gmR = sqrtf(0.5f * kGravity * (1.f / h + invR));
        const float u1m = uL - (h - hL) * gmL;
        const float u2m = uR + (h - hR) * gmR;
        m.s1m = u1m - cm;
        m.s2m = u2m + cm;
    } else if (hL > hR) {
        m.s1m = uL + 2.f * cL - 3.f * cm;
        m.s2m = uL + 2.f * cL - cm;
    } else {
        m.s1m = uR - 2.f * cR + cm;
        m.s2m = uR - 2.f * cR + 3.f * cm;
    }
    return m;
}

// ---------------------------------------------------------------------------
// Augmented Riemann solver.  MiddleFn computes (hm, s1m, s2m) for the main
// Riemann problem of a wet-wet or inundation edge; the exact Newton
// middle_state is the default, the learned surrogate plugs in here.
// ---------------------------------------------------------------------------
struct ExactMiddle {
    __device__ __forceinline__ Middle operator()(float hL, float hR, float uL, float uR) const {
        return middle_state(hL, hR, uL, uR);
    }
};

template <typename MiddleFn>
__device__ __forceinline__ EdgeOut augrie_with(float hL, float hR, float huL, float huR, float bL, float bR,
                                               const MiddleFn& middle) {
    const bool dryBoth = hL < kDryTol && hR < kDryTol;
    const bool dryLin = hL < kDryTol;
    const bool dryRin = hR < kDryTol;
    hL = dryLin ? 0.f : hL;
    huL = dryLin ? 0.f : huL;
    hR = dryRin ? 0.f : hR;
    huR = dryRin ? 0.f : huR;
    float uL = dryLin ? 0.f : huL / hL;
    float uR = dryRin ? 0.f : huR / hR;

    // wet/dry interface: wall or inundation
    bool wallL = false, wallR = false;
    if ((dryLin || dryRin) && !dryBoth) {
        if (dryRin) {
            const Middle t = middle_state(hL, hL, uL, -uL);
            const float hTest = smax(hL, t.hm);
            if (hTest + bL < bR) {
                wallR = true;
                hR = hL;
                huR = -huL;
                uR = -uL;
                bR = bL;
            } else if (hL + bL < bR) {
                bR = hL + bL;
            }
        } else {
            const Middle t = middle_state(hR, hR, -uR, uR);
            const float hTest = smax(hR, t.hm);
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
    }

    const float cL = sqrtf(kGravity * hL);
    const float cR = sqrtf(kGravity * hR);
    const float uHat = (cL * uL + cR * uR) / (cL + cR);
    const float cHat = sqrtf(kGravity * 0.5f * (hR + hL));
    const float sE1 = smin(uL - cL, uHat - cHat);
    const float sE2 = smax(uR + cR, uHat + cHat);

    const Middle mid = middle(hL, hR, uL, uR);

    const float l1 = smin(sE1, mid.s2m);
    const float l3 = smax(sE2, mid.s1m);
    const float l2 = 0.5f * (l1 + l3);
    const float dl = l3 - l1;
    const float hStarHLL = smax((huL - huR + l3 * hR - l1 * hL) / dl, 0.f);

    const float dh = hR - hL;
    const float dhu = huR - huL;
    const float phiL = huL * uL + 0.5f * kGravity * hL * hL;
    const float phiR = huR * uR + 0.5f * kGravity * hR * hR;
    const float dphi = phiR - phiL;
    const float db = bR - bL;

    const float hBar = smax(0.5f * (hL + hR), 0.f);
    const float uSum = uL + uR;
    const float s1s2bar = 0.25f * uSum * uSum - kGravity * hBar;
    const float s1s2tilde = smax(0.f, uL * uR) - kGravity * hBar;
    const float ct = kCriticalTol;

    const bool sonic = (fabsf(s1s2bar) <= ct) || (s1s2bar * s1s2tilde <= ct) || (s1s2bar * l1 * l3 <= ct) ||
                       (smin(fabsf(l1), fabsf(l3)) < ct) || ((l1 < 0.f) && (mid.s1m > 0.f)) ||
                       ((l3 > 0.f) && (mid.s2m < 0.f)) || ((uL + cL) * (uR + cR) < 0.f) ||
                       ((uL - cL) * (uR - cR) < 0.f);

    float ddh = sonic ? -db : db * kGravity * hBar / s1s2bar;
    const bool caseA = (l1 < -ct) && (l3 > ct);
    const bool caseB = !caseA && (l1 >= ct);
    const bool caseC = !caseA && !caseB && (l3 <= -ct);
    if (caseA || caseB || caseC) {
        const float dA = hStarHLL * dl / l3;
        const float dB = hStarHLL * dl / l1;
        const float hiBound = caseA ? dA : (caseB ? dB : hR);
        const float loBound = caseA ? dB : (caseB ? -hL : dA);
        ddh = smin(ddh, hiBound);
        ddh = smax(ddh, loBound);
    }

    float ddphi = sonic ? -kGravity * hBar * db : -db * kGravity * hBar * s1s2tilde / s1s2bar;
    const float pL = -hL * db;
    const float pR = -hR * db;
    ddphi = smin(ddphi, kGravity * smax(pL, pR));
    ddphi = smax(ddphi, kGravity * smin(pL, pR));

    const float del1 = dh - ddh;
    const float del2 = dhu;
    const float del3 = dphi - ddphi;

    const float l1sq = l1 * l1;
    const float l3sq = l3 * l3;
    const float beta1 = (l3 * del1 - del2) / dl;
    const float beta3 = (del2 - l1 * del1) / dl;
    const float beta2 = del3 - (beta1 * l1sq + beta3 * l3sq);

    const bool kill1 = wallL;
    const bool kill2 = wallL || wallR;
    const bool kill3 = wallR;
    const float z1h = kill1 ? 0.f : beta1 * l1;
    const float z1hu = kill1 ? 0.f : beta1 * l1sq;
    const float s1 = kill1 ? 0.f : l1;
    const float z2hu = kill2 ? 0.f : beta2;
    const float s2 = kill2 ? 0.f : l2;
    const float z3h = kill3 ? 0.f : beta3 * l3;
    const float z3hu = kill3 ? 0.f : beta3 * l3sq;
    const float s3 = kill3 ? 0.f : l3;

    EdgeOut o{0.f, 0.f, 0.f, 0.f, 0.f};
    acc_wave(s1, z1h, z1hu, o);
    acc_wave(s2, 0.f, z2hu, o);
    acc_wave(s3, z3h, z3hu, o);
    o.speed = smax(fabsf(s1), fabsf(s3));
    if (dryBoth) o = EdgeOut{0.f, 0.f, 0.f, 0.f, 0.f};
    return o;
}

__device__ __forceinline__ EdgeOut augrie(float hL, float hR, float huL, float huR, float bL, float bR) {
    return augrie_with(hL, hR, huL, huR, bL, bR, ExactMiddle());
}

// ---------------------------------------------------------------------------
// HLLE in fluctuation form
// ---------------------------------------------------------------------------
__device__ __forceinline__ EdgeOut hlle(float hL, float hR, float huL, float huR, float bL, float bR) {
    const bool dryL = hL < kDryTol;
    const bool dryR = hR < kDryTol;
    const bool mirrorR = !dryL && dryR;
    const float hl = dryL ? hR : hL;
    const float hul = dryL ? -huR : huL;
    const float bl = dryL ? bR : bL;
    const float hr = mirrorR ? hL : hR;
    const float hur = mirrorR ? -huL : huR;
    const float br = mirrorR ? bL : bR;
    const float uL = hul / hl;
    const float uR = hur / hr;
    const float cL = sqrtf(kGravity * hl);
    const float cR = sqrtf(kGravity * hr);
    const float sqhL = sqrtf(hl);
    const float sqhR = sqrtf(hr);
    const float uRoe = (uL * sqhL + uR * sqhR) / (sqhL + sqhR);
    const float cRoe = sqrtf(kGravity * 0.5f * (hl + hr));
    const float sL = smin(smin(uL - cL, uRoe - cRoe), 0.f);
    const float sR = smax(smax(uR + cR, uRoe + cRoe), 0.f);
    const float df1 = hur - hul;
    const float fR = hur * uR + 0.5f * kGravity * hr * hr;
    const float fL = hul * uL + 0.5f * kGravity * hl * hl;
    const float df2 = (fR - fL) + 0.5f * kGravity * (hl + hr) * (br - bl);
    const float dq1 = (hr + br) - (hl + bl);  // surface-gradient form: lake at rest stays at rest
    const float dq2 = hur - hul;
    const float ds = sR - sL;
    const float q1 = (sR * dq1 - df1) / ds;
    const float q2 = (sR * dq2 - df2) / ds;
    EdgeOut o;
    o.hL = sL * q1;
    o.huL = sL * q2;
    o.hR = df1 - o.hL;
    o.huR = df2 - o.huL;
    o.hL = dryL ? 0.f : o.hL;
    o.huL = dryL ? 0.f : o.huL;
    o.hR = mirrorR ? 0.f : o.hR;
    o.huR = mirrorR ? 0.f : o.huR;
    o.speed = smax(fabsf(sL), fabsf(sR));
    if (dryL && dryR) o = EdgeOut{0.f, 0.f, 0.f, 0.f, 0.f};
    return o;
}

template <int S>
struct EdgeSolver;

template <>
struct EdgeSolver<SOLVER_FWAVE> {
    __device__ __forceinline__ static EdgeOut solve(float hL, float hR, float huL, float huR, float bL, float bR) {
        return fwave(hL, hR, huL, huR, bL, bR);
    }
};
template <>
struct EdgeSolver<SOLVER_AUGRIE> {
    __device__ __forceinline__ static EdgeOut solve(float hL, float hR, float huL, float huR, float bL, float bR) {
        return augrie(hL, hR, huL, huR, bL, bR);
    }
};
template <>
struct EdgeSolver<SOLVER_HLLE> {
    __device__ __forceinline__ static EdgeOut solve(float hL, float hR, float huL, float huR, float bL, float bR) {
        return hlle(hL, hR, huL, huR, bL, bR);
    }
};
template <>
struct EdgeSolver<SOLVER_HYBRID> {
    __device__ __forceinline__ static EdgeOut solve(float hL, float hR, float huL, float huR, float bL, float bR) {
        if (hL >= kHybridTol && hR >= kHybridTol) return fwave(hL, hR, huL, huR, bL, bR);
        return augrie(hL, hR, huL, huR, bL, bR);
    }
};

}  // namespace dev
}  // namespace swe
