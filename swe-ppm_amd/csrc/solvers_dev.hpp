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

// One side of an edge: depth, normal momentum, bathymetry and the per-cell
// quantities every edge touching the cell needs.  They are computed once per
// cell (cell_side) and reused by its four edges; the values are the same
// IEEE results each edge would compute itself.
struct Side {
    float h, hu, b;
    float u;    // hu (1/h) (normal velocity; meaningless for dry cells)
    float c;    // sqrt(g h)
    float sq;   // sqrt(h)
    float inv;  // 1 / h (the augmented solver's middle state divides by both depths)
};

__device__ __forceinline__ Side cell_side(float h, float hu, float b) {
    Side s;
    s.h = h;
    s.hu = hu;
    s.b = b;
    s.inv = 1.f / h;
    s.u = hu * s.inv;
    s.sq = sqrtf(h);
    s.c = kSqrtGravity * s.sq;
    return s;
}

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

// A dry neighbour of an f-wave / HLLE edge is a reflecting wall: the wet
// state mirrored (h, -hu, b; the velocity negates exactly).
struct Mirrored {
    Side L, R;
    bool dryL, dryR, mirrorR;
};
__device__ __forceinline__ Mirrored mirror_dry(const Side& L, const Side& R) {
    Mirrored m;
    m.dryL = L.h < kDryTol;
    m.dryR = R.h < kDryTol;
    m.mirrorR = !m.dryL && m.dryR;
    m.L.h = m.dryL ? R.h : L.h;
    m.L.hu = m.dryL ? -R.hu : L.hu;
    m.L.b = m.dryL ? R.b : L.b;
    m.L.u = m.dryL ? -R.u : L.u;
    m.L.c = m.dryL ? R.c : L.c;
    m.L.sq = m.dryL ? R.sq : L.sq;
    m.L.inv = m.dryL ? R.inv : L.inv;
    m.R.h = m.mirrorR ? L.h : R.h;
    m.R.hu = m.mirrorR ? -L.hu : R.hu;
    m.R.b = m.mirrorR ? L.b : R.b;
    m.R.u = m.mirrorR ? -L.u : R.u;
    m.R.c = m.mirrorR ? L.c : R.c;
    m.R.sq = m.mirrorR ? L.sq : R.sq;
    m.R.inv = m.mirrorR ? L.inv : R.inv;
    return m;
}

// ---------------------------------------------------------------------------
// f-wave
// ---------------------------------------------------------------------------
__device__ __forceinline__ EdgeOut fwave(const Side& Lin, const Side& Rin) {
    const Mirrored m = mirror_dry(Lin, Rin);
    const Side& L = m.L;
    const Side& R = m.R;
    const float hRoe = 0.5f * (L.h + R.h);
    const float uRoe = (L.u * L.sq + R.u * R.sq) / (L.sq + R.sq);
    const float cRoe = sqrtf(kGravity * hRoe);
    const float s1 = smin(L.u - L.c, uRoe - cRoe);
    const float s2 = smax(R.u + R.c, uRoe + cRoe);

    const float df1 = R.hu - L.hu;
    const float fR = R.hu * R.u + 0.5f * kGravity * R.h * R.h;
    const float fL = L.hu * L.u + 0.5f * kGravity * L.h * L.h;
    const float df2 = (fR - fL) + 0.5f * kGravity * (L.h + R.h) * (R.b - L.b);
    const float invDs = 1.f / (s2 - s1);
    const float beta1 = (s2 * df1 - df2) * invDs;
    const float beta2 = (df2 - s1 * df1) * invDs;

    EdgeOut o{0.f, 0.f, 0.f, 0.f, 0.f};
    acc_wave(s1, beta1, beta1 * s1, o);
    acc_wave(s2, beta2, beta2 * s2, o);
    o.hL = m.dryL ? 0.f : o.hL;
    o.huL = m.dryL ? 0.f : o.huL;
    o.hR = m.mirrorR ? 0.f : o.hR;
    o.huR = m.mirrorR ? 0.f : o.huR;
    o.speed = smax(fabsf(s1), fabsf(s2));
    if (m.dryL && m.dryR) o = EdgeOut{0.f, 0.f, 0.f, 0.f, 0.f};
    return o;
}

// ---------------------------------------------------------------------------
// Middle state of the Riemann problem: phi(h) = du + f(h; hMin) + f(h; hMax)
// with shock curves (h-k) sqrt(g/2 (1/h + 1/k)) and rarefaction curves
// 2 (sqrt(g h) - sqrt(g k)).  Two rarefactions: closed form.  Otherwise one
// branch-free Newton loop serves both the two-shock case (from hMax) and the
// shock+rarefaction case (from hMin): the curve towards hMax is selected per
// lane, and its two square roots share one instruction slot.  The
// derivative's two quotients share one division, num / den, and the step
// phi / dphi is taken with both sides scaled by den, so an iteration divides
// twice (1/h and the step).  phi is increasing and
// concave on the search interval, so the iterates increase monotonically.
// cL, cR are the cells' celerities sqrt(g) sqrt(h); invL, invR are 1 / hL,
// 1 / hR, computed once per cell.
// ReuseFirst: the first iteration starts at hMax or hMin, where 1/h is known
// and, for two shocks, one of the two curve square roots (the bracket's own
// quantities); with ReuseFirst it takes them instead of recomputing them, so
// it costs no division and one square root (two shocks) or two (shock and
// rarefaction).  The values are the ones the loop computes, so both forms give
// identical results.  The compiler peels the first iteration
// for it, a second copy of the loop body; measured under the 72-register
// budget of seven waves per SIMD, that pays (the single-block sweeps and the
// batched kernels use it).
// ---------------------------------------------------------------------------
struct Middle {
    float hm, s1m, s2m;
};

template <bool ReuseFirst = false>
__device__ __forceinline__ Middle middle_state(float hL, float hR, float uL, float uR, float cL, float cR, float invL,
                                               float invR) {
    const float hMin = smin(hL, hR);
    const float hMax = smax(hL, hR);
    const float du = uR - uL;
    Middle m;
    if (hMin < kDryTol) {
        m.hm = 0.f;
        m.s1m = uR + uL - 2.f * cR + 2.f * cL;
        m.s2m = m.s1m;
        return m;
    }
    const float cMin = (hR < hL) ? cR : cL;
    const float cMax = (hL < hR) ? cR : cL;
    const float invMin = (hR < hL) ? invR : invL;  // 1 / hMin
    const float invMax = (hL < hR) ? invR : invL;  // 1 / hMax
    const float halfG = 0.5f * kGravity;
    const float quarterG = 0.25f * kGravity;
    const float fMin = du + 2.f * (cMin - cMax);
    const float gMaxMin = sqrtf(halfG * (invMax + invMin));
    const float fMax = du + (hMax - hMin) * gMaxMin;
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
    float invH = 0.f, gMin = 0.f, s2 = 0.f;
    if constexpr (ReuseFirst) {
        // at h = hMax the shock curve of hMin is gMaxMin and the new root is
        // gMax(hMax); at h = hMin the new root is gMin(hMin) and the
        // rarefaction root sqrt(g hMin) is taken as the loop takes it (the
        // cell's celerity cMin = sqrt(g) sqrt(hMin) can differ in the last bit)
        invH = shocks ? invMax : invMin;
        const float r0 = sqrtf(halfG * (invH + invH));
        gMin = shocks ? gMaxMin : r0;
        s2 = shocks ? r0 : sqrtf(kGravity * hMin);
    }
    for (int it = 0; it < kMaxNewton; ++it) {
        if (!ReuseFirst || it > 0) {
            invH = 1.f / h;
            gMin = sqrtf(halfG * (invH + invMin));
            // shocks: gMax(h); shock+rarefaction: sqrt(g h)
            s2 = sqrtf(shocks ? halfG * (invH + invMax) : kGravity * h);
        }
        const float t2 = shocks ? (h - hMax) * s2 : 2.f * (s2 - cMax);
        const float phi = (du + (h - hMin) * gMin) + t2;
        const float q = shocks ? s2 : 1.f;
        const float num = (h - hMin) * q + (shocks ? (h - hMax) * gMin : 0.f);
        const float den = gMin * q;
        const float first = gMin + (shocks ? s2 : s2 * invH);
        // phi / (first - quarterG invH^2 num / den) with both terms scaled by
        // den: one division instead of two
        const float step = (phi * den) / (first * den - quarterG * invH * invH * num);
        h = h - step;
        if (fabsf(step) <= kNewtonTol * h) break;
    }
    m.hm = h;
    const float cm = sqrtf(kGravity * h);
    if (shocks) {
        const bool leftIsMin = !(hR < hL);
        const float gL = leftIsMin ? gMin : s2;
        const float gR = leftIsMin ? s2 : gMin;
        const float u1m = uL - (h - hL) * gL;
        const float u2m = uR + (h - hR) * gR;
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
template <bool ReuseFirst>
struct ExactMiddle {
    static constexpr bool kDefers = false;
    __device__ __forceinline__ Middle operator()(float hL, float hR, float uL, float uR, float cL, float cR,
                                                 float invL, float invR, float, float) const {
        return middle_state<ReuseFirst>(hL, hR, uL, uR, cL, cR, invL, invR);
    }
};

// Wet/dry handling of the augmented solver (George 2008): a dry side next to
// a wet one is either a wall (the wet side's middle state cannot reach its
// bed: reflect) or inundated (the bed step is cut to the water surface).
struct AugPrep {
    float hL, huL, uL, cL, bL, invL;
    float hR, huR, uR, cR, bR, invR;
    bool wallL, wallR, dryBoth;
};

__device__ __forceinline__ AugPrep augrie_prepare(const Side& Lin, const Side& Rin) {
    const bool dryLin = Lin.h < kDryTol;
    const bool dryRin = Rin.h < kDryTol;
    AugPrep p;
    p.dryBoth = dryLin && dryRin;
    p.hL = dryLin ? 0.f : Lin.h;
    p.huL = dryLin ? 0.f : Lin.hu;
    p.uL = dryLin ? 0.f : Lin.u;
    p.cL = dryLin ? 0.f : Lin.c;
    p.bL = Lin.b;
    p.invL = Lin.inv;  // a dry side never reaches the divisions
    p.hR = dryRin ? 0.f : Rin.h;
    p.huR = dryRin ? 0.f : Rin.hu;
    p.uR = dryRin ? 0.f : Rin.u;
    p.cR = dryRin ? 0.f : Rin.c;
    p.bR = Rin.b;
    p.invR = Rin.inv;
    p.wallL = false;
    p.wallR = false;
    if ((dryLin || dryRin) && !p.dryBoth) {
        if (dryRin) {
            const Middle t = middle_state(p.hL, p.hL, p.uL, -p.uL, p.cL, p.cL, p.invL, p.invL);
            const float hTest = smax(p.hL, t.hm);
            if (hTest + p.bL < p.bR) {
                p.wallR = true;
                p.hR = p.hL;
                p.huR = -p.huL;
                p.uR = -p.uL;
                p.cR = p.cL;
                p.invR = p.invL;
                p.bR = p.bL;
            } else if (p.hL + p.bL < p.bR) {
                p.bR = p.hL + p.bL;
            }
        } else {
            const Middle t = middle_state(p.hR, p.hR, -p.uR, p.uR, p.cR, p.cR, p.invR, p.invR);
            const float hTest = smax(p.hR, t.hm);
            if (hTest + p.bR < p.bL) {
                p.wallL = true;
                p.hL = p.hR;
                p.huL = -p.huR;
                p.uL = -p.uR;
                p.cL = p.cR;
                p.invL = p.invR;
                p.bL = p.bR;
            } else if (p.hR + p.bR < p.bL) {
                p.bL = p.hR + p.bR;
            }
        }
    }
    return p;
}

// A still edge: both cells hold at least kStillTol of water, zero normal
// momentum on both sides, the same free surface h + b in single precision,
// and depths within a factor kStillRatio.  It is a lake at rest, whose exact
// Riemann solution carries no flux: the solver returns zero net updates and
// the speed the full decomposition reports, max(|sE1|, |sE2|) with
// sE1 = min(-cL, -cHat), sE2 = max(cR, cHat) (at rest u = uHat = 0, and for
// depths within a factor 2 the middle-state speeds never replace the
// Einfeldt speeds: s2m > 0 > sE1 and s1m < 0 < sE2).  Still water thus stays
// exactly still, and a wave whose edges are all still skips the middle state
// and the decomposition.  In a wave that mixes still and moving edges every
// lane runs the decomposition and the still lanes take the still result at
// the end: the learned middle state runs its network for the whole wave
// and must be reached with every lane active.
__device__ __forceinline__ bool still_edge(const Side& L, const Side& R) {
    return L.h >= kStillTol && R.h >= kStillTol && L.hu == 0.f && R.hu == 0.f && L.h + L.b == R.h + R.b &&
           L.h <= kStillRatio * R.h && R.h <= kStillRatio * L.h;
}

__device__ __forceinline__ float still_speed(const Side& L, const Side& R) {
    const float cHat = sqrtf(kGravity * 0.5f * (R.h + L.h));
    const float sE1 = smin(-L.c, -cHat);
    const float sE2 = smax(R.c, cHat);
    return smax(fabsf(sE1), fabsf(sE2));
}

template <typename MiddleFn>
__device__ __forceinline__ EdgeOut augrie_with(const Side& Lin, const Side& Rin, const MiddleFn& middle) {
    const bool still = still_edge(Lin, Rin);
    if (__all(still)) return EdgeOut{0.f, 0.f, 0.f, 0.f, still_speed(Lin, Rin)};
    const AugPrep p = augrie_prepare(Lin, Rin);
    const bool dryBoth = p.dryBoth;
    const bool wallL = p.wallL, wallR = p.wallR;
    const float hL = p.hL, huL = p.huL, uL = p.uL, cL = p.cL, bL = p.bL;
    const float hR = p.hR, huR = p.huR, uR = p.uR, cR = p.cR, bR = p.bR;

    const float uHat = (cL * uL + cR * uR) / (cL + cR);
    const float cHat = sqrtf(kGravity * 0.5f * (hR + hL));
    const float sE1 = smin(uL - cL, uHat - cHat);
    const float sE2 = smax(uR + cR, uHat + cHat);

    const Middle mid = middle(hL, hR, uL, uR, cL, cR, p.invL, p.invR, bL, bR);

    const float l1 = smin(sE1, mid.s2m);
    const float l3 = smax(sE2, mid.s1m);
    const float l2 = 0.5f * (l1 + l3);
    const float dl = l3 - l1;
    const float invDl = 1.f / dl;
    const float hStarHLL = smax((huL - huR + l3 * hR - l1 * hL) * invDl, 0.f);

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

    const float invBar = 1.f / s1s2bar;
    float ddh = sonic ? -db : db * kGravity * hBar * invBar;
    const bool caseA = (l1 < -ct) && (l3 > ct);
    const bool caseB = !caseA && (l1 >= ct);
    const bool caseC = !caseA && !caseB && (l3 <= -ct);
    if (caseA || caseB || caseC) {
        const float hsd = hStarHLL * dl;
        const float inv13 = 1.f / (l1 * l3);
        const float dA = hsd * l1 * inv13;  // hStarHLL dl / l3
        const float dB = hsd * l3 * inv13;  // hStarHLL dl / l1
        const float hiBound = caseA ? dA : (caseB ? dB : hR);
        const float loBound = caseA ? dB : (caseB ? -hL : dA);
        ddh = smin(ddh, hiBound);
        ddh = smax(ddh, loBound);
    }

    float ddphi = sonic ? -kGravity * hBar * db : -db * kGravity * hBar * s1s2tilde * invBar;
    const float pL = -hL * db;
    const float pR = -hR * db;
    ddphi = smin(ddphi, kGravity * smax(pL, pR));
    ddphi = smax(ddphi, kGravity * smin(pL, pR));

    const float del1 = dh - ddh;
    const float del2 = dhu;
    const float del3 = dphi - ddphi;

    const float l1sq = l1 * l1;
    const float l3sq = l3 * l3;
    const float beta1 = (l3 * del1 - del2) * invDl;
    const float beta3 = (del2 - l1 * del1) * invDl;
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
    // a middle function that defers the edge returns a negative depth: the
    // edge is marked by a negative speed and solved again elsewhere
    if constexpr (MiddleFn::kDefers)
        if (mid.hm < 0.f) o.speed = -1.f;
    if (still) o = EdgeOut{0.f, 0.f, 0.f, 0.f, still_speed(Lin, Rin)};
    return o;
}

template <bool ReuseFirst = false>
__device__ __forceinline__ EdgeOut augrie(const Side& L, const Side& R) {
    return augrie_with(L, R, ExactMiddle<ReuseFirst>());
}

// ---------------------------------------------------------------------------
// HLLE in fluctuation form, surface-gradient diffusion (well-balanced)
// ---------------------------------------------------------------------------
__device__ __forceinline__ EdgeOut hlle(const Side& Lin, const Side& Rin) {
    const Mirrored m = mirror_dry(Lin, Rin);
    const Side& L = m.L;
    const Side& R = m.R;
    const float uRoe = (L.u * L.sq + R.u * R.sq) / (L.sq + R.sq);
    const float cRoe = sqrtf(kGravity * 0.5f * (L.h + R.h));
    const float sL = smin(smin(L.u - L.c, uRoe - cRoe), 0.f);
    const float sR = smax(smax(R.u + R.c, uRoe + cRoe), 0.f);
    const float df1 = R.hu - L.hu;
    const float fR = R.hu * R.u + 0.5f * kGravity * R.h * R.h;
    const float fL = L.hu * L.u + 0.5f * kGravity * L.h * L.h;
    const float df2 = (fR - fL) + 0.5f * kGravity * (L.h + R.h) * (R.b - L.b);
    const float dq1 = (R.h + R.b) - (L.h + L.b);  // surface-gradient form: lake at rest stays at rest
    const float dq2 = R.hu - L.hu;
    const float invDs = 1.f / (sR - sL);
    const float q1 = (sR * dq1 - df1) * invDs;
    const float q2 = (sR * dq2 - df2) * invDs;
    EdgeOut o;
    o.hL = sL * q1;
    o.huL = sL * q2;
    o.hR = df1 - o.hL;
    o.huR = df2 - o.huL;
    o.hL = m.dryL ? 0.f : o.hL;
    o.huL = m.dryL ? 0.f : o.huL;
    o.hR = m.mirrorR ? 0.f : o.hR;
    o.huR = m.mirrorR ? 0.f : o.huR;
    o.speed = smax(fabsf(sL), fabsf(sR));
    if (m.dryL && m.dryR) o = EdgeOut{0.f, 0.f, 0.f, 0.f, 0.f};
    return o;
}

template <int S>
struct EdgeSolver;

// ReuseFirst selects the form of the Newton loop (middle_state); the results
// are the same.
template <>
struct EdgeSolver<SOLVER_FWAVE> {
    template <bool ReuseFirst = false>
    __device__ __forceinline__ static EdgeOut solve(const Side& L, const Side& R) {
        return fwave(L, R);
    }
};
template <>
struct EdgeSolver<SOLVER_AUGRIE> {
    template <bool ReuseFirst = false>
    __device__ __forceinline__ static EdgeOut solve(const Side& L, const Side& R) {
        return augrie<ReuseFirst>(L, R);
    }
};
template <>
struct EdgeSolver<SOLVER_HLLE> {
    template <bool ReuseFirst = false>
    __device__ __forceinline__ static EdgeOut solve(const Side& L, const Side& R) {
        return hlle(L, R);
    }
};
template <>
struct EdgeSolver<SOLVER_HYBRID> {
    template <bool ReuseFirst = false>
    __device__ __forceinline__ static EdgeOut solve(const Side& L, const Side& R) {
        if (L.h >= kHybridTol && R.h >= kHybridTol) return fwave(L, R);
        return augrie<ReuseFirst>(L, R);
    }
};

// Solver with per-kernel state, built once at kernel start from an optional
// parameter block (the learned middle state keeps its network weights in
// registers; the exact solvers have no state).
template <int S, bool ReuseFirst = false>
struct SolverCtx {
    __device__ __forceinline__ explicit SolverCtx(const unsigned*) {}
    __device__ __forceinline__ EdgeOut solve(const Side& L, const Side& R) const {
        return EdgeSolver<S>::template solve<ReuseFirst>(L, R);
    }
};

// Convenience for raw edge states (one-off evaluations).
template <int S>
__device__ __forceinline__ EdgeOut solve_raw(float hL, float hR, float huL, float huR, float bL, float bR) {
    return EdgeSolver<S>::solve(cell_side(hL, huL, bL), cell_side(hR, huR, bR));
}

}  // namespace dev
}  // namespace swe
