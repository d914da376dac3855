// mlp.hpp -- learned middle state for the augmented Riemann solver.
//
// A 6 -> 64 -> 64 -> 2 ReLU network predicts the middle depth h* and velocity
// u* of an edge's Riemann problem from scaled edge states.  In the solver it
// replaces the Newton iteration: a guess gets one FP32 Newton evaluation (the
// residual check); if the resulting step is small the corrected depth is
// taken, otherwise the exact iteration runs from its usual start (hMin or
// hMax).  The linearised (acoustic) middle state is tried first; the network
// runs for the waves in which it fails somewhere.  Dry, near-dry and
// two-rarefaction edges never use either guess.
//
// Inference runs on the matrix cores: the 64 lanes of a wave hold 64 edges and
// each layer is a chain of v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32
// accumulation) in the transposed form  H^T = W^T X^T  (rows = neurons,
// columns = edges), so that the accumulator layout of one layer is the B
// operand layout of the next one without any data movement.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (wave64, lane l, h = l / 32):
//   A (32 x 16): lane l holds A[l % 32][8h + m], m = 0..7
//   B (16 x 32): lane l holds B[8h + m][l % 32], m = 0..7
//   D (32 x 32): lane l, register r holds D[8(r/4) + 4h + r%4][l % 32]
// A K-chunk of 16 hidden neurons q of tile t is fed from registers 8q..8q+7 of
// the accumulator of tile t; position (h, m) of the chunk is neuron
//   32t + 16q + 8[m >= 4] + 4h + m % 4
// and the weight fragments are packed on the host in that order.
#pragma once

#include <hip/hip_runtime.h>

#include "solvers_dev.hpp"

namespace swe {
namespace mlp {

constexpr int kIn = 6;
constexpr int kHid = 64;
constexpr int kOut = 2;
// FP32 parameter vector: W1[kIn][kHid], b1, W2[kHid][kHid], b2, W3[kHid][kOut], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kParams = kOffB3 + kOut;  // 4738
constexpr int kSampleFloats = kIn + kOut;
// Fragment buffer: kFragDwords dwords per lane, dword-major (frag[d * 64 + lane]).
constexpr int kFragL1 = 0;   // 2 output tiles x 4 dwords
constexpr int kFragL2 = 8;   // 2 output tiles x 5 chunks (4 hidden + bias) x 4 dwords
constexpr int kFragL3 = 48;  // 4 chunks x 4 dwords
constexpr int kFragB3 = 64;  // b3[0], b3[1] as float bits (same in every lane)
constexpr int kFragDwords = 66;
// The corrected guess is accepted when the Newton step taken from it is at
// most this fraction of the depth (Newton's error after such a step is of the
// order of the step squared).
constexpr float kAcceptStep = 2e-3f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Network inputs of one edge (both sides wet): depths over the mean depth hs,
// velocities and their jump over sqrt(g hs), bed step over hs (clamped).
// The network predicts h*/hs and u*/sqrt(g hs) as corrections to the
// linearised (acoustic) Riemann solution, so that its BF16 error applies to a
// second-order term only:
//   h*/hs ~ 1 - (uR - uL)/(2 cs) = 1 - x5/2
//   u*/cs ~ (uL + uR)/(2 cs) - (hR - hL)/(2 hs) = (x2 + x3)/2 - (x1 - x0)/2
__device__ __forceinline__ void features(float hL, float hR, float uL, float uR, float bL, float bR, float* x,
                                         float& hs, float& cs) {
    hs = 0.5f * (hL + hR);
    const float invHs = 1.f / hs;
    cs = sqrtf(kGravity * hs);
    const float invCs = 1.f / cs;
    x[0] = hL * invHs;
    x[1] = hR * invHs;
    x[2] = uL * invCs;
    x[3] = uR * invCs;
    x[4] = fminf(fmaxf((bR - bL) * invHs, -1.f), 1.f);
    x[5] = (uR - uL) * invCs;
}

__device__ __forceinline__ void linear_part(const float* x, float& l0, float& l1) {
    l0 = 1.f - 0.5f * x[5];
    l1 = 0.5f * (x[2] + x[3]) - 0.5f * (x[1] - x[0]);
}

// Four fragment dwords of this lane (global memory; the network is only
// evaluated for waves in which the linearised guess failed, so its weights
// are fetched at use instead of occupying registers for the whole kernel).
__device__ __forceinline__ u32x4 frag4(const unsigned* __restrict__ f, int d) {
    const int lane = threadIdx.x & 63;
    u32x4 r;
    r.x = f[(d + 0) * 64 + lane];
    r.y = f[(d + 1) * 64 + lane];
    r.z = f[(d + 2) * 64 + lane];
    r.w = f[(d + 3) * 64 + lane];
    return r;
}

__device__ __forceinline__ bf16x8 as_bf16(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ u32x4 as_u32(bf16x8 v) { return __builtin_bit_cast(u32x4, v); }

__device__ __forceinline__ u32x4 shfl_xor4(u32x4 v, int m) {
    u32x4 r;
    r.x = (unsigned)__shfl_xor((int)v.x, m, 64);
    r.y = (unsigned)__shfl_xor((int)v.y, m, 64);
    r.z = (unsigned)__shfl_xor((int)v.z, m, 64);
    r.w = (unsigned)__shfl_xor((int)v.w, m, 64);
    return r;
}

// ReLU of accumulator registers base..base+7, rounded to BF16: one K-chunk of
// the next layer's B operand.
__device__ __forceinline__ bf16x8 relu_chunk(const f32x16& d, int base) {
    bf16x8 r;
#pragma unroll
    for (int m = 0; m < 8; ++m) r[m] = (__bf16)fmaxf(d[base + m], 0.f);
    return r;
}

__device__ __forceinline__ f32x16 mfma(u32x4 a, bf16x8 b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(a), b, c, 0, 0, 0);
}

// Network outputs (y0, y1) for the edge of every lane.  Must be called with
// the whole wave active (uniform control flow).  Per edge tile the layers
// are chained output tile by output tile, so that at most three accumulators
// (layer 1, layer 2 tile, layer 3) are live.  Not inlined: the solver calls
// it only for waves where the linearised guess failed, and a call keeps the
// network's registers out of the sweep's allocation (inlined, the sweep
// kernels needed four times the registers of the exact solver).
__device__ __attribute__((noinline)) float2 infer_wave(const unsigned* __restrict__ F, float x0, float x1, float x2,
                                                       float x3, float x4, float x5) {
    const int lane = threadIdx.x & 63;
    const bool upper = lane >= 32;
    // inputs: 6 features, the constant 1 (carries b1), 0; split into a BF16
    // high part and a BF16 residual so that the 16-deep K of layer 1 sees
    // about 16 significant bits of every input
    const float v[8] = {x0, x1, x2, x3, x4, x5, 1.f, 0.f};
    bf16x8 hi, lo;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        hi[m] = (__bf16)v[m];
        lo[m] = (__bf16)(v[m] - (float)hi[m]);
    }
    const u32x4 H = as_u32(hi), L = as_u32(lo);
    const u32x4 pH = shfl_xor4(H, 32), pL = shfl_xor4(L, 32);
    // edge tile e holds the edges of lanes 32e..32e+31: rows 0..7 of its B
    // operand are the high parts, rows 8..15 the residuals
    const u32x4 B[2] = {upper ? pL : H, upper ? L : pH};
    bf16x8 biasB;
#pragma unroll
    for (int m = 0; m < 8; ++m) biasB[m] = (__bf16)((!upper && m == 0) ? 1.f : 0.f);
    const f32x16 zero = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float out[2][2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const f32x16 d10 = mfma(frag4(F, kFragL1 + 0), as_bf16(B[e]), zero);
        const f32x16 d11 = mfma(frag4(F, kFragL1 + 4), as_bf16(B[e]), zero);
        f32x16 d3 = zero;
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
            // keep each stage's fragment loads next to their MFMAs (hoisting
            // all of them would hold 64 more registers for the whole chain)
            __builtin_amdgcn_sched_barrier(0);
            const int l2 = kFragL2 + 20 * t2;
            f32x16 acc = mfma(frag4(F, l2 + 16), biasB, zero);  // b2 via the constant chunk
            acc = mfma(frag4(F, l2 + 0), relu_chunk(d10, 0), acc);
            acc = mfma(frag4(F, l2 + 4), relu_chunk(d10, 8), acc);
            acc = mfma(frag4(F, l2 + 8), relu_chunk(d11, 0), acc);
            acc = mfma(frag4(F, l2 + 12), relu_chunk(d11, 8), acc);
            __builtin_amdgcn_sched_barrier(0);
            // layer-3 chunks 2 t2 and 2 t2 + 1 come from this output tile
            d3 = mfma(frag4(F, kFragL3 + 8 * t2 + 0), relu_chunk(acc, 0), d3);
            d3 = mfma(frag4(F, kFragL3 + 8 * t2 + 4), relu_chunk(acc, 8), d3);
        }
        __builtin_amdgcn_sched_barrier(0);
        // rows 0 and 1 (the two outputs) sit in registers 0 and 1 of lanes 0..31
        out[e][0] = d3[0];
        out[e][1] = d3[1];
    }
    const float q0 = __shfl_xor(out[1][0], 32, 64);
    const float q1 = __shfl_xor(out[1][1], 32, 64);
    float2 y;
    y.x = (upper ? q0 : out[0][0]) + __uint_as_float(F[kFragB3 * 64 + lane]);
    y.y = (upper ? q1 : out[0][1]) + __uint_as_float(F[(kFragB3 + 1) * 64 + lane]);
    return y;
}

__device__ __forceinline__ void infer(const unsigned* __restrict__ F, const float* x, float& y0, float& y1) {
    const float2 y = infer_wave(F, x[0], x[1], x[2], x[3], x[4], x[5]);
    y0 = y.x;
    y1 = y.y;
}

// Case analysis of the exact middle_state (dry, two rarefactions, Newton).
struct Bracket {
    float hMin, hMax, du, cMax, invMin, invMax;
    bool newton, shocks;
};

__device__ __forceinline__ Bracket classify(float hL, float hR, float uL, float uR, float cL, float cR, float invL,
                                            float invR) {
    using namespace dev;
    Bracket k;
    k.hMin = smin(hL, hR);
    k.hMax = smax(hL, hR);
    k.du = uR - uL;
    const float cMin = (hR < hL) ? cR : cL;
    k.cMax = (hL < hR) ? cR : cL;
    k.invMin = (hR < hL) ? invR : invL;
    k.invMax = (hL < hR) ? invR : invL;
    const float fMin = k.du + 2.f * (cMin - k.cMax);
    const float fMax = k.du + (k.hMax - k.hMin) * sqrtf(0.5f * kGravity * (k.invMax + k.invMin));
    k.newton = !(k.hMin < kDryTol) && !(fMin > 0.f);
    k.shocks = fMax <= 0.f;
    return k;
}

struct Eval {
    float h, gMin, s2;
    bool ok;
};

// One Newton evaluation of the exact iteration at a guess moved into the
// bracket the iteration starts from (two shocks: h >= hMax; shock and
// rarefaction: hMin <= h <= hMax).  The corrected depth is accepted when the
// step is at most kAcceptStep of the depth: Newton's error after such a step
// is of the order of the step squared.
__device__ __forceinline__ Eval eval_guess(const Bracket& k, float hGuess) {
    const float halfG = 0.5f * kGravity;
    const float quarterG = 0.25f * kGravity;
    const bool shocks = k.shocks;
    float h = dev::smax(hGuess, shocks ? k.hMax : k.hMin);
    if (!shocks) h = dev::smin(h, k.hMax);
    const float invH = 1.f / h;
    const float gMin = sqrtf(halfG * (invH + k.invMin));
    const float s2 = sqrtf(shocks ? halfG * (invH + k.invMax) : kGravity * h);
    const float t2 = shocks ? (h - k.hMax) * s2 : 2.f * (s2 - k.cMax);
    const float phi = (k.du + (h - k.hMin) * gMin) + t2;
    const float q = shocks ? s2 : 1.f;
    const float num = (h - k.hMin) * q + (shocks ? (h - k.hMax) * gMin : 0.f);
    const float den = gMin * q;
    const float first = gMin + (shocks ? s2 : s2 * invH);
    // the exact iteration's step form: phi / dphi with both sides scaled by den
    const float step = (phi * den) / (first * den - quarterG * invH * invH * num);
    Eval e;
    e.h = h - step;
    e.gMin = gMin;
    e.s2 = s2;
    e.ok = (fabsf(step) <= kAcceptStep * h) && (e.h > 0.f);
    return e;
}

// Speeds of the middle state from an accepted evaluation, as after the exact
// iteration (curves evaluated at the point the last step was taken from).
__device__ __forceinline__ dev::Middle finish(const Bracket& k, const Eval& e, float hL, float hR, float uL, float uR,
                                              float cL, float cR) {
    dev::Middle m;
    const float h = e.h;
    m.hm = h;
    const float cm = sqrtf(kGravity * h);
    if (k.shocks) {
        const bool leftIsMin = !(hR < hL);
        const float gL = leftIsMin ? e.gMin : e.s2;
        const float gR = leftIsMin ? e.s2 : e.gMin;
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

// Learned middle state for the edge of every lane (whole wave active).
// Newton edges first try the linearised guess; if any of them in the wave
// fails the residual check, the network runs for the wave and the failed
// edges try its guess; what still fails runs the exact iteration.
// *how: 0 exact middle_state, 1 linearised guess accepted, 2 network guess.
__device__ __forceinline__ dev::Middle learned_middle(const unsigned* __restrict__ F, float hL, float hR, float uL,
                                                      float uR, float cL, float cR, float invL, float invR, float bL,
                                                      float bR, int* how) {
    // nothing moves across the learned part: two edge solves interleaved by
    // the scheduler would each hold their network state at once
    __builtin_amdgcn_sched_barrier(0);
    const Bracket k = classify(hL, hR, uL, uR, cL, cR, invL, invR);
    float x[kIn], hs, cs;
    features(k.newton ? hL : 1.f, k.newton ? hR : 1.f, k.newton ? uL : 0.f, k.newton ? uR : 0.f, bL, bR, x, hs, cs);
    float l0, l1;
    linear_part(x, l0, l1);
    Eval ev = eval_guess(k, l0 * hs);
    int used = (k.newton && ev.ok) ? 1 : 0;
    if (__any(k.newton && !ev.ok)) {
        float y0, y1;
        infer(F, x, y0, y1);
        if (k.newton && !ev.ok) {
            const Eval e2 = eval_guess(k, (l0 + y0) * hs);
            if (e2.ok) {
                ev = e2;
                used = 2;
            }
        }
    }
    *how = used;
    __builtin_amdgcn_sched_barrier(0);
    if (used) return finish(k, ev, hL, hR, uL, uR, cL, cR);
    return dev::middle_state(hL, hR, uL, uR, cL, cR, invL, invR);
}

struct LearnedMiddle {
    static constexpr bool kDefers = false;
    const unsigned* F;
    __device__ __forceinline__ dev::Middle operator()(float hL, float hR, float uL, float uR, float cL, float cR,
                                                      float invL, float invR, float bL, float bR) const {
        int how;
        return learned_middle(F, hL, hR, uL, uR, cL, cR, invL, invR, bL, bR, &how);
    }
};

// The first stage of learned_middle only: a Newton edge whose linearised
// guess passes the residual check is finished here; a Newton edge whose guess
// fails is deferred (negative depth), to be solved by learned_middle in a
// later launch, which takes the same decisions and so gives the same result.
// Dry and two-rarefaction edges take the exact closed forms.  No network call
// and no Newton loop for wet edges stay in the sweep.
struct DeferredMiddle {
    static constexpr bool kDefers = true;
    __device__ __forceinline__ dev::Middle operator()(float hL, float hR, float uL, float uR, float cL, float cR,
                                                      float invL, float invR, float bL, float bR) const {
        const Bracket k = classify(hL, hR, uL, uR, cL, cR, invL, invR);
        float x[kIn], hs, cs;
        features(k.newton ? hL : 1.f, k.newton ? hR : 1.f, k.newton ? uL : 0.f, k.newton ? uR : 0.f, bL, bR, x, hs, cs);
        float l0, l1;
        linear_part(x, l0, l1);
        const Eval ev = eval_guess(k, l0 * hs);
        if (k.newton && ev.ok) return finish(k, ev, hL, hR, uL, uR, cL, cR);
        if (k.newton) return dev::Middle{-1.f, 0.f, 0.f};
        return dev::middle_state(hL, hR, uL, uR, cL, cR, invL, invR);
    }
};

}  // namespace mlp

namespace dev {
template <bool ReuseFirst>
struct SolverCtx<SOLVER_AUGRIE_DEFER, ReuseFirst> {
    __device__ __forceinline__ explicit SolverCtx(const unsigned*) {}
    __device__ __forceinline__ EdgeOut solve(const Side& L, const Side& R) const {
        return augrie_with(L, R, mlp::DeferredMiddle{});
    }
};

template <bool ReuseFirst>
struct SolverCtx<SOLVER_AUGRIE_MLP, ReuseFirst> {
    const unsigned* F;
    __device__ __forceinline__ explicit SolverCtx(const unsigned* frags) : F(frags) {}
    __device__ __forceinline__ EdgeOut solve(const Side& L, const Side& R) const {
        return augrie_with(L, R, mlp::LearnedMiddle{F});
    }
};
}  // namespace dev

}  // namespace swe
