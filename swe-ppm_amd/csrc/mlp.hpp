// mlp.hpp -- learned middle state for the augmented Riemann solver.
//
// A 6 -> 64 -> 64 -> 2 ReLU network predicts the middle depth h* and velocity
// u* of an edge's Riemann problem from scaled edge states.  In the solver it
// replaces the Newton iteration's starting point: the guess gets one FP32
// Newton evaluation (the residual check); if the resulting step is small the
// corrected depth is taken, otherwise the exact iteration runs from its usual
// start (hMin or hMax).  Dry, near-dry and two-rarefaction edges never use the
// network.
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

struct MlpFrags {
    u32x4 l1[2];
    u32x4 l2[2][5];
    u32x4 l3[4];
    float b3[2];

    __device__ __forceinline__ static u32x4 ld4(const unsigned* f, int d, int lane) {
        u32x4 r;
        r.x = f[(d + 0) * 64 + lane];
        r.y = f[(d + 1) * 64 + lane];
        r.z = f[(d + 2) * 64 + lane];
        r.w = f[(d + 3) * 64 + lane];
        return r;
    }
    __device__ __forceinline__ void load(const unsigned* f) {
        const int lane = threadIdx.x & 63;
        for (int t = 0; t < 2; ++t) l1[t] = ld4(f, kFragL1 + 4 * t, lane);
        for (int t = 0; t < 2; ++t)
            for (int c = 0; c < 5; ++c) l2[t][c] = ld4(f, kFragL2 + 4 * (5 * t + c), lane);
        for (int c = 0; c < 4; ++c) l3[c] = ld4(f, kFragL3 + 4 * c, lane);
        b3[0] = __uint_as_float(f[kFragB3 * 64 + lane]);
        b3[1] = __uint_as_float(f[(kFragB3 + 1) * 64 + lane]);
    }
};

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

// Network outputs (y0, y1) for the edge of every lane.  Must be called with
// the whole wave active (uniform control flow).
__device__ __forceinline__ void infer(const MlpFrags& F, const float* x, float& y0, float& y1) {
    const int lane = threadIdx.x & 63;
    const bool upper = lane >= 32;
    // inputs: 6 features, the constant 1 (carries b1), 0; split into a BF16
    // high part and a BF16 residual so that the 16-deep K of layer 1 sees
    // about 16 significant bits of every input
    const float v[8] = {x[0], x[1], x[2], x[3], x[4], x[5], 1.f, 0.f};
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
        f32x16 d1[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
            d1[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(F.l1[t]), as_bf16(B[e]), zero, 0, 0, 0);
        f32x16 d2[2];
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
            f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(F.l2[t2][4]), biasB, zero, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(F.l2[t2][c]), relu_chunk(d1[c >> 1], 8 * (c & 1)),
                                                              acc, 0, 0, 0);
            d2[t2] = acc;
        }
        f32x16 d3 = zero;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            d3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(F.l3[c]), relu_chunk(d2[c >> 1], 8 * (c & 1)), d3, 0, 0,
                                                         0);
        // rows 0 and 1 (the two outputs) sit in registers 0 and 1 of lanes 0..31
        out[e][0] = d3[0];
        out[e][1] = d3[1];
    }
    const float q0 = __shfl_xor(out[1][0], 32, 64);
    const float q1 = __shfl_xor(out[1][1], 32, 64);
    y0 = (upper ? q0 : out[0][0]) + F.b3[0];
    y1 = (upper ? q1 : out[0][1]) + F.b3[1];
}

// Middle state from a guess of h*: the same case analysis as the exact
// middle_state; for the Newton cases one evaluation at the guess (moved into
// the bracket the exact iteration starts from), the corrected depth is kept
// if the step is at most kAcceptStep of the depth, else the exact iteration
// runs.  *accepted tells which.
__device__ __forceinline__ dev::Middle middle_from_guess(float hL, float hR, float uL, float uR, float cL, float cR,
                                                        float hGuess, int* accepted) {
    using namespace dev;
    const float hMin = smin(hL, hR);
    const float hMax = smax(hL, hR);
    const float du = uR - uL;
    *accepted = 0;
    if (hMin < kDryTol) return middle_state(hL, hR, uL, uR, cL, cR);
    const float cMin = (hR < hL) ? cR : cL;
    const float cMax = (hL < hR) ? cR : cL;
    const float invMin = 1.f / hMin;
    const float invMax = 1.f / hMax;
    const float halfG = 0.5f * kGravity;
    const float quarterG = 0.25f * kGravity;
    const float fMin = du + 2.f * (cMin - cMax);
    if (fMin > 0.f) return middle_state(hL, hR, uL, uR, cL, cR);
    const float fMax = du + (hMax - hMin) * sqrtf(halfG * (invMax + invMin));
    const bool shocks = fMax <= 0.f;
    float h = smax(hGuess, shocks ? hMax : hMin);
    if (!shocks) h = smin(h, hMax);
    const float invH = 1.f / h;
    const float gMin = sqrtf(halfG * (invH + invMin));
    const float s2 = sqrtf(shocks ? halfG * (invH + invMax) : kGravity * h);
    const float t2 = shocks ? (h - hMax) * s2 : 2.f * (s2 - cMax);
    const float phi = (du + (h - hMin) * gMin) + t2;
    const float q = shocks ? s2 : 1.f;
    const float num = (h - hMin) * q + (shocks ? (h - hMax) * gMin : 0.f);
    const float den = gMin * q;
    const float first = gMin + (shocks ? s2 : s2 * invH);
    const float dphi = first - quarterG * invH * invH * (num / den);
    const float step = phi / dphi;
    const float hNew = h - step;
    if (!(fabsf(step) <= kAcceptStep * h) || !(hNew > 0.f)) return middle_state(hL, hR, uL, uR, cL, cR);
    *accepted = 1;
    h = hNew;
    Middle m;
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

// The network's guess for the edge of every lane (whole wave active); lanes
// whose edge has a dry side get a dummy wet state.
__device__ __forceinline__ float guess_depth(const MlpFrags& F, float hL, float hR, float uL, float uR, float bL,
                                             float bR) {
    const bool wet = dev::smin(hL, hR) >= kDryTol;
    float x[kIn], hs, cs;
    features(wet ? hL : 1.f, wet ? hR : 1.f, wet ? uL : 0.f, wet ? uR : 0.f, bL, bR, x, hs, cs);
    float y0, y1;
    infer(F, x, y0, y1);
    float l0, l1;
    linear_part(x, l0, l1);
    return (l0 + y0) * hs;
}

struct LearnedMiddle {
    const MlpFrags* F;
    __device__ __forceinline__ dev::Middle operator()(float hL, float hR, float uL, float uR, float cL, float cR,
                                                      float bL, float bR) const {
        const float g = guess_depth(*F, hL, hR, uL, uR, bL, bR);
        int accepted;
        return middle_from_guess(hL, hR, uL, uR, cL, cR, g, &accepted);
    }
};

}  // namespace mlp

namespace dev {
template <>
struct SolverCtx<SOLVER_AUGRIE_MLP> {
    mlp::MlpFrags f;
    __device__ __forceinline__ explicit SolverCtx(const unsigned* frags) { f.load(frags); }
    __device__ __forceinline__ EdgeOut solve(const Side& L, const Side& R) const {
        return augrie_with(L, R, mlp::LearnedMiddle{&f});
    }
};
}  // namespace dev

}  // namespace swe
