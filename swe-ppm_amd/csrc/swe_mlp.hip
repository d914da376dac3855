// swe_mlp.hip -- learned middle state: sample logging from the solver,
// training (FP32 forward / backward / Adam kernels), BF16 matrix-core
// inference entry points and the C ABI of include/swe_hip_mlp.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <mutex>
#include <vector>

#include "../../include/swe_hip_mlp.h"
#include "kernels_u.hpp"
#include "mlp.hpp"
#include "swe_internal.h"

using namespace swe;
using namespace swe::mlp;

#define HIP_TRY(expr)                                     \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

static_assert(kParams == SWE_MLP_PARAMS, "parameter count");
static_assert(kSampleFloats == SWE_MLP_SAMPLE_FLOATS, "sample size");

namespace {

// activations kept per sample for the weight gradients (SoA rows, stride = batch)
constexpr int kRowX = 0, kRowH1 = kRowX + kIn, kRowH2 = kRowH1 + kHid, kRowDH1 = kRowH2 + kHid,
              kRowDH2 = kRowDH1 + kHid, kRowDY = kRowDH2 + kHid, kRows = kRowDY + kOut;  // 264
constexpr int kWgradChunk = 128;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// ---- logging: every Newton edge of a block state, kept with probability `keep`
__global__ __launch_bounds__(256) void mlp_log_kernel(StateView v, float keep, unsigned long long seed, int hardOnly,
                                                      float* samples, long long cap, unsigned long long* count) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nV = (long long)(v.nx + 1) * v.ny;
    const long long nH = (long long)v.nx * (v.ny + 1);
    if (k >= nV + nH) return;
    if ((double)(mix64(seed ^ (unsigned long long)k) >> 11) * (1.0 / 9007199254740992.0) >= (double)keep) return;
    size_t cl, cr;
    bool vertical = k < nV;
    if (vertical) {
        const int x = (int)(k / v.ny), y = (int)(k % v.ny) + 1;
        cl = idx(v, x, y);
        cr = idx(v, x + 1, y);
    } else {
        const long long k2 = k - nV;
        const int x = (int)(k2 / (v.ny + 1)) + 1, y = (int)(k2 % (v.ny + 1));
        cl = idx(v, x, y);
        cr = idx(v, x, y + 1);
    }
    const dev::Side L = dev::cell_side(v.h[cl], vertical ? v.hu[cl] : v.hv[cl], v.b[cl]);
    const dev::Side R = dev::cell_side(v.h[cr], vertical ? v.hu[cr] : v.hv[cr], v.b[cr]);
    const dev::AugPrep p = dev::augrie_prepare(L, R);
    if (p.dryBoth) return;
    const float hMin = dev::smin(p.hL, p.hR);
    if (hMin < kDryTol) return;
    const float cMin = (p.hR < p.hL) ? p.cR : p.cL;
    const float cMax = (p.hL < p.hR) ? p.cR : p.cL;
    if ((p.uR - p.uL) + 2.f * (cMin - cMax) > 0.f) return;  // two rarefactions: closed form, no iteration
    float x[kIn], hs, cs;
    features(p.hL, p.hR, p.uL, p.uR, p.bL, p.bR, x, hs, cs);
    float l0, l1;
    linear_part(x, l0, l1);
    if (hardOnly) {
        // only edges whose linearised guess fails the residual check
        const Bracket k = classify(p.hL, p.hR, p.uL, p.uR, p.cL, p.cR, p.invL, p.invR);
        if (eval_guess(k, l0 * hs).ok) return;
    }
    const dev::Middle m = dev::middle_state(p.hL, p.hR, p.uL, p.uR, p.cL, p.cR, p.invL, p.invR);
    const unsigned long long at = atomicAdd(count, 1ULL);
    if ((long long)at >= cap) return;
    float* o = samples + (size_t)at * kSampleFloats;
    for (int i = 0; i < kIn; ++i) o[i] = x[i];
    o[kIn] = m.hm / hs - l0;
    o[kIn + 1] = 0.5f * (m.s1m + m.s2m) / cs - l1;
}

// ---- training: forward + backward of one sample per thread (FP32) ---------
__global__ __launch_bounds__(256) void mlp_fwdbwd_kernel(const float* __restrict__ P, const float* __restrict__ samples,
                                                         long long nsamples, int batch, unsigned long long seed,
                                                         int step, int randomIdx, float* __restrict__ act,
                                                         float* lossOut) {
    __shared__ float sP[kParams];
    __shared__ float sLoss[4];
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float lossv = 0.f;
    if (b < batch) {
        const long long s =
            randomIdx ? (long long)(mix64(seed * 0x100000001B3ULL + (unsigned long long)step * 1000003ULL +
                                          (unsigned long long)b) %
                                    (unsigned long long)nsamples)
                      : (long long)b;
        const float* smp = samples + (size_t)s * kSampleFloats;
        float x[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) x[k] = smp[k];
        const float t0 = smp[kIn], t1 = smp[kIn + 1];
        float h1[kHid];
#pragma unroll
        for (int i = 0; i < kHid; ++i) {
            float a = sP[kOffB1 + i];
#pragma unroll
            for (int k = 0; k < kIn; ++k) a += x[k] * sP[kOffW1 + k * kHid + i];
            h1[i] = fmaxf(a, 0.f);
        }
        float h2[kHid];
#pragma unroll
        for (int j = 0; j < kHid; ++j) h2[j] = sP[kOffB2 + j];
#pragma unroll
        for (int i = 0; i < kHid; ++i) {
            const float hv = h1[i];
#pragma unroll
            for (int j = 0; j < kHid; ++j) h2[j] += hv * sP[kOffW2 + i * kHid + j];
        }
        float y0 = sP[kOffB3], y1 = sP[kOffB3 + 1];
#pragma unroll
        for (int j = 0; j < kHid; ++j) {
            h2[j] = fmaxf(h2[j], 0.f);
            y0 += h2[j] * sP[kOffW3 + 2 * j];
            y1 += h2[j] * sP[kOffW3 + 2 * j + 1];
        }
        const float e0 = y0 - t0, e1 = y1 - t1;
        lossv = e0 * e0 + e1 * e1;
        const float dy0 = 2.f * e0 / batch, dy1 = 2.f * e1 / batch;
        float* a = act + b;
#pragma unroll
        for (int k = 0; k < kIn; ++k) a[(size_t)(kRowX + k) * batch] = x[k];
#pragma unroll
        for (int i = 0; i < kHid; ++i) a[(size_t)(kRowH1 + i) * batch] = h1[i];
#pragma unroll
        for (int j = 0; j < kHid; ++j) a[(size_t)(kRowH2 + j) * batch] = h2[j];
        a[(size_t)kRowDY * batch] = dy0;
        a[(size_t)(kRowDY + 1) * batch] = dy1;
        // dh2 (in place of h2), then dh1 (in place of h1)
#pragma unroll
        for (int j = 0; j < kHid; ++j)
            h2[j] = h2[j] > 0.f ? sP[kOffW3 + 2 * j] * dy0 + sP[kOffW3 + 2 * j + 1] * dy1 : 0.f;
#pragma unroll
        for (int j = 0; j < kHid; ++j) a[(size_t)(kRowDH2 + j) * batch] = h2[j];
#pragma unroll
        for (int i = 0; i < kHid; ++i) {
            float g = 0.f;
#pragma unroll
            for (int j = 0; j < kHid; ++j) g += sP[kOffW2 + i * kHid + j] * h2[j];
            h1[i] = h1[i] > 0.f ? g : 0.f;
        }
#pragma unroll
        for (int i = 0; i < kHid; ++i) a[(size_t)(kRowDH1 + i) * batch] = h1[i];
    }
    // workgroup sum of the per-sample losses
    for (int off = 32; off >= 1; off >>= 1) lossv += __shfl_xor(lossv, off, 64);
    if ((threadIdx.x & 63) == 0) sLoss[threadIdx.x >> 6] = lossv;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(lossOut, (sLoss[0] + sLoss[1] + sLoss[2] + sLoss[3]) / batch);
}

// ---- weight gradients: partial sums over a chunk of samples per workgroup
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const float* __restrict__ act, int batch,
                                                        float* __restrict__ partials) {
    __shared__ float s[kRows][65];
    const int t = threadIdx.x;
    const int s0 = blockIdx.x * kWgradChunk;
    const int s1 = min(s0 + kWgradChunk, batch);
    const int i2 = t >> 2, jb = (t & 3) * 16;  // dW2[i2][jb..jb+15]
    float w2[16];
    for (int u = 0; u < 16; ++u) w2[u] = 0.f;
    float w1a = 0.f, w1b = 0.f, bia = 0.f, w3 = 0.f, b3 = 0.f;
    for (int base = s0; base < s1; base += 64) {
        const int n = min(64, s1 - base);
        __syncthreads();
        for (int e = t; e < kRows * 64; e += blockDim.x) {
            const int row = e >> 6, c = e & 63;
            s[row][c] = c < n ? act[(size_t)row * batch + base + c] : 0.f;
        }
        __syncthreads();
        for (int c = 0; c < n; ++c) {
            const float hv = s[kRowH1 + i2][c];
#pragma unroll
            for (int u = 0; u < 16; ++u) w2[u] += hv * s[kRowDH2 + jb + u][c];
            // dW1[k][i] = sum x[k] dh1[i], entries t and t + 256 of 384
            w1a += s[kRowX + (t >> 6)][c] * s[kRowDH1 + (t & 63)][c];
            if (t < kIn * kHid - 256) w1b += s[kRowX + ((t + 256) >> 6)][c] * s[kRowDH1 + ((t + 256) & 63)][c];
            if (t < 64)
                bia += s[kRowDH1 + t][c];
            else if (t < 128)
                bia += s[kRowDH2 + t - 64][c];
            else
                w3 += s[kRowH2 + ((t - 128) >> 1)][c] * s[kRowDY + ((t - 128) & 1)][c];
            if (t < 2) b3 += s[kRowDY + t][c];
        }
    }
    float* out = partials + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int u = 0; u < 16; ++u) out[kOffW2 + i2 * kHid + jb + u] = w2[u];
    out[kOffW1 + t] = w1a;
    if (t < kIn * kHid - 256) out[kOffW1 + t + 256] = w1b;
    if (t < 64)
        out[kOffB1 + t] = bia;
    else if (t < 128)
        out[kOffB2 + t - 64] = bia;
    else
        out[kOffW3 + (t - 128)] = w3;
    if (t < 2) out[kOffB3 + t] = b3;
}

// ---- gradient reduction and Adam update (or gradient output) --------------
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* P, float* M, float* V, const float* partials, int nparts,
                                                       float lr, float bc1, float bc2, float* gradOut) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float g = 0.f;
    for (int w = 0; w < nparts; ++w) g += partials[(size_t)w * kParams + p];
    if (gradOut) {
        gradOut[p] = g;
        return;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// ---- inference entry points (BF16 matrix cores) ----------------------------
// One wave per 64 samples; every lane of a wave stays active (lanes past the
// end repeat the last sample and do not store).
__global__ __launch_bounds__(256) void mlp_infer_kernel(const unsigned* __restrict__ frags,
                                                        const float* __restrict__ feats, long long n,
                                                        float* __restrict__ out) {
    const long long base = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (base >= n) return;  // uniform per wave
    const long long s = base + (threadIdx.x & 63);
    const long long sc = s < n ? s : n - 1;
    float x[kIn];
    for (int k = 0; k < kIn; ++k) x[k] = feats[sc * kIn + k];
    float y0, y1;
    infer(frags, x, y0, y1);
    if (s < n) {
        out[2 * s] = y0;
        out[2 * s + 1] = y1;
    }
}

__global__ __launch_bounds__(256) void mlp_features_kernel(long long n, const float* hL, const float* hR,
                                                           const float* uL, const float* uR, const float* bL,
                                                           const float* bR, float* feats) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[kIn], hs, cs;
    features(hL[s], hR[s], uL[s], uR[s], bL[s], bR[s], x, hs, cs);
    for (int k = 0; k < kIn; ++k) feats[s * kIn + k] = x[k];
}

__global__ __launch_bounds__(256) void mlp_middle_kernel(const unsigned* __restrict__ frags, long long n,
                                                         const float* in, float* out, int* accepted) {
    const long long base = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (base >= n) return;
    const long long s = base + (threadIdx.x & 63);
    const long long sc = s < n ? s : n - 1;
    const float hL = in[sc], hR = in[n + sc], uL = in[2 * n + sc], uR = in[3 * n + sc];
    const float bL = in[4 * n + sc], bR = in[5 * n + sc];
    const float cL = kSqrtGravity * sqrtf(hL), cR = kSqrtGravity * sqrtf(hR);
    int acc;
    const dev::Middle m = learned_middle(frags, hL, hR, uL, uR, cL, cR, 1.f / hL, 1.f / hR, bL, bR, &acc);
    if (s < n) {
        out[s] = m.hm;
        out[n + s] = m.s1m;
        out[2 * n + s] = m.s2m;
        accepted[s] = acc;
    }
}

__global__ __launch_bounds__(256) void mlp_solve_edges_kernel(const unsigned* __restrict__ frags, long long n,
                                                              const float* __restrict__ in, float* __restrict__ out) {
    const long long base = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (base >= n) return;
    const dev::SolverCtx<SOLVER_AUGRIE_MLP> solver(frags);
    const long long s = base + (threadIdx.x & 63);
    const long long k = s < n ? s : n - 1;
    const dev::EdgeOut e = solver.solve(dev::cell_side(in[k], in[2 * n + k], in[4 * n + k]),
                                        dev::cell_side(in[n + k], in[3 * n + k], in[5 * n + k]));
    if (s < n) {
        out[s] = e.hL;
        out[n + s] = e.hR;
        out[2 * n + s] = e.huL;
        out[3 * n + s] = e.huR;
        out[4 * n + s] = e.speed;
    }
}

// ---- host helpers ----------------------------------------------------------
uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// neuron of a hidden K-chunk position (see mlp.hpp): tile t, half q, lane half h, element m
int chunk_neuron(int t, int q, int h, int m) { return 32 * t + 16 * q + 8 * (m >= 4) + 4 * h + (m & 3); }

void pack_frags(const float* P, std::vector<uint32_t>& F) {
    F.assign((size_t)kFragDwords * 64, 0u);
    auto put = [&](int d, int lane, int m, float val) {
        uint32_t& w = F[(size_t)(d + (m >> 1)) * 64 + lane];
        const uint32_t b = bf16_rne(val);
        w = (m & 1) ? ((w & 0x0000ffffu) | (b << 16)) : ((w & 0xffff0000u) | b);
    };
    for (int lane = 0; lane < 64; ++lane) {
        const int h = lane >> 5, i = lane & 31;
        for (int m = 0; m < 8; ++m) {
            for (int t = 0; t < 2; ++t) {
                const int o = 32 * t + i;
                const float w1 = m < kIn ? P[kOffW1 + m * kHid + o] : (m == kIn ? P[kOffB1 + o] : 0.f);
                put(kFragL1 + 4 * t, lane, m, w1);
            }
            for (int t2 = 0; t2 < 2; ++t2) {
                const int o = 32 * t2 + i;
                for (int c = 0; c < 4; ++c) {
                    const int in = chunk_neuron(c >> 1, c & 1, h, m);
                    put(kFragL2 + 4 * (5 * t2 + c), lane, m, P[kOffW2 + in * kHid + o]);
                }
                put(kFragL2 + 4 * (5 * t2 + 4), lane, m, (h == 0 && m == 0) ? P[kOffB2 + o] : 0.f);
            }
            for (int c = 0; c < 4; ++c) {
                const int in = chunk_neuron(c >> 1, c & 1, h, m);
                put(kFragL3 + 4 * c, lane, m, i < kOut ? P[kOffW3 + in * kOut + i] : 0.f);
            }
        }
        uint32_t b0, b1;
        std::memcpy(&b0, &P[kOffB3], 4);
        std::memcpy(&b1, &P[kOffB3 + 1], 4);
        F[(size_t)kFragB3 * 64 + lane] = b0;
        F[(size_t)(kFragB3 + 1) * 64 + lane] = b1;
    }
}

}  // namespace

struct swe_hip_mlp {
    float* P = nullptr;  // FP32 master parameters
    float* M = nullptr;  // Adam moments
    float* V = nullptr;
    unsigned* frags = nullptr;
    float* samples = nullptr;
    long long cap = 0;
    unsigned long long* dCount = nullptr;
    float* act = nullptr;
    int actBatch = 0;
    float* partials = nullptr;
    float* lossLog = nullptr;
    int lossCap = 0;
    int adamStep = 0;
    hipStream_t stream = nullptr;
    // streams of the blocks and batches the network is attached to (one
    // entry per attachment): destroy is refused while any remain, and a
    // repack waits for their queued work before rewriting the weights
    std::mutex usersMu;
    std::vector<hipStream_t> users;
};

namespace swe {
const unsigned* mlp_frags(swe_hip_mlp* m) { return m->frags; }

void mlp_attach(swe_hip_mlp* m, hipStream_t user) {
    std::lock_guard<std::mutex> lock(m->usersMu);
    m->users.push_back(user);
}

void mlp_detach(swe_hip_mlp* m, hipStream_t user) {
    std::lock_guard<std::mutex> lock(m->usersMu);
    auto it = std::find(m->users.begin(), m->users.end(), user);
    if (it != m->users.end()) m->users.erase(it);
}
}  // namespace swe

namespace {

int repack(swe_hip_mlp* m) {
    {
        // no kernel of an attached block may read the fragments while they change
        std::lock_guard<std::mutex> lock(m->usersMu);
        for (hipStream_t s : m->users) HIP_TRY(hipStreamSynchronize(s));
    }
    std::vector<float> P(kParams);
    HIP_TRY(hipMemcpyAsync(P.data(), m->P, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    std::vector<uint32_t> F;
    pack_frags(P.data(), F);
    HIP_TRY(hipMemcpyAsync(m->frags, F.data(), sizeof(uint32_t) * F.size(), hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

long long count_of(swe_hip_mlp* m) {
    unsigned long long c = 0;
    if (hipMemcpy(&c, m->dCount, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return (long long)std::min<unsigned long long>(c, (unsigned long long)m->cap);
}

int ensure_act(swe_hip_mlp* m, int batch) {
    if (batch <= m->actBatch) return 0;
    if (m->act) (void)hipFree(m->act);
    if (m->partials) (void)hipFree(m->partials);
    m->act = nullptr;
    m->partials = nullptr;
    m->actBatch = 0;
    const int nparts = (batch + kWgradChunk - 1) / kWgradChunk;
    HIP_TRY(hipMalloc(&m->act, sizeof(float) * (size_t)kRows * batch));
    HIP_TRY(hipMalloc(&m->partials, sizeof(float) * (size_t)nparts * kParams));
    m->actBatch = batch;
    return 0;
}

int ensure_loss(swe_hip_mlp* m, int n) {
    if (n <= m->lossCap) return 0;
    if (m->lossLog) (void)hipFree(m->lossLog);
    m->lossLog = nullptr;
    HIP_TRY(hipMalloc(&m->lossLog, sizeof(float) * n));
    m->lossCap = n;
    return 0;
}

// forward/backward + weight gradients of one batch; then Adam (grad == null)
// or the gradient into grad (device)
int train_step(swe_hip_mlp* m, const float* samples, long long nsamples, int batch, unsigned long long seed, int step,
               int randomIdx, float* lossSlot, float lr, float* gradDev) {
    const int groups = (batch + 255) / 256;
    hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(groups), dim3(256), 0, m->stream, m->P, samples, nsamples, batch, seed,
                       step, randomIdx, m->act, lossSlot);
    const int nparts = (batch + kWgradChunk - 1) / kWgradChunk;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(nparts), dim3(256), 0, m->stream, m->act, batch, m->partials);
    float bc1 = 1.f, bc2 = 1.f;
    if (!gradDev) {
        m->adamStep += 1;
        bc1 = 1.f - std::pow(0.9f, (float)m->adamStep);
        bc2 = 1.f - std::pow(0.999f, (float)m->adamStep);
    }
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->P, m->M, m->V,
                       m->partials, nparts, lr, bc1, bc2, gradDev);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
int upload_tmp(T** dst, const T* src, size_t n, hipStream_t s) {
    HIP_TRY(hipMalloc(dst, sizeof(T) * std::max<size_t>(n, 1)));
    HIP_TRY(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return 0;
}

}  // namespace

extern "C" {

int swe_hip_mlp_create(int64_t capacity, uint64_t seed, swe_hip_mlp** out) {
    if (!out || capacity < 0 || capacity > (1LL << 33)) {
        set_error("swe_hip_mlp_create: invalid argument");
        return SWE_HIP_EINVAL;
    }
    swe_hip_mlp* m = new (std::nothrow) swe_hip_mlp();
    if (!m) {
        set_error("swe_hip_mlp_create: out of host memory");
        return SWE_HIP_EINVAL;
    }
    m->cap = capacity;
    // He initialisation (normal, std sqrt(2 / fan_in)) from a seeded generator, zero biases
    std::vector<float> P(kParams, 0.f);
    unsigned long long st = seed ^ 0x5DEECE66DULL;
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int i = 0; i < kIn * kHid; ++i) P[kOffW1 + i] = (float)(gauss() * std::sqrt(2.0 / kIn));
    for (int i = 0; i < kHid * kHid; ++i) P[kOffW2 + i] = (float)(gauss() * std::sqrt(2.0 / kHid));
    for (int i = 0; i < kHid * kOut; ++i) P[kOffW3 + i] = (float)(gauss() * std::sqrt(1.0 / kHid));
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&m->P, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&m->M, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&m->V, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMalloc(&m->frags, sizeof(unsigned) * kFragDwords * 64);
    if (e == hipSuccess) e = hipMalloc(&m->dCount, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&m->samples, sizeof(float) * kSampleFloats * (size_t)std::max<int64_t>(capacity, 1));
    if (e == hipSuccess) e = hipMemcpy(m->P, P.data(), sizeof(float) * kParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->M, 0, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMemset(m->V, 0, sizeof(float) * kParams);
    if (e == hipSuccess) e = hipMemset(m->dCount, 0, sizeof(unsigned long long));
    if (e != hipSuccess) {
        hip_fail(e, "swe_hip_mlp_create");
        swe_hip_mlp_destroy(m);
        return SWE_HIP_EDEVICE;
    }
    if (int r = repack(m)) {
        swe_hip_mlp_destroy(m);
        return r;
    }
    *out = m;
    return SWE_HIP_OK;
}

int swe_hip_mlp_destroy(swe_hip_mlp* m) {
    if (!m) return SWE_HIP_OK;
    {
        std::lock_guard<std::mutex> lock(m->usersMu);
        if (!m->users.empty()) {
            set_error("swe_hip_mlp_destroy: the network is attached to %zu block(s) or batch(es); detach it "
                      "(swe_hip_set_surrogate / swe_hip_batch_set_surrogate with NULL) or destroy them first",
                      m->users.size());
            return SWE_HIP_ESTATE;
        }
    }
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : {(void*)m->P, (void*)m->M, (void*)m->V, (void*)m->frags, (void*)m->samples, (void*)m->dCount,
                    (void*)m->act, (void*)m->partials, (void*)m->lossLog})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return SWE_HIP_OK;
}

int swe_hip_mlp_get_params(swe_hip_mlp* m, float* params) {
    if (!m || !params) {
        set_error("swe_hip_mlp_get_params: null argument");
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipMemcpyAsync(params, m->P, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return SWE_HIP_OK;
}

int swe_hip_mlp_set_params(swe_hip_mlp* m, const float* params) {
    if (!m || !params) {
        set_error("swe_hip_mlp_set_params: null argument");
        return SWE_HIP_EINVAL;
    }
    for (int i = 0; i < kParams; ++i)
        if (!std::isfinite(params[i])) {
            set_error("swe_hip_mlp_set_params: parameter %d is not finite", i);
            return SWE_HIP_EINVAL;
        }
    HIP_TRY(hipMemcpyAsync(m->P, params, sizeof(float) * kParams, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemsetAsync(m->M, 0, sizeof(float) * kParams, m->stream));
    HIP_TRY(hipMemsetAsync(m->V, 0, sizeof(float) * kParams, m->stream));
    m->adamStep = 0;
    return repack(m);
}

int swe_hip_mlp_log_block(swe_hip_mlp* m, swe_hip_block* blk, float keep, uint64_t seed, int hard_only,
                          int64_t* added) {
    if (!m || !blk || !(keep >= 0.f) || keep > 1.f) {
        set_error("swe_hip_mlp_log_block: invalid argument");
        return SWE_HIP_EINVAL;
    }
    StateView v;
    hipStream_t bs;
    if (int r = block_view(blk, &v, &bs)) return r;
    const long long before = count_of(m);
    if (before < 0) return hip_fail(hipErrorUnknown, "swe_hip_mlp_log_block: count");
    HIP_TRY(hipStreamSynchronize(bs));  // the block's queued steps have written the state
    const long long edges = (long long)(v.nx + 1) * v.ny + (long long)v.nx * (v.ny + 1);
    hipLaunchKernelGGL(mlp_log_kernel, dim3((unsigned)((edges + 255) / 256)), dim3(256), 0, m->stream, v, keep,
                       (unsigned long long)seed, hard_only ? 1 : 0, m->samples, m->cap, m->dCount);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    const long long after = count_of(m);
    if (added) *added = after - before;
    return SWE_HIP_OK;
}

int swe_hip_mlp_add_samples(swe_hip_mlp* m, int64_t n, const float* samples) {
    if (!m || n < 0 || (n > 0 && !samples)) {
        set_error("swe_hip_mlp_add_samples: invalid argument");
        return SWE_HIP_EINVAL;
    }
    const long long c = count_of(m);
    if (c < 0 || c + n > m->cap) {
        set_error("swe_hip_mlp_add_samples: %lld samples do not fit (%lld of %lld used)", (long long)n, c, m->cap);
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipMemcpy(m->samples + (size_t)c * kSampleFloats, samples, sizeof(float) * kSampleFloats * n,
                      hipMemcpyHostToDevice));
    const unsigned long long nc = (unsigned long long)(c + n);
    HIP_TRY(hipMemcpy(m->dCount, &nc, sizeof(nc), hipMemcpyHostToDevice));
    return SWE_HIP_OK;
}

int swe_hip_mlp_read_samples(swe_hip_mlp* m, int64_t first, int64_t n, float* samples) {
    const long long c = m ? count_of(m) : -1;
    if (!m || first < 0 || n < 0 || first + n > c || (n > 0 && !samples)) {
        set_error("swe_hip_mlp_read_samples: invalid range");
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipMemcpy(samples, m->samples + (size_t)first * kSampleFloats, sizeof(float) * kSampleFloats * n,
                      hipMemcpyDeviceToHost));
    return SWE_HIP_OK;
}

int swe_hip_mlp_sample_count(swe_hip_mlp* m, int64_t* count) {
    if (!m || !count) {
        set_error("swe_hip_mlp_sample_count: null argument");
        return SWE_HIP_EINVAL;
    }
    const long long c = count_of(m);
    if (c < 0) return hip_fail(hipErrorUnknown, "swe_hip_mlp_sample_count");
    *count = c;
    return SWE_HIP_OK;
}

int swe_hip_mlp_clear_samples(swe_hip_mlp* m) {
    if (!m) {
        set_error("swe_hip_mlp_clear_samples: null handle");
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipMemset(m->dCount, 0, sizeof(unsigned long long)));
    return SWE_HIP_OK;
}

int swe_hip_mlp_train(swe_hip_mlp* m, int steps, int batch, float lr, uint64_t seed, float* loss_log) {
    if (!m || steps < 0 || batch < 1 || batch > (1 << 20) || !(lr > 0.f)) {
        set_error("swe_hip_mlp_train: invalid argument");
        return SWE_HIP_EINVAL;
    }
    const long long n = count_of(m);
    if (n < 1) {
        set_error("swe_hip_mlp_train: no samples");
        return SWE_HIP_ESTATE;
    }
    if (int r = ensure_act(m, batch)) return r;
    if (int r = ensure_loss(m, std::max(steps, 1))) return r;
    HIP_TRY(hipMemsetAsync(m->lossLog, 0, sizeof(float) * std::max(steps, 1), m->stream));
    for (int s = 0; s < steps; ++s) {
        const float cosw = 0.5f * (1.f + std::cos(3.14159265f * (float)s / (float)std::max(steps, 1)));
        const float lrs = lr * (0.05f + 0.95f * cosw);
        if (int r = train_step(m, m->samples, n, batch, (unsigned long long)seed, s, 1, m->lossLog + s, lrs, nullptr))
            return r;
    }
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (loss_log && steps > 0)
        HIP_TRY(hipMemcpy(loss_log, m->lossLog, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return repack(m);
}

int swe_hip_mlp_gradient(swe_hip_mlp* m, int64_t n, const float* samples, float* grad, float* loss) {
    if (!m || n < 1 || n > (1 << 20) || !samples || !grad) {
        set_error("swe_hip_mlp_gradient: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (int r = ensure_act(m, (int)n)) return r;
    if (int r = ensure_loss(m, 1)) return r;
    float *ds = nullptr, *dg = nullptr;
    int rc = upload_tmp(&ds, samples, (size_t)n * kSampleFloats, m->stream);
    if (rc == 0 && hipMalloc(&dg, sizeof(float) * kParams) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "grad");
    if (rc == 0 && hipMemsetAsync(m->lossLog, 0, sizeof(float), m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "grad loss");
    if (rc == 0) rc = train_step(m, ds, n, (int)n, 0, 0, 0, m->lossLog, 0.f, dg);
    if (rc == 0 && hipMemcpyAsync(grad, dg, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "grad copy");
    if (rc == 0 && loss &&
        hipMemcpyAsync(loss, m->lossLog, sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "loss copy");
    if (rc == 0 && hipStreamSynchronize(m->stream) != hipSuccess) rc = hip_fail(hipErrorUnknown, "grad sync");
    if (ds) (void)hipFree(ds);
    if (dg) (void)hipFree(dg);
    return rc;
}

int swe_hip_mlp_infer(swe_hip_mlp* m, int64_t n, const float* features, float* out) {
    if (!m || n < 0 || (n > 0 && (!features || !out))) {
        set_error("swe_hip_mlp_infer: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (n == 0) return SWE_HIP_OK;
    float *df = nullptr, *dout = nullptr;
    int rc = upload_tmp(&df, features, (size_t)n * kIn, m->stream);
    if (rc == 0 && hipMalloc(&dout, sizeof(float) * 2 * n) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "infer");
    if (rc == 0) {
        const long long waves = (n + 63) / 64;
        hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, m->stream, m->frags, df,
                           (long long)n, dout);
        if (hipGetLastError() != hipSuccess) rc = hip_fail(hipErrorUnknown, "mlp_infer_kernel launch");
    }
    if (rc == 0 && hipMemcpyAsync(out, dout, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "infer copy");
    if (rc == 0 && hipStreamSynchronize(m->stream) != hipSuccess) rc = hip_fail(hipErrorUnknown, "infer sync");
    if (df) (void)hipFree(df);
    if (dout) (void)hipFree(dout);
    return rc;
}

int swe_hip_mlp_features(int64_t n, const float* hL, const float* hR, const float* uL, const float* uR,
                         const float* bL, const float* bR, float* features) {
    if (n < 0 || (n > 0 && (!hL || !hR || !uL || !uR || !bL || !bR || !features))) {
        set_error("swe_hip_mlp_features: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (n == 0) return SWE_HIP_OK;
    std::vector<float> in((size_t)6 * n);
    const float* src[6] = {hL, hR, uL, uR, bL, bR};
    for (int a = 0; a < 6; ++a) std::memcpy(in.data() + (size_t)a * n, src[a], sizeof(float) * n);
    float *din = nullptr, *dout = nullptr;
    int rc = upload_tmp(&din, in.data(), in.size(), nullptr);
    if (rc == 0 && hipMalloc(&dout, sizeof(float) * kIn * n) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "features");
    if (rc == 0) {
        hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, (long long)n,
                           din, din + n, din + 2 * n, din + 3 * n, din + 4 * n, din + 5 * n, dout);
        if (hipGetLastError() != hipSuccess) rc = hip_fail(hipErrorUnknown, "mlp_features_kernel launch");
    }
    if (rc == 0 && hipMemcpy(features, dout, sizeof(float) * kIn * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "features copy");
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return rc;
}

int swe_hip_mlp_middle(swe_hip_mlp* m, int64_t n, const float* hL, const float* hR, const float* uL, const float* uR,
                       const float* bL, const float* bR, float* hm, float* s1m, float* s2m, int32_t* accepted) {
    if (!m || n < 0 || (n > 0 && (!hL || !hR || !uL || !uR || !bL || !bR || !hm || !s1m || !s2m || !accepted))) {
        set_error("swe_hip_mlp_middle: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (n == 0) return SWE_HIP_OK;
    std::vector<float> in((size_t)6 * n);
    const float* src[6] = {hL, hR, uL, uR, bL, bR};
    for (int a = 0; a < 6; ++a) std::memcpy(in.data() + (size_t)a * n, src[a], sizeof(float) * n);
    float *din = nullptr, *dout = nullptr;
    int* dacc = nullptr;
    int rc = upload_tmp(&din, in.data(), in.size(), m->stream);
    if (rc == 0 && hipMalloc(&dout, sizeof(float) * 3 * n) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "middle");
    if (rc == 0 && hipMalloc(&dacc, sizeof(int) * n) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "middle");
    if (rc == 0) {
        const long long waves = (n + 63) / 64;
        hipLaunchKernelGGL(mlp_middle_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, m->stream, m->frags,
                           (long long)n, din, dout, dacc);
        if (hipGetLastError() != hipSuccess) rc = hip_fail(hipErrorUnknown, "mlp_middle_kernel launch");
    }
    std::vector<float> out((size_t)3 * n);
    if (rc == 0 && hipMemcpyAsync(out.data(), dout, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "middle copy");
    if (rc == 0 && hipMemcpyAsync(accepted, dacc, sizeof(int) * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "middle copy");
    if (rc == 0 && hipStreamSynchronize(m->stream) != hipSuccess) rc = hip_fail(hipErrorUnknown, "middle sync");
    if (rc == 0) {
        std::memcpy(hm, out.data(), sizeof(float) * n);
        std::memcpy(s1m, out.data() + n, sizeof(float) * n);
        std::memcpy(s2m, out.data() + 2 * n, sizeof(float) * n);
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (dacc) (void)hipFree(dacc);
    return rc;
}

int swe_hip_mlp_solve_edges(swe_hip_mlp* m, int64_t n, const float* hL, const float* hR, const float* huL,
                            const float* huR, const float* bL, const float* bR, float* hNetL, float* hNetR,
                            float* huNetL, float* huNetR, float* maxSpeed) {
    if (!m || n < 0 ||
        (n > 0 && (!hL || !hR || !huL || !huR || !bL || !bR || !hNetL || !hNetR || !huNetL || !huNetR || !maxSpeed))) {
        set_error("swe_hip_mlp_solve_edges: invalid argument");
        return SWE_HIP_EINVAL;
    }
    if (n == 0) return SWE_HIP_OK;
    std::vector<float> in((size_t)6 * n);
    const float* src[6] = {hL, hR, huL, huR, bL, bR};
    for (int a = 0; a < 6; ++a) std::memcpy(in.data() + (size_t)a * n, src[a], sizeof(float) * n);
    float *din = nullptr, *dout = nullptr;
    int rc = upload_tmp(&din, in.data(), in.size(), m->stream);
    if (rc == 0 && hipMalloc(&dout, sizeof(float) * 5 * n) != hipSuccess) rc = hip_fail(hipErrorOutOfMemory, "solve");
    if (rc == 0) {
        const long long waves = (n + 63) / 64;
        hipLaunchKernelGGL(mlp_solve_edges_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, m->stream, m->frags,
                           (long long)n, din, dout);
        if (hipGetLastError() != hipSuccess) rc = hip_fail(hipErrorUnknown, "mlp_solve_edges_kernel launch");
    }
    std::vector<float> out((size_t)5 * n);
    if (rc == 0 && hipMemcpyAsync(out.data(), dout, sizeof(float) * 5 * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        rc = hip_fail(hipErrorUnknown, "solve copy");
    if (rc == 0 && hipStreamSynchronize(m->stream) != hipSuccess) rc = hip_fail(hipErrorUnknown, "solve sync");
    if (rc == 0) {
        float* dst[5] = {hNetL, hNetR, huNetL, huNetR, maxSpeed};
        for (int a = 0; a < 5; ++a) std::memcpy(dst[a], out.data() + (size_t)a * n, sizeof(float) * n);
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
