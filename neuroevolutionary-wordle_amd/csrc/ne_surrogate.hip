// ne_surrogate.hip -- the learned fitness surrogate of screened evolution: a 192 -> 128 -> 64 -> 1
// ReLU MLP over child fingerprints, its forward pass, backpropagation and Adam, and the replay
// buffer of fully evaluated (fingerprint, fitness) pairs it trains on.  Every sum runs in a fixed
// order (no atomics), so training is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ne_device.hpp"
#include "ne_kernels.hpp"

namespace ne {

constexpr uint32_t kTagSurr = 0x53555247u;
constexpr uint32_t kSgRows = 8;  // samples per forward workgroup

// He initialisation (N(0, 2 / fan_in) for the ReLU layers, N(0, 1 / 64) for the output), biases 0,
// Adam moments 0.  Draw of parameter i: Philox(i; seed, tag).
__global__ void SurrogateInitKernel(float* th, float* m, float* v, uint32_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kSgParams) return;
    float std = 0.0f;
    if (i < kSgOffB1)
        std = 0.10206207261596575f;  // sqrt(2 / 192)
    else if (i >= kSgOffW2 && i < kSgOffB2)
        std = 0.125f;  // sqrt(2 / 128)
    else if (i >= kSgOffW3 && i < kSgOffB3)
        std = 0.125f;  // sqrt(1 / 64)
    const U4 w = Philox4x32(i, 0, 0, 0, seed, kTagSurr);
    th[i] = std == 0.0f ? 0.0f : std * NormalFromWords(w.x, w.y);
    m[i] = 0.0f;
    v[i] = 0.0f;
}

// Forward of kSgRows samples per workgroup of 128 threads: thread h computes hidden unit h of layer 1
// for all rows (W1T row i is read once per workgroup, coalesced over h), then layer 2 (unit t % 64,
// rows 4 (t / 64) ..), then 8 threads the outputs.  Sums run in increasing input order.
__global__ __launch_bounds__(128) void SurrogateForwardKernel(const float* th, const float* x, uint32_t n, float* y,
                                                              float* h1o, float* h2o) {
    __shared__ float xs[kSgRows][kSgIn];
    __shared__ float h1s[kSgRows][kSgH1];
    __shared__ float h2s[kSgRows][kSgH2];
    const uint32_t m0 = blockIdx.x * kSgRows, tid = threadIdx.x;
    for (uint32_t i = tid; i < kSgRows * kSgIn; i += blockDim.x) {
        const uint32_t r = i / kSgIn, c = i - r * kSgIn;
        xs[r][c] = m0 + r < n ? x[uint64_t(m0 + r) * kSgIn + c] : 0.0f;
    }
    __syncthreads();
    {
        const uint32_t h = tid;
        float acc[kSgRows];
#pragma unroll
        for (uint32_t r = 0; r < kSgRows; ++r) acc[r] = th[kSgOffB1 + h];
#pragma unroll 8
        for (uint32_t i = 0; i < kSgIn; ++i) {
            const float w = th[kSgOffW1 + i * kSgH1 + h];
#pragma unroll
            for (uint32_t r = 0; r < kSgRows; ++r) acc[r] = fmaf(w, xs[r][i], acc[r]);
        }
#pragma unroll
        for (uint32_t r = 0; r < kSgRows; ++r) {
            const float v = fmaxf(acc[r], 0.0f);
            h1s[r][h] = v;
            if (h1o != nullptr && m0 + r < n) h1o[uint64_t(m0 + r) * kSgH1 + h] = v;
        }
    }
    __syncthreads();
    {
        constexpr uint32_t kHalf = kSgRows / 2;
        const uint32_t o = tid % kSgH2, rb = (tid / kSgH2) * kHalf;
        float acc[kHalf];
#pragma unroll
        for (uint32_t r = 0; r < kHalf; ++r) acc[r] = th[kSgOffB2 + o];
#pragma unroll 8
        for (uint32_t h = 0; h < kSgH1; ++h) {
            const float w = th[kSgOffW2 + h * kSgH2 + o];
#pragma unroll
            for (uint32_t r = 0; r < kHalf; ++r) acc[r] = fmaf(w, h1s[rb + r][h], acc[r]);
        }
#pragma unroll
        for (uint32_t r = 0; r < kHalf; ++r) {
            const float v = fmaxf(acc[r], 0.0f);
            h2s[rb + r][o] = v;
            if (h2o != nullptr && m0 + rb + r < n) h2o[uint64_t(m0 + rb + r) * kSgH2 + o] = v;
        }
    }
    __syncthreads();
    if (tid < kSgRows && m0 + tid < n) {
        float acc = th[kSgOffB3];
        for (uint32_t o = 0; o < kSgH2; ++o) acc = fmaf(th[kSgOffW3 + o], h2s[tid][o], acc);
        y[m0 + tid] = acc;
    }
}

// One sample per workgroup of 128 threads.  With L = 1/(2n) sum (y - t)^2: dy = (y - t) / n,
// dh2 = dy w3 [h2 > 0], dh1 = (W2 dh2) [h1 > 0] (ReLU derivative 0 at 0).
__global__ __launch_bounds__(128) void SurrogateBackpropKernel(const float* th, const float* h1, const float* h2,
                                                               const float* y, const float* t, uint32_t n, float* dy,
                                                               float* dh1, float* dh2) {
    __shared__ float s_dh2[kSgH2];
    const uint32_t m = blockIdx.x, tid = threadIdx.x;
    const float g = (y[m] - t[m]) / float(n);
    if (tid == 0) dy[m] = g;
    if (tid < kSgH2) {
        const float v = h2[uint64_t(m) * kSgH2 + tid] > 0.0f ? g * th[kSgOffW3 + tid] : 0.0f;
        s_dh2[tid] = v;
        dh2[uint64_t(m) * kSgH2 + tid] = v;
    }
    __syncthreads();
    float acc = 0.0f;
    for (uint32_t o = 0; o < kSgH2; ++o) acc = fmaf(th[kSgOffW2 + tid * kSgH2 + o], s_dh2[o], acc);
    dh1[uint64_t(m) * kSgH1 + tid] = h1[uint64_t(m) * kSgH1 + tid] > 0.0f ? acc : 0.0f;
}

// One thread per parameter: its gradient summed over the batch in sample order (loads issued 8 samples
// ahead of the accumulation), then Adam.
__global__ __launch_bounds__(256) void SurrogateAdamKernel(float* th, float* am, float* av, const float* x,
                                                           const float* h1, const float* h2, const float* dy,
                                                           const float* dh1, const float* dh2, uint32_t n, float lr,
                                                           const float* coef) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kSgParams) return;
    const float c1 = coef[0], c2 = coef[1];
    float g = 0.0f;
    if (p < kSgOffB1) {
        const uint32_t i = p / kSgH1, h = p - i * kSgH1;
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g = fmaf(x[uint64_t(m) * kSgIn + i], dh1[uint64_t(m) * kSgH1 + h], g);
    } else if (p < kSgOffW2) {
        const uint32_t h = p - kSgOffB1;
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g += dh1[uint64_t(m) * kSgH1 + h];
    } else if (p < kSgOffB2) {
        const uint32_t q = p - kSgOffW2, h = q / kSgH2, o = q - h * kSgH2;
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g = fmaf(h1[uint64_t(m) * kSgH1 + h], dh2[uint64_t(m) * kSgH2 + o], g);
    } else if (p < kSgOffW3) {
        const uint32_t o = p - kSgOffB2;
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g += dh2[uint64_t(m) * kSgH2 + o];
    } else if (p < kSgOffB3) {
        const uint32_t o = p - kSgOffW3;
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g = fmaf(h2[uint64_t(m) * kSgH2 + o], dy[m], g);
    } else {
#pragma unroll 8
        for (uint32_t m = 0; m < n; ++m) g += dy[m];
    }
    const float m1 = 0.9f * am[p] + 0.1f * g;
    const float v1 = 0.999f * av[p] + 0.001f * (g * g);
    am[p] = m1;
    av[p] = v1;
    th[p] -= lr * (m1 * c1) / (sqrtf(v1 * c2) + 1e-8f);
}

__global__ void ReplayAppendKernel(const float* fp, const uint32_t* src_row, const float* fitness, uint32_t count,
                                   float* rx, float* rf, uint32_t head, uint32_t cap) {
    const uint32_t e = blockIdx.x;
    if (e >= count) return;
    const uint32_t dst = (head + e) % cap;
    for (uint32_t c = threadIdx.x; c < kSgIn; c += blockDim.x)
        rx[uint64_t(dst) * kSgIn + c] = fp[uint64_t(src_row[e]) * kSgIn + c];
    if (threadIdx.x == 0) rf[dst] = fitness[e];
}

__global__ void ReplayGatherKernel(const float* rx, const float* rf, const uint32_t* idx, uint32_t n,
                                   const float* stats, float* x, float* t) {
    const uint32_t b = blockIdx.x;
    if (b >= n) return;
    const uint32_t r = idx[b];
    for (uint32_t c = threadIdx.x; c < kSgIn; c += blockDim.x) x[uint64_t(b) * kSgIn + c] = rx[uint64_t(r) * kSgIn + c];
    if (threadIdx.x == 0) t[b] = (rf[r] - stats[0]) * stats[1];
}

void LaunchSurrogateInit(float* params, float* adam_m, float* adam_v, uint32_t seed, hipStream_t s) {
    hipLaunchKernelGGL(SurrogateInitKernel, dim3((kSgParams + 255) / 256), dim3(256), 0, s, params, adam_m, adam_v,
                       seed);
}

void LaunchSurrogateForward(const float* params, const float* x, uint32_t n, float* y, float* h1, float* h2,
                            hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(SurrogateForwardKernel, dim3((n + kSgRows - 1) / kSgRows), dim3(128), 0, s, params, x, n, y,
                       h1, h2);
}

void LaunchSurrogateBackprop(const float* params, const float* h1, const float* h2, const float* y, const float* t,
                             uint32_t n, float* dy, float* dh1, float* dh2, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(SurrogateBackpropKernel, dim3(n), dim3(128), 0, s, params, h1, h2, y, t, n, dy, dh1, dh2);
}

void LaunchSurrogateAdam(float* params, float* adam_m, float* adam_v, const float* x, const float* h1, const float* h2,
                         const float* dy, const float* dh1, const float* dh2, uint32_t n, float lr, const float* coef,
                         hipStream_t s) {
    hipLaunchKernelGGL(SurrogateAdamKernel, dim3((kSgParams + 255) / 256), dim3(256), 0, s, params, adam_m, adam_v,
                       x, h1, h2, dy, dh1, dh2, n, lr, coef);
}

void LaunchReplayAppend(const float* fp, const uint32_t* src_row, const float* fitness, uint32_t count, float* rx,
                        float* rf, uint32_t head, uint32_t cap, hipStream_t s) {
    if (count == 0) return;
    hipLaunchKernelGGL(ReplayAppendKernel, dim3(count), dim3(64), 0, s, fp, src_row, fitness, count, rx, rf, head, cap);
}

void LaunchReplayGather(const float* rx, const float* rf, const uint32_t* idx, uint32_t n, const float* stats,
                        float* x, float* t, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(ReplayGatherKernel, dim3(n), dim3(64), 0, s, rx, rf, idx, n, stats, x, t);
}

}  // namespace ne
