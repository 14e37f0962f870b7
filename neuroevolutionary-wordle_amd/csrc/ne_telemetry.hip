// ne_telemetry.hip -- device side of the generation telemetry that drives the curriculum controller:
// fitness order statistics (one workgroup sorts the generation's fitness in LDS) and the sampled
// genetic-convergence distances.  Sums that the reference accumulates sequentially run sequentially
// here too, in the same order and precision, so every figure is the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ne_kernels.hpp"

namespace ne {

// GenerationFitnessTelemetryRow's inputs (src/cli/run_genetic_algorithm.cu:1080-1186): the fitness
// sorted ascending (bitonic network over the next power of two, +inf padding), the order statistics the
// host asks for, the distinct-value count, the sequential double sums over the sorted values, and, when
// reference counts are given, the parents with 0, 1 and >= 2 children.
__global__ __launch_bounds__(1024) void FitnessStatsKernel(const float* fit, const uint8_t* has, uint32_t n,
                                                          uint32_t pow2, uint32_t i90, uint32_t i99,
                                                          const uint32_t* refcounts, FitnessStatsRecord* out) {
    extern __shared__ float s_sorted[];
    __shared__ uint32_t s_count[5];
    const uint32_t tid = threadIdx.x;
    if (tid < 5) s_count[tid] = 0;
    for (uint32_t i = tid; i < pow2; i += blockDim.x) s_sorted[i] = i < n ? fit[i] : __builtin_inff();
    uint32_t missing = 0, zero = 0, one = 0, many = 0;
    for (uint32_t i = tid; i < n; i += blockDim.x) {
        missing += has[i] == 0 ? 1u : 0u;
        if (refcounts) {
            const uint32_t r = refcounts[i];
            zero += r == 0 ? 1u : 0u;
            one += r == 1 ? 1u : 0u;
            many += r >= 2 ? 1u : 0u;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= pow2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < pow2; i += blockDim.x) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const float a = s_sorted[i], b = s_sorted[p];
                    if ((i & k) == 0 ? a > b : a < b) {
                        s_sorted[i] = b;
                        s_sorted[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    uint32_t distinct = 0;
    for (uint32_t i = tid + 1; i < n; i += blockDim.x) distinct += s_sorted[i] != s_sorted[i - 1] ? 1u : 0u;
    atomicAdd(&s_count[0], distinct);
    atomicAdd(&s_count[1], missing);
    atomicAdd(&s_count[2], zero);
    atomicAdd(&s_count[3], one);
    atomicAdd(&s_count[4], many);
    __syncthreads();
    if (tid == 0) {
        // the reference's accumulation order: ascending values, one double at a time
        double sum = 0.0;
        for (uint32_t i = 0; i < n; ++i) sum += double(s_sorted[i]);
        const double mean = sum / double(n);
        double squares = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double d = double(s_sorted[i]) - mean;
            squares += d * d;
        }
        FitnessStatsRecord r{};
        r.sum = sum;
        r.squared_deltas = squares;
        r.min = s_sorted[0];
        r.max = s_sorted[n - 1];
        r.median_low = s_sorted[n >= 2 ? n / 2 - 1 : 0];
        r.median_high = s_sorted[n / 2];
        r.p90 = s_sorted[i90];
        r.p99 = s_sorted[i99];
        r.distinct = s_count[0] + 1;
        r.missing = s_count[1];
        r.childless = s_count[2];
        r.one_child = s_count[3];
        r.multiple_children = s_count[4];
        *out = r;
    }
}

// GatherGeneticConvergenceSamplesKernel (run_genetic_algorithm.cu:887-935) on fp32 slots: value w of
// sampled organism o is trainable scalar weights[w] of its genome (policy parameters, then tail rows), 0
// past the genome's last scalar.  Written twice: [o][w] for the centroid pass, [w][o] for the distances.
__global__ void ConvergenceGatherKernel(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n_org,
                                        const uint32_t* weights, uint32_t n_w, uint64_t scalars, float* by_org,
                                        float* by_weight) {
    const uint64_t flat = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (flat >= uint64_t(n_org) * n_w) return;
    const uint32_t o = uint32_t(flat / n_w), w = uint32_t(flat % n_w);
    const float* g = reinterpret_cast<const float*>(slab + uint64_t(slots[o]) * stride);
    const uint32_t idx = weights[w];
    const float v = idx < scalars ? g[idx] : 0.0f;
    by_org[flat] = v;
    by_weight[uint64_t(w) * n_org + o] = v;
}

// TryComputeMetrics' centroid (genetic_convergence.hpp:130-144): per weight, the organisms' values added
// in organism order in float, then scaled by the float reciprocal of the organism count.
__global__ void ConvergenceCentroidKernel(const float* by_org, uint32_t n_org, uint32_t n_w, float inverse,
                                          float* centroid) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_w) return;
    float c = 0.0f;
    for (uint32_t o = 0; o < n_org; ++o) c += by_org[uint64_t(o) * n_w + w];
    centroid[w] = c * inverse;
}

// The squared distances of TryComputeMetrics (genetic_convergence.hpp:146-191), each summed over the
// weights in weight order in float: thread t < n_org holds organism t against the centroid, thread
// n_org + p holds pair p.  The square roots and the means are taken on the host.
__global__ void ConvergenceDistanceKernel(const float* by_weight, const float* centroid, uint32_t n_org, uint32_t n_w,
                                          const uint32_t* pairs, uint32_t n_pairs, float* squared) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_org + n_pairs) return;
    float s = 0.0f;
    if (t < n_org) {
        for (uint32_t w = 0; w < n_w; ++w) {
            const float d = by_weight[uint64_t(w) * n_org + t] - centroid[w];
            s += d * d;
        }
    } else {
        const uint32_t a = pairs[2 * (t - n_org)], b = pairs[2 * (t - n_org) + 1];
        for (uint32_t w = 0; w < n_w; ++w) {
            const float d = by_weight[uint64_t(w) * n_org + a] - by_weight[uint64_t(w) * n_org + b];
            s += d * d;
        }
    }
    squared[t] = s;
}

void LaunchFitnessStats(const float* fit, const uint8_t* has, uint32_t n, uint32_t pow2, uint32_t i90, uint32_t i99,
                        const uint32_t* refcounts, FitnessStatsRecord* out, hipStream_t s) {
    FitnessStatsKernel<<<1, 1024, pow2 * sizeof(float), s>>>(fit, has, n, pow2, i90, i99, refcounts, out);
}

void LaunchConvergence(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n_org,
                       const uint32_t* weights, uint32_t n_w, uint64_t scalars, const uint32_t* pairs, uint32_t n_pairs,
                       float inverse, float* by_org, float* by_weight, float* centroid, float* squared,
                       hipStream_t s) {
    const uint64_t values = uint64_t(n_org) * n_w;
    ConvergenceGatherKernel<<<uint32_t((values + 255) / 256), 256, 0, s>>>(slab, stride, slots, n_org, weights, n_w,
                                                                          scalars, by_org, by_weight);
    ConvergenceCentroidKernel<<<(n_w + 255) / 256, 256, 0, s>>>(by_org, n_org, n_w, inverse, centroid);
    ConvergenceDistanceKernel<<<(n_org + n_pairs + 63) / 64, 64, 0, s>>>(by_weight, centroid, n_org, n_w, pairs,
                                                                        n_pairs, squared);
}

}  // namespace ne
