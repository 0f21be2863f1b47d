// score.hip - FP64 completion of the per-scale norms and the final score.
//
// One workgroup per pair (grid-stride over pairs), one wave per active map job
// of the pair: the wave sums the per-tile float partials in double, divides by the
// scale's pixel count (L1 = mean, L4 = mean(x^4)^(1/4), reference
// src/ssimu2/score.hpp:315-320,422-438), then forms the weighted sum of the
// 108 features (one wave, fixed-order reduction) and the polynomial remap
// (score.hpp:442-460) in double. The reference accumulates the weighted sum in
// float and returns a float; this kernel keeps every step in double.
// Summation order is fixed, so results are bit-reproducible run to run.
#include "ssimu2_common.hpp"
#include "weights.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace vship {

namespace {

__constant__ double d_weights[108] = VSHIP_SSIMU2_WEIGHTS;

constexpr int kScoreThreads = 1024;  // 16 waves: one wave per map job of a pair

__device__ inline double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kScoreThreads)
void score_kernel(const ScoreJob* __restrict__ pairs, const int* __restrict__ counts,
                  const MapJob* __restrict__ jobs, const float* __restrict__ partials,
                  double* __restrict__ features, double* __restrict__ scores)
{
    __shared__ double feat[108];
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int npairs = counts[kCountPairs];
    for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const ScoreJob pj = pairs[pi];
        for (int i = t; i < 108; i += kScoreThreads) feat[i] = 0.0;
        __syncthreads();
        for (int j = wave; j < pj.job_count; j += kScoreThreads / 64) {
            const MapJob job = jobs[pj.job_begin + j];
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int k = lane; k < job.ntiles; k += 64) {
                const float* p = partials + (int64_t)(job.part_start + k) * kPartialStride;
#pragma unroll
                for (int m = 0; m < 6; m++) acc[m] += (double)p[m];
            }
#pragma unroll
            for (int m = 0; m < 6; m++) acc[m] = wave_sum_d(acc[m]);
            if (lane < 3) {
                const int m = lane;
                const int flag = (m == 0) ? kMapSsim : (m == 1 ? kMapArtifact : kMapDetail);
                if (job.flags & flag) {
                    const double n = (double)job.w * (double)job.h;
                    const double s1 = m == 0 ? acc[0] : (m == 1 ? acc[1] : acc[2]);
                    const double s4 = m == 0 ? acc[3] : (m == 1 ? acc[4] : acc[5]);
                    feat[job.feature_base + m] = s1 / n;
                    feat[job.feature_base + 3 + m] = sqrt(sqrt(s4 / n));
                }
            }
        }
        __syncthreads();
        if (wave == 0) {
            // weighted sum of the 108 features: lane k takes features k and
            // k + 64, then a fixed-order wave reduction (all in double)
            double s = lane < 108 ? d_weights[lane] * feat[lane] : 0.0;
            if (lane + 64 < 108) s = fma(d_weights[lane + 64], feat[lane + 64], s);
            s = wave_sum_d(s);
            if (lane == 0) {
                // no contraction: the single-pair handler forms the same
                // polynomial on the host with these operations
#pragma clang fp contract(off)
                s *= 0.9562382616834844;
                s = (6.248496625763138e-5 * s * s) * s + 2.326765642916932 * s - 0.020884521182843837 * s * s;
                scores[pj.pair] = (s > 0.0) ? pow(s, 0.6276336467831387) * -10.0 + 100.0 : 100.0;
            }
        }
        if (features)
            for (int i = t; i < 108; i += kScoreThreads) features[(int64_t)pj.pair * 108 + i] = feat[i];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t stream)
{
    const int64_t n = bytes / 16;
    if (n <= 0) return hipSuccess;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, stream, static_cast<float4*>(dst),
                       static_cast<const float4*>(src), n);
    return hipGetLastError();
}

hipError_t launch_score(const ScoreJob* pairs, const int* counts, int grid, const MapJob* jobs,
                        const float* partials, double* features, double* scores, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(score_kernel, dim3(grid), dim3(kScoreThreads), 0, stream, pairs, counts, jobs, partials,
                       features, scores);
    return hipGetLastError();
}

}  // namespace vship
