// score.hip - FP64 completion of the per-scale norms and the final score.
//
// One workgroup per pair (grid-stride over pairs): for every active map job
// of the pair it sums the per-tile float partials in double, divides by the
// scale's pixel count (L1 = mean, L4 = mean(x^4)^(1/4), reference
// src/ssimu2/score.hpp:315-320,422-438), then forms the weighted sum of the
// 108 features and the polynomial remap (score.hpp:442-460) in double. The
// reference accumulates the weighted sum in float and returns a float; this
// kernel keeps every step in double. Summation order is fixed, so results are
// bit-reproducible run to run.
#include "ssimu2_common.hpp"
#include "weights.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

__constant__ double d_weights[108] = VSHIP_SSIMU2_WEIGHTS;

constexpr int kScoreThreads = 256;

__device__ inline double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(kScoreThreads)
void score_kernel(const ScoreJob* __restrict__ pairs, const int* __restrict__ npairs_ptr,
                  const MapJob* __restrict__ jobs, const float* __restrict__ partials,
                  double* __restrict__ features, double* __restrict__ scores)
{
    __shared__ double red[4][6];
    __shared__ double feat[108];
    const int t = threadIdx.x;
    const int npairs = *npairs_ptr;
    for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const ScoreJob pj = pairs[pi];
        for (int i = t; i < 108; i += kScoreThreads) feat[i] = 0.0;
        __syncthreads();
        for (int j = pj.job_begin; j < pj.job_begin + pj.job_count; j++) {
            const MapJob job = jobs[j];
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int k = t; k < job.ntiles; k += kScoreThreads) {
                const float* p = partials + (int64_t)(job.tile_start + k) * kPartialStride;
#pragma unroll
                for (int m = 0; m < 6; m++) acc[m] += (double)p[m];
            }
#pragma unroll
            for (int m = 0; m < 6; m++) acc[m] = wave_sum_d(acc[m]);
            if ((t & 63) == 0) {
#pragma unroll
                for (int m = 0; m < 6; m++) red[t >> 6][m] = acc[m];
            }
            __syncthreads();
            if (t < 3) {
                const int m = t;
                const int flag = (m == 0) ? kMapSsim : (m == 1 ? kMapArtifact : kMapDetail);
                if (job.flags & flag) {
                    const double n = (double)job.w * (double)job.h;
                    const double s1 = ((red[0][m] + red[1][m]) + red[2][m]) + red[3][m];
                    const double s4 = ((red[0][3 + m] + red[1][3 + m]) + red[2][3 + m]) + red[3][3 + m];
                    feat[job.feature_base + m] = s1 / n;
                    feat[job.feature_base + 3 + m] = sqrt(sqrt(s4 / n));
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            double s = 0.0;
            for (int i = 0; i < 108; i++) s = fma(d_weights[i], feat[i], s);
            s *= 0.9562382616834844;
            s = (6.248496625763138e-5 * s * s) * s + 2.326765642916932 * s - 0.020884521182843837 * s * s;
            scores[pi] = (s > 0.0) ? pow(s, 0.6276336467831387) * -10.0 + 100.0 : 100.0;
        }
        if (features)
            for (int i = t; i < 108; i += kScoreThreads) features[(int64_t)pi * 108 + i] = feat[i];
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_score(const ScoreJob* pairs, const int* npairs, int grid, const MapJob* jobs,
                        const float* partials, double* features, double* scores, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(score_kernel, dim3(grid), dim3(kScoreThreads), 0, stream, pairs, npairs, jobs, partials,
                       features, scores);
    return hipGetLastError();
}

}  // namespace vship
