// resize.hip - separable cubic-Hermite resize of linear-RGB planes for the
// target_width / target_height path, reference
// src/gpuColorToLinear/resize.hpp:26-65.
//
// Output sample x of a line of n source samples resized to r samples sits at
// pos = x * n / r (float arithmetic as in the reference, no half-pixel
// offset); with b = floor(pos) and t = pos - b it is the Catmull-Rom segment
// over (s[b-1], s[b], s[b+1], s[b+2]) where a missing left neighbour is s[b],
// a missing s[b+1] is s[b] and a missing s[b+2] is s[b+1]. The horizontal pass
// runs first, then the vertical one (the reference transposes twice; the
// arithmetic per sample is the same).
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

__device__ inline float resample(const float* line, int64_t step, int n, int r, int x)
{
    const float pos = (float)x * (float)n / (float)r;
    const int b = (int)pos;
    const float v1 = line[(int64_t)b * step];
    const float v0 = (b - 1 >= 0) ? line[(int64_t)(b - 1) * step] : v1;
    const float v2 = (b + 1 < n) ? line[(int64_t)(b + 1) * step] : v1;
    const float v3 = (b + 2 < n) ? line[(int64_t)(b + 2) * step] : v2;
    return catmull_rom(v0, v1, v2, v3, pos - (float)b);
}

// grid: (ceil(rw*h/256), 3 planes)
__global__ __launch_bounds__(256) void resize_h_kernel(float* __restrict__ dst, const float* __restrict__ src, int sw,
                                                       int h, int rw)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)rw * h) return;
    const int y = (int)(i / rw), x = (int)(i % rw);
    const int p = blockIdx.y;
    const float* line = src + (int64_t)p * sw * h + (int64_t)y * sw;
    dst[(int64_t)p * rw * h + i] = resample(line, 1, sw, rw, x);
}

// grid: (ceil(w*rh/256), 3 planes)
__global__ __launch_bounds__(256) void resize_v_kernel(float* __restrict__ dst, const float* __restrict__ src, int w,
                                                       int sh, int rh)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)w * rh) return;
    const int y = (int)(i / w), x = (int)(i % w);
    const int p = blockIdx.y;
    const float* col = src + (int64_t)p * w * sh + x;
    dst[(int64_t)p * w * rh + i] = resample(col, w, sh, rh, y);
}

}  // namespace

hipError_t launch_resize_h(float* dst, const float* src, int sw, int h, int rw, hipStream_t stream)
{
    const int64_t n = (int64_t)rw * h;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(resize_h_kernel, dim3((unsigned)((n + 255) / 256), 3), dim3(256), 0, stream, dst, src, sw, h,
                       rw);
    return hipGetLastError();
}

hipError_t launch_resize_v(float* dst, const float* src, int w, int sh, int rh, hipStream_t stream)
{
    const int64_t n = (int64_t)w * rh;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(resize_v_kernel, dim3((unsigned)((n + 255) / 256), 3), dim3(256), 0, stream, dst, src, w, sh,
                       rh);
    return hipGetLastError();
}

}  // namespace vship
