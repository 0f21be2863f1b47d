// upload.hip - input planes in pinned host memory copied to the device by one
// kernel (the handler's host-plane call, reference VshipLib.cpp:247-268 with
// the copies of vshipColor.hpp:96): one launch reads all six planes of a pair
// over PCIe instead of six stream-ordered copies, whose per-copy overhead
// (tens of microseconds each) and contention between host threads bound the
// call's rate.
#include "kernels.hpp"

namespace vship {

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kGatherThreads = 256;
constexpr int kGatherBlocksPerPlane = 16;

// blockIdx.y picks the plane; 16-byte vectors, then the plane's last
// bytes % 16 one per thread.
__global__ __launch_bounds__(kGatherThreads) void gather_kernel(GatherPlanes g)
{
    const int s = blockIdx.y;
    if (s >= g.n) return;
    const int64_t nv = g.bytes[s] / 16;
    const auto* src = (const __attribute__((address_space(1))) u32x4*)g.src[s];
    auto* dst = (__attribute__((address_space(1))) u32x4*)g.dst[s];
    // four loads in flight per thread (PCIe latency), then their stores
    const int64_t stride = (int64_t)gridDim.x * kGatherThreads;
    for (int64_t i = (int64_t)blockIdx.x * kGatherThreads + threadIdx.x; i < nv; i += 4 * stride) {
        u32x4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + k * stride < nv) v[k] = src[i + k * stride];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i + k * stride < nv) dst[i + k * stride] = v[k];
    }
    const int64_t tail = g.bytes[s] - nv * 16;
    if (blockIdx.x == 0 && threadIdx.x < tail) {
        const auto* sb = (const __attribute__((address_space(1))) uint8_t*)g.src[s];
        auto* db = (__attribute__((address_space(1))) uint8_t*)g.dst[s];
        db[nv * 16 + threadIdx.x] = sb[nv * 16 + threadIdx.x];
    }
}

}  // namespace

hipError_t launch_gather(const GatherPlanes& g, hipStream_t stream)
{
    if (g.n <= 0) return hipSuccess;
    if (g.n > kGatherPlanes) return hipErrorInvalidValue;
    for (int s = 0; s < g.n; s++)
        if (((uintptr_t)g.src[s] | (uintptr_t)g.dst[s]) % 16 != 0 || g.bytes[s] < 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_kernel, dim3(kGatherBlocksPerPlane, g.n), dim3(kGatherThreads), 0, stream, g);
    return hipGetLastError();
}

}  // namespace vship
