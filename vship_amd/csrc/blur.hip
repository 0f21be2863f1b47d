// blur.hip - fused 17-tap blur + SSIM/edge maps + per-tile partial sums.
//
// One workgroup (256 threads) produces a 64x32 output tile of one active
// (pair, plane, scale). It stages the 80x48 input window of both images in
// LDS (zero outside the image, reference gaussianblur.hpp:24-36), runs the
// horizontal 17-tap pass register-blocked (each thread: 4 adjacent columns
// of 3 rows, 20-sample windows read as 128-bit LDS loads), writes the
// horizontal results over the staging buffer, runs the vertical pass (each
// thread: one column, 8 rows), and evaluates the maps in registers. Each axis
// is divided by the sum of its in-bounds taps (gaussianblur.hpp:52-54,68-70).
// The blurred quantities are a, b, a*b and a*a + b*b (4 blurs when the SSIM
// map is active, otherwise a and b only). Sums of x and x^4 of the three maps
// are reduced across the workgroup and written as one 8-float partial per
// tile; the FP64 score kernel finishes the norms.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "blur_maps.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

constexpr int kSW = kBlurTW + 2 * kHalo;   // 80 staged columns
constexpr int kSH = kBlurTH + 2 * kHalo;   // 48 staged rows
constexpr int kHRowsPerThread = kSH / 16;  // 3 rows of the horizontal pass
constexpr int kVRows = kBlurTH / 4;        // 8 output rows per thread

__device__ inline float inv_norm(int x, int w, const BlurTaps& taps)
{
    if (x >= w) return 0.0f;
    const int beg = x < 8 ? 8 - x : 0;
    const int end = (w - x + 8) < 17 ? (w - x + 8) : 17;
    return 1.0f / (taps.integral[end] - taps.integral[beg]);
}

template <typename StoreT>
__device__ inline void load4(const StoreT* p, float v[4]);

template <>
__device__ inline void load4<float>(const float* p, float v[4])
{
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

template <>
__device__ inline void load4<__half>(const __half* p, float v[4])
{
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    v[0] = __half2float(__ushort_as_half((unsigned short)(q.x & 0xffff)));
    v[1] = __half2float(__ushort_as_half((unsigned short)(q.x >> 16)));
    v[2] = __half2float(__ushort_as_half((unsigned short)(q.y & 0xffff)));
    v[3] = __half2float(__ushort_as_half((unsigned short)(q.y >> 16)));
}

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename StoreT, int NQ>
__device__ void blur_tile(const MapJob& job, int tile_x, int tile_y, const BlurTaps& taps,
                          float* lds, float* red, float* __restrict__ partial)
{
    const int t = threadIdx.x;
    const int w = job.w, h = job.h, pitch = job.pitch;
    const int tx0 = tile_x * kBlurTW, ty0 = tile_y * kBlurTH;
    const StoreT* A = static_cast<const StoreT*>(job.a);
    const StoreT* B = static_cast<const StoreT*>(job.b);
    float* sA = lds;
    float* sB = lds + kSH * kSW;

    // ---- stage the (kSH x kSW) input windows, zero outside the image
    for (int e = t; e < kSH * (kSW / 4); e += kBlurThreads) {
        const int r = e / (kSW / 4);
        const int c4 = e - r * (kSW / 4);
        const int gy = ty0 - kHalo + r;
        const int gx = tx0 - kHalo + 4 * c4;
        float va[4] = {0.f, 0.f, 0.f, 0.f}, vb[4] = {0.f, 0.f, 0.f, 0.f};
        if (gy >= 0 && gy < h && gx >= 0 && gx < pitch) {
            const int64_t o = (int64_t)gy * pitch + gx;
            load4<StoreT>(A + o, va);
            load4<StoreT>(B + o, vb);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (gx + i >= w) { va[i] = 0.f; vb[i] = 0.f; }
        }
        *reinterpret_cast<float4*>(sA + r * kSW + 4 * c4) = make_float4(va[0], va[1], va[2], va[3]);
        *reinterpret_cast<float4*>(sB + r * kSW + 4 * c4) = make_float4(vb[0], vb[1], vb[2], vb[3]);
    }
    __syncthreads();

    // ---- centre samples for the edge maps (vertical-pass ownership)
    const int vc = t & 63;
    const int vr0 = (t >> 6) * kVRows;
    float ca[kVRows], cb[kVRows];
#pragma unroll
    for (int j = 0; j < kVRows; j++) {
        ca[j] = sA[(kHalo + vr0 + j) * kSW + kHalo + vc];
        cb[j] = sB[(kHalo + vr0 + j) * kSW + kHalo + vc];
    }

    // ---- horizontal pass: rows hr = (t>>4) + 16*it, columns 4cg..4cg+3
    const int cg = t & 15;
    float inx[4];
#pragma unroll
    for (int i = 0; i < 4; i++) inx[i] = inv_norm(tx0 + 4 * cg + i, w, taps);
    float hres[kHRowsPerThread][NQ][4];
#pragma unroll
    for (int it = 0; it < kHRowsPerThread; it++) {
        const int hr = (t >> 4) + 16 * it;
        float wa[20], wb[20];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const float4 qa = *reinterpret_cast<const float4*>(sA + hr * kSW + 4 * cg + 4 * k);
            const float4 qb = *reinterpret_cast<const float4*>(sB + hr * kSW + 4 * cg + 4 * k);
            wa[4 * k] = qa.x; wa[4 * k + 1] = qa.y; wa[4 * k + 2] = qa.z; wa[4 * k + 3] = qa.w;
            wb[4 * k] = qb.x; wb[4 * k + 1] = qb.y; wb[4 * k + 2] = qb.z; wb[4 * k + 3] = qb.w;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = 0.0f;
#pragma unroll
            for (int k = 0; k < 17; k++) {
                const float g = taps.g[k];
                const float a = wa[i + k], b = wb[i + k];
                acc[0] = fmaf(g, a, acc[0]);
                acc[1] = fmaf(g, b, acc[1]);
                if constexpr (NQ == 4) {
                    float p, s;
                    {
#pragma clang fp contract(off)
                        p = a * b;
                        s = a * a + b * b;
                    }
                    acc[2] = fmaf(g, p, acc[2]);
                    acc[3] = fmaf(g, s, acc[3]);
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) hres[it][q][i] = acc[q] * inx[i];
        }
    }
    __syncthreads();  // every staging read is done; reuse LDS for the horizontal results
    float* sH = lds;  // [NQ][kSH][kBlurTW]
#pragma unroll
    for (int it = 0; it < kHRowsPerThread; it++) {
        const int hr = (t >> 4) + 16 * it;
#pragma unroll
        for (int q = 0; q < NQ; q++)
            *reinterpret_cast<float4*>(sH + (q * kSH + hr) * kBlurTW + 4 * cg) =
                make_float4(hres[it][q][0], hres[it][q][1], hres[it][q][2], hres[it][q][3]);
    }
    __syncthreads();

    // ---- vertical pass: column vc, output rows vr0..vr0+7
    float iny[kVRows];
#pragma unroll
    for (int j = 0; j < kVRows; j++) iny[j] = inv_norm(ty0 + vr0 + j, h, taps);
    float vres[NQ][kVRows];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        float col[kVRows + 16];
#pragma unroll
        for (int r = 0; r < kVRows + 16; r++) col[r] = sH[(q * kSH + vr0 + r) * kBlurTW + vc];
#pragma unroll
        for (int j = 0; j < kVRows; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 17; k++) acc = fmaf(taps.g[k], col[j + k], acc);
            vres[q][j] = acc * iny[j];
        }
    }

    // ---- maps and sums
    MapSums s;
#pragma unroll
    for (int m = 0; m < 3; m++) { s.s1[m] = 0.0f; s.s4[m] = 0.0f; }
    const int x = tx0 + vc;
    const bool do_ssim = (job.flags & kMapSsim) != 0;
    const bool do_art = (job.flags & kMapArtifact) != 0;
    const bool do_det = (job.flags & kMapDetail) != 0;
#pragma unroll
    for (int j = 0; j < kVRows; j++) {
        const int y = ty0 + vr0 + j;
        if (x < w && y < h) {
            if constexpr (NQ == 4) {
                if (do_ssim) {
                    const float v = ssim_value(vres[0][j], vres[1][j], vres[2][j], vres[3][j]);
                    s.s1[0] += v;
                    s.s4[0] += pow4(v);
                }
            }
            if (do_art || do_det) {
                const float d = edge_value(ca[j], cb[j], vres[0][j], vres[1][j]);
                const float art = fmaxf(d, 0.0f), det = fmaxf(-d, 0.0f);
                s.s1[1] += art; s.s4[1] += pow4(art);
                s.s1[2] += det; s.s4[2] += pow4(det);
            }
        }
    }

    // ---- workgroup reduction of the 6 sums
    float v6[6] = {s.s1[0], s.s1[1], s.s1[2], s.s4[0], s.s4[1], s.s4[2]};
#pragma unroll
    for (int m = 0; m < 6; m++) v6[m] = wave_sum(v6[m]);
    const int wave = t >> 6, lane = t & 63;
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < 6; m++) red[wave * 8 + m] = v6[m];
    }
    __syncthreads();
    if (t < 6) {
        float tot = red[t];
        tot += red[8 + t];
        tot += red[16 + t];
        tot += red[24 + t];
        partial[t] = tot;
    }
}

template <typename StoreT>
__global__ __launch_bounds__(kBlurThreads)
void blur_map_kernel(const MapJob* __restrict__ jobs, int njobs, const int* __restrict__ total_tiles,
                     BlurTaps taps, float* __restrict__ partials)
{
    __shared__ __attribute__((aligned(16))) float lds[4 * kSH * kBlurTW];
    __shared__ float red[32];
    const int total = *total_tiles;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].tile_start <= item) lo = mid; else hi = mid - 1;
        }
        const MapJob job = jobs[lo];
        const int local = item - job.tile_start;
        const int tile_x = local % job.tiles_x;
        const int tile_y = local / job.tiles_x;
        __syncthreads();  // LDS of the previous item is free
        if (job.flags & kMapSsim)
            blur_tile<StoreT, 4>(job, tile_x, tile_y, taps, lds, red, partials + (int64_t)item * kPartialStride);
        else
            blur_tile<StoreT, 2>(job, tile_x, tile_y, taps, lds, red, partials + (int64_t)item * kPartialStride);
    }
}

static_assert(2 * kSH * kSW <= 4 * kSH * kBlurTW, "staging must fit in the horizontal-result buffer");

}  // namespace

template <typename StoreT>
hipError_t launch_blur(const MapJob* jobs, int njobs, const int* total_tiles, int grid, const BlurTaps& taps,
                       float* partials, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(blur_map_kernel<StoreT>, dim3(grid), dim3(kBlurThreads), 0, stream,
                       jobs, njobs, total_tiles, taps, partials);
    return hipGetLastError();
}

template hipError_t launch_blur<float>(const MapJob*, int, const int*, int, const BlurTaps&, float*, hipStream_t);
template hipError_t launch_blur<__half>(const MapJob*, int, const int*, int, const BlurTaps&, float*, hipStream_t);

}  // namespace vship
