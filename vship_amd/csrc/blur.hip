// blur.hip - fused 17-tap blur + SSIM/edge maps + per-tile partial sums.
//
// One workgroup (256 threads) produces a TW x TH output tile of one active
// (pair, plane, scale). It stages the (TW+16) x (TH+16) input window of both
// images in LDS (zero outside the image, reference gaussianblur.hpp:24-36),
// runs the horizontal 17-tap pass register-blocked (each item: 4 adjacent
// columns of one row, a 20-sample window read as five 128-bit LDS loads),
// writes the horizontal results over the staging buffer, runs the vertical
// pass (each thread: one column, TW*TH/256 rows) and evaluates the maps in
// registers. Each axis is divided by the sum of its in-bounds taps
// (gaussianblur.hpp:52-54,68-70). The blurred quantities are a, b, a*b and
// a*a + b*b when the SSIM map is active (4 blurs), otherwise a and b only.
// Sums of x and x^4 of the three maps are reduced across the workgroup and
// written as one 8-float partial per tile; the FP64 score kernel finishes the
// norms. Tile shape is a template parameter chosen per launch (autotuned for
// the batched path).
//
// Workgroups are dispatched round-robin over the 8 XCDs; the kernel gives each
// XCD a contiguous range of tiles so that neighbouring tiles, which share
// 16 halo columns and rows, read them from the same L2.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "blur_maps.hpp"
#include "kernels.hpp"

namespace vship {

namespace {


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

template <int TW, int TH>
struct BlurShape {
    static constexpr int SW = TW + 2 * kHalo;               // staged columns
    static constexpr int SH = TH + 2 * kHalo;               // staged rows
    static constexpr int HITEMS = SH * (TW / 4);            // horizontal items (row, 4 columns)
    static constexpr int HPER = (HITEMS + kBlurThreads - 1) / kBlurThreads;
    static constexpr int VR = TW * TH / kBlurThreads;        // vertical rows per thread
    static constexpr int STAGE = 2 * SH * SW;
    static constexpr int HOUT = 4 * SH * TW;
    static constexpr int LDS = STAGE > HOUT ? STAGE : HOUT;  // floats
    static_assert(TW % 4 == 0 && kBlurThreads % TW == 0 && kBlurThreads % (TW / 4) == 0, "tile width");
    static_assert((TW * TH) % kBlurThreads == 0, "tile area");
};

template <typename StoreT, int TW, int TH, int NQ>
__device__ void blur_tile(const MapJob& job, int tile_x, int tile_y, const BlurTaps& taps, float* lds,
                          float* red, float* __restrict__ partial)
{
    using Sh = BlurShape<TW, TH>;
    const int t = threadIdx.x;
    const int w = job.w, h = job.h, pitch = job.pitch;
    const int tx0 = tile_x * TW, ty0 = tile_y * TH;
    const StoreT* A = static_cast<const StoreT*>(job.a);
    const StoreT* B = static_cast<const StoreT*>(job.b);
    float* sA = lds;
    float* sB = lds + Sh::SH * Sh::SW;

    // ---- stage the input windows, centred on c = a(tx0, ty0) (a sample
    // inside the image, the same for both images so that a == b stays
    // bit-symmetric), zero outside the image. The renormalised blur averages
    // in-bounds samples only, so G(a - c) = G(a) - c and every moment below is
    // the same as with uncentred values; centring keeps E[x^2] - mu^2 from
    // cancelling where the window is flat (small scales, smooth content).
    const float c = load_cast<StoreT>(A[(int64_t)ty0 * pitch + tx0]);
    for (int e = t; e < Sh::SH * (Sh::SW / 4); e += kBlurThreads) {
        const int r = e / (Sh::SW / 4);
        const int c4 = e - r * (Sh::SW / 4);
        const int gy = ty0 - kHalo + r;
        const int gx = tx0 - kHalo + 4 * c4;
        float va[4] = {0.f, 0.f, 0.f, 0.f}, vb[4] = {0.f, 0.f, 0.f, 0.f};
        if (gy >= 0 && gy < h && gx >= 0 && gx < pitch) {
            const int64_t o = (int64_t)gy * pitch + gx;
            load4<StoreT>(A + o, va);
            load4<StoreT>(B + o, vb);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                // FP16 planes hold the distorted image as its difference
                // from the reference: b = a + d
                const float bv = (sizeof(StoreT) == 2) ? va[i] + vb[i] : vb[i];
                va[i] = (gx + i < w) ? va[i] - c : 0.f;
                vb[i] = (gx + i < w) ? bv - c : 0.f;
            }
        }
        *reinterpret_cast<float4*>(sA + r * Sh::SW + 4 * c4) = make_float4(va[0], va[1], va[2], va[3]);
        *reinterpret_cast<float4*>(sB + r * Sh::SW + 4 * c4) = make_float4(vb[0], vb[1], vb[2], vb[3]);
    }
    __syncthreads();

    // ---- centred samples for the edge maps (vertical-pass ownership)
    const int vc = t % TW;
    const int vr0 = (t / TW) * Sh::VR;
    float ca[Sh::VR], cb[Sh::VR];
#pragma unroll
    for (int j = 0; j < Sh::VR; j++) {
        ca[j] = sA[(kHalo + vr0 + j) * Sh::SW + kHalo + vc];
        cb[j] = sB[(kHalo + vr0 + j) * Sh::SW + kHalo + vc];
    }

    // ---- horizontal pass; 256 is a multiple of TW/4, so every item of a
    // thread has the same column group
    const int cg = t % (TW / 4);
    float inx[4];
#pragma unroll
    for (int i = 0; i < 4; i++) inx[i] = inv_norm(tx0 + 4 * cg + i, w, taps);
    float hres[Sh::HPER][NQ][4];
#pragma unroll
    for (int it = 0; it < Sh::HPER; it++) {
        const int item = t + it * kBlurThreads;
        if (Sh::HITEMS % kBlurThreads != 0 && item >= Sh::HITEMS) break;
        const int hr = item / (TW / 4);
        float wa[20], wb[20];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const float4 qa = *reinterpret_cast<const float4*>(sA + hr * Sh::SW + 4 * cg + 4 * k);
            const float4 qb = *reinterpret_cast<const float4*>(sB + hr * Sh::SW + 4 * cg + 4 * k);
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
    float* sH = lds;  // [NQ][SH][TW]
#pragma unroll
    for (int it = 0; it < Sh::HPER; it++) {
        const int item = t + it * kBlurThreads;
        if (Sh::HITEMS % kBlurThreads != 0 && item >= Sh::HITEMS) break;
        const int hr = item / (TW / 4);
#pragma unroll
        for (int q = 0; q < NQ; q++)
            *reinterpret_cast<float4*>(sH + (q * Sh::SH + hr) * TW + 4 * cg) =
                make_float4(hres[it][q][0], hres[it][q][1], hres[it][q][2], hres[it][q][3]);
    }
    __syncthreads();

    // ---- vertical pass: column vc, output rows vr0 .. vr0+VR-1
    float vres[NQ][Sh::VR];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        float col[Sh::VR + 16];
#pragma unroll
        for (int r = 0; r < Sh::VR + 16; r++) col[r] = sH[(q * Sh::SH + vr0 + r) * TW + vc];
#pragma unroll
        for (int j = 0; j < Sh::VR; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 17; k++) acc = fmaf(taps.g[k], col[j + k], acc);
            vres[q][j] = acc;
        }
    }

    // ---- maps and sums
    float s1[3] = {0.f, 0.f, 0.f}, s4[3] = {0.f, 0.f, 0.f};
    const int x = tx0 + vc;
    const bool do_ssim = (job.flags & kMapSsim) != 0;
    const bool do_edge = (job.flags & (kMapArtifact | kMapDetail)) != 0;
#pragma unroll
    for (int j = 0; j < Sh::VR; j++) {
        const int y = ty0 + vr0 + j;
        if (x < w && y < h) {
            const float iny = inv_norm(y, h, taps);
            const float m1 = vres[0][j] * iny, m2 = vres[1][j] * iny;
            if constexpr (NQ == 4) {
                if (do_ssim) {
                    const float v = ssim_value(m1, m2, vres[2][j] * iny, vres[3][j] * iny);
                    s1[0] += v;
                    s4[0] += pow4(v);
                }
            }
            if (do_edge) {
                const float d = edge_value(ca[j], cb[j], m1, m2);
                const float art = fmaxf(d, 0.0f), det = fmaxf(-d, 0.0f);
                s1[1] += art; s4[1] += pow4(art);
                s1[2] += det; s4[2] += pow4(det);
            }
        }
    }

    // ---- workgroup reduction of the 6 sums
    float v6[6] = {s1[0], s1[1], s1[2], s4[0], s4[1], s4[2]};
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

template <typename StoreT, int TW, int TH>
__global__ __launch_bounds__(kBlurThreads)
void blur_map_kernel(const MapJob* __restrict__ jobs, const int* __restrict__ counts, BlurTaps taps,
                     float* __restrict__ partials)
{
    __shared__ __attribute__((aligned(16))) float lds[BlurShape<TW, TH>::LDS];
    __shared__ float red[32];
    const int total = counts[kCountBlurTiles];
    const int njobs = counts[kCountBlurJobs];
    const int per = (total + kXcds - 1) / kXcds;
    const int xcd = blockIdx.x % kXcds;
    const int slots = gridDim.x / kXcds;
    const int end = min((xcd + 1) * per, total);
    for (int item = xcd * per + blockIdx.x / kXcds; item < end; item += slots) {
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
            blur_tile<StoreT, TW, TH, 4>(job, tile_x, tile_y, taps, lds, red, partials + (int64_t)(job.part_start + local) * kPartialStride);
        else
            blur_tile<StoreT, TW, TH, 2>(job, tile_x, tile_y, taps, lds, red, partials + (int64_t)(job.part_start + local) * kPartialStride);
    }
}

template <typename StoreT, int TW, int TH>
hipError_t launch_one(const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps, float* partials,
                      hipStream_t stream)
{
    hipLaunchKernelGGL((blur_map_kernel<StoreT, TW, TH>), dim3(grid), dim3(kBlurThreads), 0, stream, jobs, counts,
                       taps, partials);
    return hipGetLastError();
}

}  // namespace

BlurTile blur_tile_shape(int config)
{
    switch (config) {
    case 1: return {64, 16};
    case 2: return {32, 64};
    case 3: return {32, 32};
    case 4: return {64, 64};
    case 5: return {64, 48};
    case 6: return {64, 64};
    case 7: return {64, 128};
    case 8: return {64, 256};
    default: return {64, 32};
    }
}

int blur_tile_configs() { return 9; }

int blur_grid_for(int total_tiles)
{
    return (total_tiles + kXcds - 1) / kXcds * kXcds;
}

template <typename StoreT>
hipError_t launch_blur(int config, const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps,
                       float* partials, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (grid % kXcds) return hipErrorInvalidValue;
    switch (config) {
    case 1: return launch_one<StoreT, 64, 16>(jobs, counts, grid, taps, partials, stream);
    case 2: return launch_one<StoreT, 32, 64>(jobs, counts, grid, taps, partials, stream);
    case 3: return launch_one<StoreT, 32, 32>(jobs, counts, grid, taps, partials, stream);
    case 4: return launch_one<StoreT, 64, 64>(jobs, counts, grid, taps, partials, stream);
    case 5: return launch_blur_mfma<StoreT>(jobs, counts, grid, taps, partials, stream);
    case 6: return launch_blur_strip<StoreT>(4, jobs, counts, grid, taps, partials, stream);
    case 7: return launch_blur_strip<StoreT>(8, jobs, counts, grid, taps, partials, stream);
    case 8: return launch_blur_strip<StoreT>(16, jobs, counts, grid, taps, partials, stream);
    default: return launch_one<StoreT, 64, 32>(jobs, counts, grid, taps, partials, stream);
    }
}

template hipError_t launch_blur<float>(int, const MapJob*, const int*, int, const BlurTaps&, float*, hipStream_t);
template hipError_t launch_blur<__half>(int, const MapJob*, const int*, int, const BlurTaps&, float*, hipStream_t);

}  // namespace vship
