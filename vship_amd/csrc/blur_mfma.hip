// blur_mfma.hip - the 17-tap blur of the SSIMULACRA2 moments as banded
// Toeplitz products on the gfx950 matrix cores, fused with the SSIM/edge maps
// and the per-tile partial sums (same outputs as blur_map_kernel).
//
// Tile: 64 x 48 output pixels of one (pair, plane, scale), 256 threads.
// The (64+16) x (48+16) input window of both images is staged in LDS as in
// blur.hip (centred on the tile's first reference sample, zero outside the
// image), widened to 96 columns so that the horizontal pass runs over three
// aligned 32-column chunks.
//
// Each pass is  out[j] = sum_k g[k] in[j + k]  (reference
// ssimu2/gaussianblur.hpp:24-77 before its renormalisation), i.e. a product
// with a banded Toeplitz matrix. With v_mfma_f32_16x16x32_f16 a 16-wide output
// block needs the 32 inputs starting at its first column, which fall into one
// or two aligned chunks; the chunk's 32 x 16 Toeplitz block is
// B_d[k][n] = g[k + d - n] with d in {0, -16, +16}. The same three blocks, in
// the A-operand layout, serve the vertical pass.
//
// Precision: FP16 operands carry FP32 values split as hi + lo (hi = fp16(x),
// lo = fp16(x - hi)), and the taps likewise; each product is
// hi*hi + lo*hi + hi*lo with FP32 accumulation, about 22 significant bits.
// Values are scaled by a power of two per tile (from the largest centred
// magnitude in the window) so that hi and lo stay in FP16's normal range;
// scaling by a power of two is exact and is undone at the end.
//
// Identity: the blurred quantities are a, b, a*b and (a*a + b*b) / 2. For
// a == b the last two are the same float, so their blurs are the same bits
// and sigma11 + sigma22 = 2 * sigma12 exactly, as in the FP32 kernel.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "blur_maps.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTW = 64, kTH = 48;                 // output tile
constexpr int kWR = kTH + 2 * kHalo;              // window rows (64)
constexpr int kWC = 96;                           // window columns (80 used, 3 chunks of 32)
constexpr int kWP = kWC + 4;                      // LDS row pitch of the window (floats)
constexpr int kHP = kWR + 8;                      // LDS column pitch of the horizontal results (halves)
constexpr int kWindowFloats = 2 * kWR * kWP;      // both images
constexpr int kHHalves = 4 * 2 * kTW * kHP;       // [quantity][hi, lo][column][row]
constexpr int kLdsBytes = (kWindowFloats * 4 > kHHalves * 2) ? kWindowFloats * 4 : kHHalves * 2;
static_assert(kBlurThreads == 256, "four waves: one window row block each");

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline void split(float x, _Float16& hi, _Float16& lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

__device__ inline floatx4 mma3(const half8& ah, const half8& al, const half8& bh, const half8& bl, floatx4 c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
    return c;
}

template <typename StoreT>
__device__ inline float load_elem(const StoreT* p) { return load_cast<StoreT>(*p); }

template <typename StoreT>
__device__ inline void load4v(const StoreT* p, float v[4]);

template <>
__device__ inline void load4v<float>(const float* p, float v[4])
{
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

template <>
__device__ inline void load4v<__half>(const __half* p, float v[4])
{
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    v[0] = __half2float(__ushort_as_half((unsigned short)(q.x & 0xffff)));
    v[1] = __half2float(__ushort_as_half((unsigned short)(q.x >> 16)));
    v[2] = __half2float(__ushort_as_half((unsigned short)(q.y & 0xffff)));
    v[3] = __half2float(__ushort_as_half((unsigned short)(q.y >> 16)));
}

template <typename StoreT, int NQ>
__device__ void blur_tile_mfma(const MapJob& job, int tile_x, int tile_y, const BlurTaps& taps,
                               const half8 (&toe_hi)[3], const half8 (&toe_lo)[3], char* lds, float* red,
                               float* __restrict__ partial)
{
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int ln = lane & 15, lg = lane >> 4;  // MFMA lane coordinates: column/row 0..15, k group 0..3
    const int w = job.w, h = job.h, pitch = job.pitch;
    const int tx0 = tile_x * kTW, ty0 = tile_y * kTH;
    const StoreT* A = static_cast<const StoreT*>(job.a);
    const StoreT* B = static_cast<const StoreT*>(job.b);
    float* sA = reinterpret_cast<float*>(lds);
    float* sB = sA + kWR * kWP;

    // ---- stage the window (u, v) = (x - tx0 + 8, y - ty0 + 8), centred on c,
    // four columns per item (x - 8 is a multiple of 4 and plane rows are
    // padded to 64 elements, so every vector load stays inside its row)
    const float c = load_elem<StoreT>(A + (int64_t)ty0 * pitch + tx0);
    float vmax = 0.0f;
    // every thread stages exactly kStageIters items; all loads are issued
    // before any is used
    constexpr int kItems = kWR * (kWC / 4);
    constexpr int kStageIters = kItems / kBlurThreads;
    static_assert(kItems % kBlurThreads == 0, "staging items per thread");
    float ra[kStageIters][4], rb[kStageIters][4];
    bool ok[kStageIters];
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        const int v = e / (kWC / 4);
        const int u = 4 * (e - v * (kWC / 4));
        const int gy = ty0 - kHalo + v, gx = tx0 - kHalo + u;
        ok[it] = gy >= 0 && gy < h && gx >= 0 && gx < pitch && u < kTW + 2 * kHalo;
        if (ok[it]) {
            const int64_t o = (int64_t)gy * pitch + gx;
            load4v<StoreT>(A + o, ra[it]);
            load4v<StoreT>(B + o, rb[it]);
        }
    }
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        const int v = e / (kWC / 4);
        const int u = 4 * (e - v * (kWC / 4));
        const int gx = tx0 - kHalo + u;
        float va[4] = {0.f, 0.f, 0.f, 0.f}, vb[4] = {0.f, 0.f, 0.f, 0.f};
        if (ok[it]) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                // FP16 planes hold the distorted image as its difference from
                // the reference: b = a + d
                const float bv = (sizeof(StoreT) == 2) ? ra[it][i] + rb[it][i] : rb[it][i];
                va[i] = (gx + i < w) ? ra[it][i] - c : 0.f;
                vb[i] = (gx + i < w) ? bv - c : 0.f;
            }
        }
        *reinterpret_cast<float4*>(sA + v * kWP + u) = make_float4(va[0], va[1], va[2], va[3]);
        *reinterpret_cast<float4*>(sB + v * kWP + u) = make_float4(vb[0], vb[1], vb[2], vb[3]);
#pragma unroll
        for (int i = 0; i < 4; i++) vmax = fmaxf(vmax, fmaxf(fabsf(va[i]), fabsf(vb[i])));
    }
    vmax = wave_max(vmax);
    if (lane == 0) red[wave] = vmax;
    __syncthreads();
    vmax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    // power-of-two scales: values to at most 2^14, products to at most 2^14
    int e1 = 14, e2 = 14;
    if (vmax > 0.0f) {
        int ex;
        frexpf(vmax, &ex);  // vmax < 2^ex
        e1 = 14 - ex;
        e2 = 14 - 2 * ex;
    }
    e1 = max(-60, min(e1, 60));
    e2 = max(-60, min(e2, 60));
    const float s1 = ldexpf(1.0f, e1), s2 = ldexpf(1.0f, e2);
    const float u1 = ldexpf(1.0f, -e1), u2 = ldexpf(1.0f, -e2);

    // ---- centred samples at the vertical-pass outputs (for the edge maps):
    // wave owns columns 16*wave + ln; rows 16*rb + 4*lg + i
    float ca[3][4], cb[3][4];
#pragma unroll
    for (int rb = 0; rb < 3; rb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int v = kHalo + 16 * rb + 4 * lg + i, u = kHalo + 16 * wave + ln;
            ca[rb][i] = sA[v * kWP + u];
            cb[rb][i] = sB[v * kWP + u];
        }

    // ---- horizontal pass: wave owns window rows 16*wave .. +15
    floatx4 hacc[4][NQ];
#pragma unroll
    for (int cbk = 0; cbk < 4; cbk++)
#pragma unroll
        for (int q = 0; q < NQ; q++) hacc[cbk][q] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        const int v = 16 * wave + ln;         // A operand: row = lane % 16
        const int u0 = 32 * ch + 8 * lg;      // k = 8 * (lane / 16) + j
        half8 ah[NQ], al[NQ];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float a = sA[v * kWP + u0 + j], b = sB[v * kWP + u0 + j];
            float x[4];
            x[0] = a * s1;
            x[1] = b * s1;
            if constexpr (NQ == 4) {
#pragma clang fp contract(off)
                const float p = a * b, hs = (a * a + b * b) * 0.5f;
                x[2] = p * s2;
                x[3] = hs * s2;
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                _Float16 hi, lo;
                split(x[q], hi, lo);
                ah[q][j] = hi;
                al[q][j] = lo;
            }
        }
        // output column block x0 = 16*cbk reads chunk ch with offset
        // d = 32*ch - x0 in {0, -16, +16}; fragment index 0: d=0, 1: d=-16, 2: d=+16
#pragma unroll
        for (int cbk = 0; cbk < 4; cbk++) {
            const int d = 32 * ch - 16 * cbk;
            if (d != 0 && d != -16 && d != 16) continue;
            const int f = d == 0 ? 0 : (d < 0 ? 1 : 2);
#pragma unroll
            for (int q = 0; q < NQ; q++) hacc[cbk][q] = mma3(ah[q], al[q], toe_hi[f], toe_lo[f], hacc[cbk][q]);
        }
    }
    __syncthreads();  // every window read is done: the LDS now takes the horizontal results

    // ---- horizontal results as hi/lo FP16, [q][hl][column][row]: the lane
    // holds rows 16*wave + 4*lg + i of column 16*cbk + ln
    _Float16* sH = reinterpret_cast<_Float16*>(lds);
#pragma unroll
    for (int cbk = 0; cbk < 4; cbk++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            half4 hi, lo;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                _Float16 h16, l16;
                split(hacc[cbk][q][i], h16, l16);
                hi[i] = h16;
                lo[i] = l16;
            }
            const int col = 16 * cbk + ln, row = 16 * wave + 4 * lg;
            *reinterpret_cast<half4*>(sH + ((q * 2 + 0) * kTW + col) * kHP + row) = hi;
            *reinterpret_cast<half4*>(sH + ((q * 2 + 1) * kTW + col) * kHP + row) = lo;
        }
    __syncthreads();

    // ---- vertical pass: wave owns output columns 16*wave .. +15, all 48
    // rows; output row block y0 = 16*rb reads row chunk rc with offset
    // d = 32*rc - y0 (the Toeplitz block is the A operand here)
    floatx4 vacc[3][NQ];
#pragma unroll
    for (int rb = 0; rb < 3; rb++)
#pragma unroll
        for (int q = 0; q < NQ; q++) vacc[rb][q] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rc = 0; rc < 2; rc++) {
        const int col = 16 * wave + ln, row0 = 32 * rc + 8 * lg;  // B operand: k = 8*(lane/16)+j, n = lane%16
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const half8 bh = *reinterpret_cast<const half8*>(sH + ((q * 2 + 0) * kTW + col) * kHP + row0);
            const half8 bl = *reinterpret_cast<const half8*>(sH + ((q * 2 + 1) * kTW + col) * kHP + row0);
#pragma unroll
            for (int rb = 0; rb < 3; rb++) {
                const int d = 32 * rc - 16 * rb;
                if (d != 0 && d != -16 && d != 16) continue;
                const int f = d == 0 ? 0 : (d < 0 ? 1 : 2);
                vacc[rb][q] = mma3(toe_hi[f], toe_lo[f], bh, bl, vacc[rb][q]);
            }
        }
    }

    // ---- maps and sums: the lane holds output (x, y) = (tx0 + 16*wave + ln,
    // ty0 + 16*rb + 4*lg + i)
    float s1m[3] = {0.f, 0.f, 0.f}, s4m[3] = {0.f, 0.f, 0.f};
    const int x = tx0 + 16 * wave + ln;
    const float inx = inv_norm(x, w, taps);
    const bool do_ssim = (job.flags & kMapSsim) != 0;
    const bool do_edge = (job.flags & (kMapArtifact | kMapDetail)) != 0;
#pragma unroll
    for (int rb = 0; rb < 3; rb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int y = ty0 + 16 * rb + 4 * lg + i;
            if (x < w && y < h) {
                const float nrm = inx * inv_norm(y, h, taps);
                const float k1 = u1 * nrm;
                const float m1 = vacc[rb][0][i] * k1, m2 = vacc[rb][1][i] * k1;
                if constexpr (NQ == 4) {
                    if (do_ssim) {
                        const float k2 = u2 * nrm;
                        const float s12 = vacc[rb][2][i] * k2;
                        const float sq = (vacc[rb][3][i] * k2) * 2.0f;
                        const float val = ssim_value(m1, m2, s12, sq);
                        s1m[0] += val;
                        s4m[0] += pow4(val);
                    }
                }
                if (do_edge) {
                    const float dd = edge_value(ca[rb][i], cb[rb][i], m1, m2);
                    const float art = fmaxf(dd, 0.0f), det = fmaxf(-dd, 0.0f);
                    s1m[1] += art;
                    s4m[1] += pow4(art);
                    s1m[2] += det;
                    s4m[2] += pow4(det);
                }
            }
        }

    // ---- workgroup reduction of the 6 sums
    float v6[6] = {s1m[0], s1m[1], s1m[2], s4m[0], s4m[1], s4m[2]};
#pragma unroll
    for (int m = 0; m < 6; m++) v6[m] = wave_sum(v6[m]);
    __syncthreads();  // red[] held the window maximum
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
void blur_mfma_kernel(const MapJob* __restrict__ jobs, const int* __restrict__ counts, BlurTaps taps,
                      float* __restrict__ partials)
{
    __shared__ __attribute__((aligned(16))) char lds[kLdsBytes];
    __shared__ float red[32];
    // Toeplitz fragments, identical in the B layout (k = 8*(lane/16)+j,
    // n = lane%16) and the A layout (row = lane%16, k = 8*(lane/16)+j):
    // value g[k + d - n] for d = 0, -16, +16, zero outside the 17 taps
    const int lane = threadIdx.x & 63;
    half8 toe_hi[3], toe_lo[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const int d = f == 0 ? 0 : (f == 1 ? -16 : 16);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = 8 * (lane >> 4) + j + d - (lane & 15);
            float g = 0.0f;  // select chain: a lane-varying index would spill the taps to scratch
#pragma unroll
            for (int m = 0; m < 17; m++) g = (k == m) ? taps.g[m] : g;
            _Float16 hi, lo;
            split(g, hi, lo);
            toe_hi[f][j] = hi;
            toe_lo[f][j] = lo;
        }
    }
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
            blur_tile_mfma<StoreT, 4>(job, tile_x, tile_y, taps, toe_hi, toe_lo, lds, red,
                                      partials + (int64_t)(job.part_start + local) * kPartialStride);
        else
            blur_tile_mfma<StoreT, 2>(job, tile_x, tile_y, taps, toe_hi, toe_lo, lds, red,
                                      partials + (int64_t)(job.part_start + local) * kPartialStride);
    }
}

}  // namespace

template <typename StoreT>
hipError_t launch_blur_mfma(const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps, float* partials,
                            hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (grid % kXcds) return hipErrorInvalidValue;
    hipLaunchKernelGGL(blur_mfma_kernel<StoreT>, dim3(grid), dim3(kBlurThreads), 0, stream, jobs, counts, taps,
                       partials);
    return hipGetLastError();
}

template hipError_t launch_blur_mfma<float>(const MapJob*, const int*, int, const BlurTaps&, float*, hipStream_t);
template hipError_t launch_blur_mfma<__half>(const MapJob*, const int*, int, const BlurTaps&, float*, hipStream_t);

}  // namespace vship
