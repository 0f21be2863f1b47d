// blur_strip.hip - the 17-tap blur of the SSIMULACRA2 moments on the gfx950
// matrix cores, streamed down 64-column strips, fused with the SSIM/edge maps
// and the per-item partial sums (same outputs as blur_map_kernel).
//
// Item: 64 output columns x 16*SEGB output rows of one (pair, plane, scale),
// 256 threads. Wave w owns output columns 16w..16w+15 and walks down the
// item's window one 16-row block at a time:
//   horizontal  H_kb = W_kb * G          v_mfma_f32_16x16x32_f16
//               (16 window rows x 32 window columns 16w..16w+31 times the
//               32x16 Toeplitz block G[k][n] = g[k - n])
//   vertical    V_rb = T0 * H_rb + T1 * H_rb+1     v_mfma_f32_16x16x16_f16
//               (T0[m][k] = g[k - m], T1[m][k] = g[16 + k - m])
// which is reference ssimu2/gaussianblur.hpp:24-77 before its
// renormalisation. The accumulator layout of a 16x16 product (lane holds rows
// 4*(lane/16)+i of column lane%16) is exactly the B-operand layout of the
// 16x16x16 product, so the horizontal results feed the vertical products from
// registers: they never go through LDS, and each window row is converted and
// multiplied once per item (no vertical halo recompute between blocks).
//
// The window (80 columns, zero outside the image, centred on the item's first
// reference sample) streams through a ring of three 16-row LDS slots: while
// block kb is multiplied, block kb+1 is loaded into registers and written to
// the slot that block kb-2 held.
//
// Precision: FP16 operands carry FP32 values as hi + lo (hi = fp16 of x
// rounded toward zero, lo = fp16(x - hi)), the taps likewise (round to
// nearest); each product is hi*hi + lo*hi + hi*lo with FP32 accumulation,
// about 22 significant bits. Every wave scales each window block by a power of
// two from the largest magnitude in its own 16x32 operand block, so hi and lo
// stay in FP16's normal range; the vertical products of the two blocks that
// meet in one output block go to separate accumulators and are unscaled
// (exactly) before the one FP32 add that joins them.
//
// Identity: the blurred quantities are a, b, a*b and (a*a + b*b) / 2. For
// a == b the last two are the same float and every operation on a and b is
// the same, so sigma11 + sigma22 = 2 * sigma12 bit for bit and the map is 0.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "blur_maps.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int uintx2 __attribute__((ext_vector_type(2)));

constexpr int kSW = 64;                       // output columns per item (4 waves x 16)
constexpr int kWinC = kSW + 2 * kHalo;        // window columns (80)
constexpr int kRP = kWinC + 4;                // LDS row pitch (floats)
constexpr int kSlots = 3;                     // ring of 16-row window blocks
constexpr int kSlotFloats = 16 * kRP;         // one image, one block
constexpr int kVecPerRow = kWinC / 4;         // 4-sample vectors per window row (20)
constexpr int kVecPerBlock = 16 * kVecPerRow; // per image and block (320)
constexpr int kStageIters = (kVecPerBlock + kBlurThreads - 1) / kBlurThreads;
static_assert(kBlurThreads == 256, "four waves, one 16-column strip each");

// Wave reductions on the DPP network: row shifts 1/2/4/8 reduce each 16-lane
// row into its lane 15, two row broadcasts (15, 31) fold the rows into lane
// 63, whose value is returned to every lane. Out-of-row sources read 0
// (bound_ctrl), the identity of both reductions here (sums, and maxima of
// non-negative values compared as their bit patterns).
__device__ inline float wave_sum(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xf, 0xf, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xf, 0xf, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xf, 0xf, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xf, 0xf, true));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xf, 0xf, true));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline uint32_t wave_max_bits(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xf, 0xf, true));
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xf, 0xf, true));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Global-memory view of a plane pointer read from a job table (loads through
// it are global loads, not flat loads that also wait on the LDS counter).
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* global_ptr(const void* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

// x = v * s (s a power of two: exact) as hi + lo FP16 for two values: hi
// rounded toward zero by one packed conversion, lo = fp16(x - hi) from one
// mixed-precision FMA each (x - hi is exact in FP32).
__device__ inline void split2(float v0, float v1, float s, _Float16& h0, _Float16& h1, _Float16& l0, _Float16& l1)
{
    const float x0 = v0 * s, x1 = v1 * s;
    const fp16x2 hh = __builtin_amdgcn_cvt_pkrtz(x0, x1);
    h0 = (_Float16)hh[0];
    h1 = (_Float16)hh[1];
    l0 = (_Float16)(x0 - (float)hh[0]);
    l1 = (_Float16)(x1 - (float)hh[1]);
}

// Four accumulator values (already scaled) as hi + lo FP16, hi rounded
// toward zero.
__device__ inline void split4(const float (&x)[4], half4& hi, half4& lo)
{
    const fp16x2 h01 = __builtin_amdgcn_cvt_pkrtz(x[0], x[1]);
    const fp16x2 h23 = __builtin_amdgcn_cvt_pkrtz(x[2], x[3]);
    hi[0] = (_Float16)h01[0];
    hi[1] = (_Float16)h01[1];
    hi[2] = (_Float16)h23[0];
    hi[3] = (_Float16)h23[1];
    lo[0] = (_Float16)(x[0] - (float)h01[0]);
    lo[1] = (_Float16)(x[1] - (float)h01[1]);
    lo[2] = (_Float16)(x[2] - (float)h23[0]);
    lo[3] = (_Float16)(x[3] - (float)h23[1]);
}

__device__ inline void split_rn(float x, _Float16& hi, _Float16& lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

__device__ inline floatx4 mma32(const half8& ah, const half8& al, const half8& bh, const half8& bl)
{
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
    return c;
}

__device__ inline floatx4 mma16(const half4& ah, const half4& al, const half4& bh, const half4& bl)
{
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, c, 0, 0, 0);
    return c;
}

// Toeplitz fragments of one lane (select chains: a lane-varying index into
// the taps would put them in scratch memory).
struct Toeplitz {
    half8 g_hi, g_lo;    // horizontal B operand: k = 8*(lane/16)+j, n = lane%16, g[k - n]
    half4 t0_hi, t0_lo;  // vertical A operand: m = lane%16, k = 4*(lane/16)+j, g[k - m]
    half4 t1_hi, t1_lo;  //                                               g[16 + k - m]
};

__device__ inline float tap_at(const BlurTaps& taps, int k)
{
    float g = 0.0f;
#pragma unroll
    for (int m = 0; m < 17; m++) g = (k == m) ? taps.g[m] : g;
    return g;
}

__device__ inline Toeplitz make_toeplitz(const BlurTaps& taps, int lane)
{
    Toeplitz T;
    const int ln = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        _Float16 h, l;
        split_rn(tap_at(taps, 8 * lg + j - ln), h, l);
        T.g_hi[j] = h;
        T.g_lo[j] = l;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        _Float16 h, l;
        split_rn(tap_at(taps, 4 * lg + j - ln), h, l);
        T.t0_hi[j] = h;
        T.t0_lo[j] = l;
        split_rn(tap_at(taps, 16 + 4 * lg + j - ln), h, l);
        T.t1_hi[j] = h;
        T.t1_lo[j] = l;
    }
    return T;
}

// The fragments of every lane, shared by the workgroup's waves through LDS
// (structure of arrays, one 16-byte vector per lane and fragment pair, so the
// reads are conflict-free); loaded right before the products that use them,
// which keeps 16 registers free across the item.
struct ToeplitzLds {
    half8 g_hi[64], g_lo[64];
    half8 t01[64];  // t0_hi, t0_lo
    half8 t1[64];   // t1_hi, t1_lo
};

__device__ inline void toeplitz_to_lds(const Toeplitz& T, int lane, ToeplitzLds* L)
{
    L->g_hi[lane] = T.g_hi;
    L->g_lo[lane] = T.g_lo;
    half8 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        a[j] = T.t0_hi[j];
        a[4 + j] = T.t0_lo[j];
        b[j] = T.t1_hi[j];
        b[4 + j] = T.t1_lo[j];
    }
    L->t01[lane] = a;
    L->t1[lane] = b;
}

__device__ inline void split_half8(const half8& v, half4& lo4, half4& hi4)
{
#pragma unroll
    for (int j = 0; j < 4; j++) {
        lo4[j] = v[j];
        hi4[j] = v[4 + j];
    }
}

template <typename StoreT>
struct StageRegs;

template <>
struct StageRegs<float> {
    typedef floatx4 Vec;
    Vec a[kStageIters], b[kStageIters];
};

template <>
struct StageRegs<__half> {
    typedef uintx2 Vec;
    Vec a[kStageIters], b[kStageIters];
};

__device__ inline void unpack4(const floatx4& q, float v[4])
{
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

__device__ inline void unpack4(const uintx2& q, float v[4])
{
    v[0] = __half2float(__ushort_as_half((unsigned short)(q.x & 0xffff)));
    v[1] = __half2float(__ushort_as_half((unsigned short)(q.x >> 16)));
    v[2] = __half2float(__ushort_as_half((unsigned short)(q.y & 0xffff)));
    v[3] = __half2float(__ushort_as_half((unsigned short)(q.y >> 16)));
}

// Window geometry of one item.
struct StripGeom {
    int w, h, pitch;
    int tx0, ty0;  // first output pixel
};

// Global loads of window block kb (rows ty0 - 8 + 16kb ..) into registers.
// Columns tx0 - 8 + 4v..+3 are a multiple of 4 and plane rows are padded to
// 64 elements, so every vector lies entirely inside or outside its row.
template <typename StoreT>
__device__ inline void stage_load(const void* A, const void* B, const StripGeom& g, int kb, StageRegs<StoreT>& r)
{
    typedef typename StageRegs<StoreT>::Vec Vec;
    const auto* ga = global_ptr<Vec>(A);
    const auto* gb = global_ptr<Vec>(B);
    const int t = threadIdx.x;
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        const int row = e / kVecPerRow, v = e - row * kVecPerRow;
        const int gy = g.ty0 - kHalo + 16 * kb + row, gx = g.tx0 - kHalo + 4 * v;
        const bool ok = e < kVecPerBlock && gy >= 0 && gy < g.h && gx >= 0 && gx < g.pitch;
        Vec va = {}, vb = {};
        if (ok) {
            // vector index inside the plane (< 2^31 elements per plane)
            const uint32_t o = ((uint32_t)gy * (uint32_t)g.pitch + (uint32_t)gx) >> 2;
            va = ga[o];
            vb = gb[o];
        }
        r.a[it] = va;
        r.b[it] = vb;
    }
}

// Centred values of the staged block into its ring slot: a - c and b - c
// inside the image, 0 outside (rows outside were loaded as zeros, columns at
// or beyond w are masked here). FP16 planes hold the distorted image as its
// difference from the reference: b = a + d.
template <typename StoreT>
__device__ inline void stage_store(const StageRegs<StoreT>& r, const StripGeom& g, int kb, float c, float* sA,
                                   float* sB)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        if (e >= kVecPerBlock) continue;
        const int row = e / kVecPerRow, v = e - row * kVecPerRow;
        const int gy = g.ty0 - kHalo + 16 * kb + row, gx = g.tx0 - kHalo + 4 * v;
        const bool inside = gy >= 0 && gy < g.h && gx >= 0;
        float av[4], bv[4];
        unpack4(r.a[it], av);
        unpack4(r.b[it], bv);
        float va[4], vb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const bool in = inside && gx + i < g.w;
            const float bb = (sizeof(StoreT) == 2) ? av[i] + bv[i] : bv[i];
            va[i] = in ? av[i] - c : 0.0f;
            vb[i] = in ? bb - c : 0.0f;
        }
        const int off = (kb % kSlots) * kSlotFloats + row * kRP + 4 * v;
        *reinterpret_cast<float4*>(sA + off) = make_float4(va[0], va[1], va[2], va[3]);
        *reinterpret_cast<float4*>(sB + off) = make_float4(vb[0], vb[1], vb[2], vb[3]);
    }
}

template <typename StoreT, int NQ>
__device__ void blur_item(const MapJob& job, int tile_x, int tile_y, int segb, const BlurTaps& taps,
                          const ToeplitzLds* TL, float* sA, float* sB, float* red, float* __restrict__ partial)
{
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int ln = lane & 15, lg = lane >> 4;
    StripGeom g;
    g.w = job.w;
    g.h = job.h;
    g.pitch = job.pitch;
    g.tx0 = tile_x * kSW;
    g.ty0 = tile_y * 16 * segb;
    const uint32_t c_at = (uint32_t)g.ty0 * (uint32_t)g.pitch + (uint32_t)g.tx0;
    float c;
    if constexpr (sizeof(StoreT) == 4) c = global_ptr<float>(job.a)[c_at];
    else c = __half2float(__ushort_as_half(global_ptr<unsigned short>(job.a)[c_at]));
    // output row blocks of this item that hold image rows; window blocks = that + 1
    const int nrb = min(segb, (g.h - g.ty0 + 15) / 16);
    const int nwb = nrb + 1;

    StageRegs<StoreT> st;
    stage_load<StoreT>(job.a, job.b, g, 0, st);
    stage_store<StoreT>(st, g, 0, c, sA, sB);
    __syncthreads();

    const int x = g.tx0 + 16 * wave + ln;
    const float inx = inv_norm(x, g.w, taps);
    const bool do_ssim = (job.flags & kMapSsim) != 0;
    const bool do_edge = (job.flags & (kMapArtifact | kMapDetail)) != 0;
    float s1m[3] = {0.f, 0.f, 0.f}, s4m[3] = {0.f, 0.f, 0.f};
    floatx4 prev[NQ];  // T0 part of the next output block (scale of its window block)
    float prev_u1 = 0.f, prev_u2 = 0.f;

    for (int kb = 0; kb < nwb; kb++) {
        const bool more = kb + 1 < nwb;
        if (more) stage_load<StoreT>(job.a, job.b, g, kb + 1, st);

        // ---- horizontal operand: window row 16kb + ln, columns 16w + 8lg .. +7
        const float* ra = sA + (kb % kSlots) * kSlotFloats + ln * kRP + 16 * wave + 8 * lg;
        const float* rb = sB + (kb % kSlots) * kSlotFloats + ln * kRP + 16 * wave + 8 * lg;
        float a[8], b[8];
        {
            const float4 a0 = *reinterpret_cast<const float4*>(ra), a1 = *reinterpret_cast<const float4*>(ra + 4);
            const float4 b0 = *reinterpret_cast<const float4*>(rb), b1 = *reinterpret_cast<const float4*>(rb + 4);
            a[0] = a0.x; a[1] = a0.y; a[2] = a0.z; a[3] = a0.w; a[4] = a1.x; a[5] = a1.y; a[6] = a1.z; a[7] = a1.w;
            b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w; b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
        }
        uint32_t mbits = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            mbits = max(mbits, max(__float_as_uint(a[j]) & 0x7fffffffu, __float_as_uint(b[j]) & 0x7fffffffu));
        const float vmax = __uint_as_float(wave_max_bits(mbits));
        // power-of-two scales: values to below 2^14, products to below 2^14
        int e1 = 14, e2 = 14;
        if (vmax > 0.0f) {
            int ex;
            frexpf(vmax, &ex);  // vmax < 2^ex
            e1 = 14 - ex;
            e2 = 14 - 2 * ex;
        }
        e1 = max(-60, min(e1, 60));
        e2 = max(-60, min(e2, 60));
        const float s1 = ldexpf(1.0f, e1), s2 = ldexpf(1.0f, e2), s2h = ldexpf(1.0f, e2 - 1);
        const float u1 = ldexpf(1.0f, -e1), u2 = ldexpf(1.0f, -e2);

        // ---- per quantity: operand hi/lo, horizontal product, hi/lo of the
        // results (the vertical B operand)
        half4 bh[NQ], bl[NQ];
        const half8 g_hi = TL->g_hi[lane], g_lo = TL->g_lo[lane];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            half8 ah, al;
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                float v0, v1;
                if (q == 0) { v0 = a[j]; v1 = a[j + 1]; }
                else if (q == 1) { v0 = b[j]; v1 = b[j + 1]; }
                else {
#pragma clang fp contract(off)
                    if (q == 2) { v0 = a[j] * b[j]; v1 = a[j + 1] * b[j + 1]; }
                    else {
                        // (a^2 + b^2) / 2: the halving is folded into the
                        // power-of-two scale (exact either way)
                        v0 = a[j] * a[j] + b[j] * b[j];
                        v1 = a[j + 1] * a[j + 1] + b[j + 1] * b[j + 1];
                    }
                }
                _Float16 h0, h1, l0, l1;
                split2(v0, v1, q < 2 ? s1 : (q == 2 ? s2 : s2h), h0, h1, l0, l1);
                ah[j] = h0; ah[j + 1] = h1; al[j] = l0; al[j + 1] = l1;
            }
            const floatx4 hq = mma32(ah, al, g_hi, g_lo);
            const float hv[4] = {hq[0], hq[1], hq[2], hq[3]};
            split4(hv, bh[q], bl[q]);
        }

        // ---- output block kb - 1 is complete: T1 part, unscale, maps
        if (kb >= 1) {
            const int rbk = kb - 1;
            half4 t1_hi, t1_lo;
            split_half8(TL->t1[lane], t1_hi, t1_lo);
            floatx4 v[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const floatx4 t1 = mma16(t1_hi, t1_lo, bh[q], bl[q]);
                const float up = q < 2 ? prev_u1 : prev_u2, uc = q < 2 ? u1 : u2;
#pragma unroll
                for (int i = 0; i < 4; i++) v[q][i] = prev[q][i] * up + t1[i] * uc;
            }
            // centred samples at the outputs: window rows 16rbk + 8 + 4lg + i
            const float* ca;
            const float* cb;
            {
                const int r0 = 8 + 4 * lg;  // < 16 for lg < 2
                const int slot = r0 < 16 ? rbk % kSlots : (rbk + 1) % kSlots;
                const int rr = r0 < 16 ? r0 : r0 - 16;
                ca = sA + slot * kSlotFloats + rr * kRP + 16 * wave + kHalo + ln;
                cb = sB + slot * kSlotFloats + rr * kRP + 16 * wave + kHalo + ln;
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int y = g.ty0 + 16 * rbk + 4 * lg + i;
                if (x < g.w && y < g.h) {
                    const float nrm = inx * inv_norm(y, g.h, taps);
                    const float m1 = v[0][i] * nrm, m2 = v[1][i] * nrm;
                    if constexpr (NQ == 4) {
                        if (do_ssim) {
                            const float s12 = v[2][i] * nrm;
                            const float sq = (v[3][i] * nrm) * 2.0f;
                            const float val = ssim_value(m1, m2, s12, sq);
                            s1m[0] += val;
                            s4m[0] += pow4(val);
                        }
                    }
                    if (do_edge) {
                        const float dd = edge_value(ca[i * kRP], cb[i * kRP], m1, m2);
                        const float art = fmaxf(dd, 0.0f), det = fmaxf(-dd, 0.0f);
                        s1m[1] += art;
                        s4m[1] += pow4(art);
                        s1m[2] += det;
                        s4m[2] += pow4(det);
                    }
                }
            }
        }
        // ---- T0 part of output block kb
        if (kb < nrb) {
            half4 t0_hi, t0_lo;
            split_half8(TL->t01[lane], t0_hi, t0_lo);
#pragma unroll
            for (int q = 0; q < NQ; q++) prev[q] = mma16(t0_hi, t0_lo, bh[q], bl[q]);
            prev_u1 = u1;
            prev_u2 = u2;
        }
        if (more) stage_store<StoreT>(st, g, kb + 1, c, sA, sB);
        __syncthreads();
    }

    // ---- workgroup reduction of the 6 sums
    float v6[6] = {s1m[0], s1m[1], s1m[2], s4m[0], s4m[1], s4m[2]};
#pragma unroll
    for (int m = 0; m < 6; m++) v6[m] = wave_sum(v6[m]);
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
// 4 waves per SIMD: 128 registers per lane, 36.4 KB of LDS per workgroup.
__global__ __launch_bounds__(kBlurThreads, 4)
void blur_strip_kernel(const MapJob* __restrict__ jobs, const int* __restrict__ counts, BlurTaps taps, int segb,
                       float* __restrict__ partials)
{
    __shared__ __attribute__((aligned(16))) float ring[2 * kSlots * kSlotFloats];
    __shared__ float red[32];
    __shared__ ToeplitzLds toe;
    const int total = counts[kCountBlurTiles];
    const int njobs = counts[kCountBlurJobs];
    const int per = (total + kXcds - 1) / kXcds;
    const int xcd = blockIdx.x % kXcds;
    const int slots = gridDim.x / kXcds;
    const int end = min((xcd + 1) * per, total);
    const int first = xcd * per + blockIdx.x / kXcds;
    if (first >= end) return;  // a grid sized for larger batches: nothing to set up
    if (threadIdx.x < 64) toeplitz_to_lds(make_toeplitz(taps, threadIdx.x), threadIdx.x, &toe);
    __syncthreads();
    float* sA = ring;
    float* sB = ring + kSlots * kSlotFloats;
    for (int item = first; item < end; item += slots) {
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
            blur_item<StoreT, 4>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red,
                                 partials + (int64_t)(job.part_start + local) * kPartialStride);
        else
            blur_item<StoreT, 2>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red,
                                 partials + (int64_t)(job.part_start + local) * kPartialStride);
    }
}

}  // namespace

template <typename StoreT>
hipError_t launch_blur_strip(int segb, const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps,
                             float* partials, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (grid % kXcds || segb < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(blur_strip_kernel<StoreT>, dim3(grid), dim3(kBlurThreads), 0, stream, jobs, counts, taps, segb,
                       partials);
    return hipGetLastError();
}

template hipError_t launch_blur_strip<float>(int, const MapJob*, const int*, int, const BlurTaps&, float*,
                                             hipStream_t);
template hipError_t launch_blur_strip<__half>(int, const MapJob*, const int*, int, const BlurTaps&, float*,
                                              hipStream_t);

}  // namespace vship
