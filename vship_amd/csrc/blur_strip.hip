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
// about 22 significant bits. Every sample of an item is centred on the
// item's first sample and scaled by one power of two (fixed, or from a bound
// over the window when the fixed one would leave FP16's range), so hi and lo
// stay in range for all four quantities and the two vertical products of an
// output block share one accumulator.
//
// Identity: the blurred quantities are a, b, a*b and (a - b)^2. For a == b
// every operation on a and b is the same and (a - b)^2 is 0, so the SSIM
// variance and the edge difference are 0 and so are the maps.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "blur_maps.hpp"
#include "kernels.hpp"

#include <type_traits>

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
constexpr int kMaxSegb = 16;                  // items of at most 16 row blocks (256 rows)
constexpr int kFirstScaleExp = 6;             // first attempt's scale 2^6 ...
constexpr float kFirstScaleMax = 96.0f;       // ... and the largest |A|, |B| it accepts
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

// Element i of a plane as a 32-bit byte offset from the plane pointer (an
// image's stored planes hold fewer than 2^30 elements, so the offset fits):
// the load takes the pointer as its scalar base and the offset as one lane
// register, with no per-lane 64-bit address arithmetic.
template <typename T>
__device__ inline T global_at(const void* p, uint32_t i)
{
    return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) uint8_t*)p +
                                                         i * (uint32_t)sizeof(T));
}

// lo = fp16(x - hi) of two values whose packed FP16 hi parts are hh: x - hi
// is exact in FP32 (hi is x truncated to 11 bits) and rounded once to FP16.
// one is 1.0 hidden from the optimiser: fma(x, one, -hi) is then selected as
// one mixed-precision FMA per value (hi read as FP16 directly) plus one pack,
// where x - hi would cost a conversion and a subtraction per value. The
// sequence is plain compiler code on purpose, so that the compiler inserts
// the wait states an MFMA needs before it reads a VALU result.
__device__ inline uint32_t lo_pair(float x0, float x1, fp16x2 hh, float one)
{
    typedef _Float16 half2v __attribute__((ext_vector_type(2)));
    half2v l;
    l[0] = (_Float16)fmaf(x0, one, -(float)hh[0]);
    l[1] = (_Float16)fmaf(x1, one, -(float)hh[1]);
    return __builtin_bit_cast(uint32_t, l);
}

// Four accumulator values as hi + lo FP16, hi rounded toward zero.
__device__ inline void split4(const float (&x)[4], float one, half4& hi, half4& lo)
{
    const fp16x2 h01 = __builtin_amdgcn_cvt_pkrtz(x[0], x[1]);
    const fp16x2 h23 = __builtin_amdgcn_cvt_pkrtz(x[2], x[3]);
    hi[0] = (_Float16)h01[0];
    hi[1] = (_Float16)h01[1];
    hi[2] = (_Float16)h23[0];
    hi[3] = (_Float16)h23[1];
    typedef _Float16 half2v __attribute__((ext_vector_type(2)));
    const half2v l01 = __builtin_bit_cast(half2v, lo_pair(x[0], x[1], h01, one));
    const half2v l23 = __builtin_bit_cast(half2v, lo_pair(x[2], x[3], h23, one));
    lo[0] = l01[0];
    lo[1] = l01[1];
    lo[2] = l23[0];
    lo[3] = l23[1];
}

__device__ inline floatx4 mma32(const half8& ah, const half8& al, const half8& bh, const half8& bl)
{
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
    return c;
}

// Toeplitz fragments of one lane, looked up in an LDS copy of the host-split
// band of FP16 tap pairs.
//
// Vertical pass: output block rb is Tfull * S with the 16x32 band
// Tfull[m][r] = g[r - m] and S the 32 window rows of blocks rb and rb+1
// (16 + 16). A 16x16x32 product contracts over k = 8*(lane/16) + j; the
// horizontal results hold rows 4*(lane/16) + i of their block in each lane,
// so k is mapped to window row r(k) = 4*(lane/16) + j for j < 4 (block rb)
// and 16 + 4*(lane/16) + j - 4 for j >= 4 (block rb + 1). The B operand is
// then the lane's four results of block rb followed by its four of block
// rb + 1, straight from registers, and the A operand is Tfull with its
// columns in that order (tv; tw holds the order with the two blocks'
// halves swapped, for a register tuple whose halves alternate).
struct Toeplitz {
    half8 g_hi, g_lo;    // horizontal B operand: k = 8*(lane/16)+j, n = lane%16, g[k - n]
    half8 tv_hi, tv_lo;  // vertical A operand: m = lane%16, g[r(k) - m]
    half8 tw_hi, tw_lo;  // the same with the two blocks' halves of k swapped
};

__device__ inline _Float16 half_bits(uint32_t v)
{
    return __builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
}

// band: the LDS copy of BlurTaps::band (g[k] as hi | lo << 16 at k + 16).
__device__ inline Toeplitz make_toeplitz(const uint32_t* band, int lane)
{
    Toeplitz T;
    const int ln = lane & 15, lg = lane >> 4;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        // band offsets k - ln + 16 with k <= 31 and ln <= 15: all inside 1..47
        const uint32_t gv = band[16 + 8 * lg + j - ln];
        T.g_hi[j] = half_bits(gv);
        T.g_lo[j] = half_bits(gv >> 16);
        const int r = j < 4 ? 4 * lg + j : 16 + 4 * lg + (j - 4);
        const uint32_t tv = band[16 + r - ln];
        T.tv_hi[j] = half_bits(tv);
        T.tv_lo[j] = half_bits(tv >> 16);
        const int rw = j < 4 ? 16 + 4 * lg + j : 4 * lg + (j - 4);
        const uint32_t tw = band[16 + rw - ln];
        T.tw_hi[j] = half_bits(tw);
        T.tw_lo[j] = half_bits(tw >> 16);
    }
    return T;
}

// The fragments of every lane, shared by the workgroup's waves through LDS
// (structure of arrays, one 16-byte vector per lane and fragment, so the
// reads are conflict-free); loaded right before the products that use them,
// which keeps 16 registers free across the item.
struct ToeplitzLds {
    half8 g_hi[64], g_lo[64];
    half8 tv_hi[64], tv_lo[64];
    half8 tw_hi[64], tw_lo[64];
};

__device__ inline void toeplitz_to_lds(const Toeplitz& T, int lane, ToeplitzLds* L)
{
    L->g_hi[lane] = T.g_hi;
    L->g_lo[lane] = T.g_lo;
    L->tv_hi[lane] = T.tv_hi;
    L->tv_lo[lane] = T.tv_lo;
    L->tw_hi[lane] = T.tw_hi;
    L->tw_lo[lane] = T.tw_lo;
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
    typedef uintx2 Vec;  // four FP16 samples
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

// Samples nv.. (nv in 1..3) of a vector that straddles the image's last
// column are replaced by the fill values (the row padding is not defined).
__device__ inline void fill_tail(floatx4& a, floatx4& b, int nv, const floatx4& fa, const floatx4& fb)
{
    a.w = fa.w;
    b.w = fb.w;
    if (nv < 3) { a.z = fa.z; b.z = fb.z; }
    if (nv < 2) { a.y = fa.y; b.y = fb.y; }
}

__device__ inline void fill_tail(uintx2& a, uintx2& b, int nv, const uintx2& fa, const uintx2& fb)
{
    // samples 0 1 2 3 = low and high halves of x, then of y
    a.y = nv < 3 ? fa.y : ((a.y & 0xffffu) | (fa.y & 0xffff0000u));
    b.y = nv < 3 ? fb.y : ((b.y & 0xffffu) | (fb.y & 0xffff0000u));
    if (nv < 2) {
        a.x = (a.x & 0xffffu) | (fa.x & 0xffff0000u);
        b.x = (b.x & 0xffffu) | (fb.x & 0xffff0000u);
    }
}

// Window geometry of one item.
struct StripGeom {
    int w, h, pitch;
    int tx0, ty0;  // first output pixel
};

// Global loads of window block kb (rows ty0 - 8 + 16kb ..) into registers.
// Columns tx0 - 8 + 4v..+3 are a multiple of 4 and plane rows are padded to
// 64 elements, so every vector that starts inside a row lies inside its
// padded row. Samples outside the image take the fill values fa, fb: the
// item's centring sample c (and, for FP16 planes, whose second image is
// stored as its difference from the first, a zero difference), which the
// staging maps to exactly 0, the value the renormalised blur gives samples
// outside the image.
template <typename StoreT>
__device__ inline void stage_load(const void* A, const void* B, const StripGeom& g, int kb, StageRegs<StoreT>& r,
                                  const typename StageRegs<StoreT>::Vec& fa, const typename StageRegs<StoreT>::Vec& fb)
{
    typedef typename StageRegs<StoreT>::Vec Vec;
    const int t = threadIdx.x;
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        const int row = e / kVecPerRow, v = e - row * kVecPerRow;
        const int gy = g.ty0 - kHalo + 16 * kb + row, gx = g.tx0 - kHalo + 4 * v;
        const bool ok = e < kVecPerBlock && gy >= 0 && gy < g.h && gx >= 0 && gx < g.w;
        Vec va = fa, vb = fb;
        if (ok) {
            // vector index inside the plane (< 2^31 elements per plane)
            const uint32_t o = ((uint32_t)gy * (uint32_t)g.pitch + (uint32_t)gx) >> 2;
            va = global_at<Vec>(A, o);
            vb = global_at<Vec>(B, o);
            if (gx + 4 > g.w) fill_tail(va, vb, g.w - gx, fa, fb);
        }
        r.a[it] = va;
        r.b[it] = vb;
    }
}

// Bound on the centred magnitudes |x - c| of a staged window (fill samples
// are c: they add nothing). Any bound within a small factor of the maximum
// serves: it only picks the item's power-of-two scale.
template <typename StoreT>
struct WindowBound;

template <>
struct WindowBound<float> {
    float m = 0.0f;
    __device__ inline explicit WindowBound(float) {}
    __device__ inline void add(const StageRegs<float>& r, float c)
    {
#pragma unroll
        for (int it = 0; it < kStageIters; it++) {
            float av[4], bv[4];
            unpack4(r.a[it], av);
            unpack4(r.b[it], bv);
#pragma unroll
            for (int i = 0; i < 4; i++) m = fmaxf(m, fmaxf(fabsf(av[i] - c), fabsf(bv[i] - c)));
        }
    }
    __device__ inline float value() const { return m; }
};

template <>
struct WindowBound<__half> {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    h2 mx = {0, 0}, mn = {0, 0};  // of a - c and of a + d - c, two samples per lane
    h2 c2;
    __device__ inline explicit WindowBound(float c) { c2 = h2{(_Float16)c, (_Float16)c}; }
    __device__ inline void add(const StageRegs<__half>& r, float)
    {
#pragma unroll
        for (int it = 0; it < kStageIters; it++) {
            const uint32_t aw[2] = {r.a[it].x, r.a[it].y}, dw[2] = {r.b[it].x, r.b[it].y};
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const h2 da = __builtin_bit_cast(h2, aw[k]) - c2;  // packed FP16: a bound needs no more
                const h2 db = da + __builtin_bit_cast(h2, dw[k]);
                mx = __builtin_elementwise_max(mx, __builtin_elementwise_max(da, db));
                mn = __builtin_elementwise_min(mn, __builtin_elementwise_min(da, db));
            }
        }
    }
    __device__ inline float value() const
    {
        return fmaxf(fmaxf((float)mx[0], (float)mx[1]), -fminf((float)mn[0], (float)mn[1]));
    }
};

// Scaled centred values of a staged block into its ring slot: A = (a - c) s
// and B = (b - c) s (0 for fill samples). s is a power of two, so
// A = fma(a, s, -c s) is s * (a - c) rounded once. FP16 planes hold the
// second image as d = b - a: B = A + d s. mag collects the largest |A|, |B|
// written (NaN samples are passed over, as by the window bound).
template <typename StoreT>
__device__ inline void stage_store(const StageRegs<StoreT>& r, int kb, float s, float cs, float* sA, float* sB,
                                   float& mag)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int it = 0; it < kStageIters; it++) {
        const int e = t + it * kBlurThreads;
        if (e >= kVecPerBlock) continue;
        const int row = e / kVecPerRow, v = e - row * kVecPerRow;
        float av[4], bv[4];
        unpack4(r.a[it], av);
        unpack4(r.b[it], bv);
        float va[4], vb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            va[i] = fmaf(av[i], s, -cs);
            vb[i] = (sizeof(StoreT) == 2) ? fmaf(bv[i], s, va[i]) : fmaf(bv[i], s, -cs);
            mag = fmaxf(mag, fmaxf(fabsf(va[i]), fabsf(vb[i])));
        }
        const int off = (kb % kSlots) * kSlotFloats + row * kRP + 4 * v;
        *reinterpret_cast<float4*>(sA + off) = make_float4(va[0], va[1], va[2], va[3]);
        *reinterpret_cast<float4*>(sB + off) = make_float4(vb[0], vb[1], vb[2], vb[3]);
    }
}

// Two values as hi + lo FP16 (hi rounded toward zero, lo = fp16(x - hi)).
__device__ inline void split_pair(float x0, float x1, float one, _Float16& h0, _Float16& h1, _Float16& l0,
                                  _Float16& l1)
{
    typedef _Float16 half2v __attribute__((ext_vector_type(2)));
    const fp16x2 hh = __builtin_amdgcn_cvt_pkrtz(x0, x1);
    h0 = (_Float16)hh[0];
    h1 = (_Float16)hh[1];
    const half2v ll = __builtin_bit_cast(half2v, lo_pair(x0, x1, hh, one));
    l0 = ll[0];
    l1 = ll[1];
}

// One item. Every quantity is blurred in a scaled domain: the window values
// are A = (a - c) s, B = (b - c) s with one power of two s for the whole item.
// The first attempt (kExact false) takes s = 2^6, which keeps |A|, |B| <= 96
// for any window whose centred XYB values stay within 1.5 (all SDR content:
// its X', Y' and B' lie in about [0, 1.3]); the staging checks every value it
// writes, and an item that exceeds 96 anywhere (HDR or float input, or
// inf) reports it and is run again with kExact, whose s comes from a bound
// over the whole window (|A|, |B| < 2^6). Either way |A B| <= 2^13.2 and
// (A - B)^2 <= 2^15.2 < 65504 stay in FP16's range; hi/lo keep 22 significant
// bits down to lo's subnormal floor of 2^-24 / s (2^-24 / s^2 for products),
// absolute errors of 1e-9 and 1e-11 at s = 2^6, which the maps cannot resolve
// against their constants C2 = 9e-4 and 1. The blurred quantities are A, B, A B and (A - B)^2, and
// sigma1^2 + sigma2^2 = 2 sigma12 + var(A - B), var(A - B) =
// E[(A - B)^2] - (mu1 - mu2)^2: for a == b it is 0 exactly and so is the map.
// The maps are evaluated on the scaled
// quantities: the SSIM ratio's numerator and denominator both carry s^2
// exactly (C2 becomes C2 s^2, (mu1 - mu2)^2 is unscaled by s^-2) and the edge
// ratio (|B - M2| + s) / (|A - M1| + s) equals (|b - mu2| + 1) / (|a - mu1| + 1),
// so the maps are those of the unscaled values bit for bit. Outputs outside
// the image have zero normalisers and zero samples, which makes both maps 0
// exactly (the SSIM variance and the edge difference are 0), so no output
// is masked by a branch. Both edge maps are always summed (the
// scorers read only the maps a job's flags select); NQ = 4 jobs add SSIM.
template <typename StoreT, int NQ, bool kExact>
__device__ bool blur_item(const MapJob& job, int tile_x, int tile_y, int segb, const BlurTaps& taps,
                          const ToeplitzLds* TL, float* sA, float* sB, float* red, float* rowinv,
                          float* __restrict__ partial)
{
    typedef typename StageRegs<StoreT>::Vec Vec;
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
    Vec fa, fb;
    if constexpr (sizeof(StoreT) == 4) {
        c = global_ptr<float>(job.a)[c_at];
        fa = Vec{c, c, c, c};
        fb = fa;
    } else {
        const uint32_t cb = global_ptr<unsigned short>(job.a)[c_at];
        c = __half2float(__ushort_as_half((unsigned short)cb));
        fa = Vec{cb | (cb << 16), cb | (cb << 16)};
        fb = Vec{0u, 0u};
    }
    // output row blocks of this item that hold image rows; window blocks = that + 1
    const int nrb = min(segb, (g.h - g.ty0 + 15) / 16);
    const int nwb = nrb + 1;

    // 1 / (in-bounds tap sum) of each output row of the item, 0 below the image
    if (t < 16 * segb) rowinv[t] = rel_norm(g.ty0 + t, g.h, taps);

    StageRegs<StoreT> st;
    int e1 = kFirstScaleExp;
    if constexpr (kExact) {
        // the item's scale from a bound over the whole window: two blocks per
        // step so that their loads are in flight together, in reverse order so
        // that block 0 is the one left in st
        {
            WindowBound<StoreT> bnd(c);
            StageRegs<StoreT> st2;
            for (int kb = nwb - 1; kb > 0; kb -= 2) {
                stage_load<StoreT>(job.a, job.b, g, kb, st2, fa, fb);
                stage_load<StoreT>(job.a, job.b, g, kb - 1, st, fa, fb);
                bnd.add(st2, c);
                bnd.add(st, c);
            }
            if (nwb & 1) {
                stage_load<StoreT>(job.a, job.b, g, 0, st, fa, fb);
                bnd.add(st, c);
            }
            const uint32_t mbits = wave_max_bits(__float_as_uint(bnd.value()));
            uint32_t* redu = reinterpret_cast<uint32_t*>(red);
            if (lane == 0) redu[wave] = mbits;
        }
        __syncthreads();
        const uint32_t* redu = reinterpret_cast<const uint32_t*>(red);
        const float vmax = __uint_as_float(max(max(redu[0], redu[1]), max(redu[2], redu[3])));
        e1 = 6;
        if (vmax > 0.0f) {
            int ex;
            frexpf(vmax, &ex);  // vmax < 2^ex
            e1 = 6 - ex;
        }
        e1 = max(-60, min(e1, 60));
    } else {
        stage_load<StoreT>(job.a, job.b, g, 0, st, fa, fb);
    }
    const float s = ldexpf(1.0f, e1);
    const float cs = c * s;
    const float inv_s2 = ldexpf(1.0f, -2 * e1);
    const float c2s2 = kC2 * ldexpf(1.0f, 2 * e1);
    float mag = 0.0f;

    stage_store<StoreT>(st, 0, s, cs, sA, sB, mag);
    __syncthreads();

    const int x = g.tx0 + 16 * wave + ln;
    const float inx = rel_norm(x, g.w, taps);  // 0 at or beyond w
    // the wave's 16 columns all have their 17 taps in bounds
    const bool cols_inside = g.tx0 + 16 * wave >= kHalo && g.tx0 + 16 * wave + 15 + kHalo < g.w;
    float one = 1.0f;
    asm volatile("" : "+s"(one));  // emits nothing: keeps fma(x, one, -hi) an FMA
    float s1m[3] = {0.f, 0.f, 0.f}, s4m[3] = {0.f, 0.f, 0.f};
    // hi/lo horizontal results of two consecutive window blocks per quantity:
    // block kb sits in elements 4 (kb & 1) .. +3, so the vertical B operand
    // of output block kb - 1 is the register tuple itself, read in the K order
    // of the table for kb's parity (the loop runs two blocks per turn so that
    // the parity is a constant)
    half8 xh[NQ], xl[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        xh[q] = half8{0, 0, 0, 0, 0, 0, 0, 0};
        xl[q] = xh[q];
    }

    auto step = [&](auto parity, int kb) {
        constexpr int P = decltype(parity)::value;
        const bool more = kb + 1 < nwb;
        if (more) stage_load<StoreT>(job.a, job.b, g, kb + 1, st, fa, fb);

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

        // ---- per quantity (A, B, A B, (A - B)^2): hi/lo operand, horizontal
        // product, hi/lo of the results (the vertical B operand)
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
                        const float d0 = a[j] - b[j], d1 = a[j + 1] - b[j + 1];
                        v0 = d0 * d0;
                        v1 = d1 * d1;
                    }
                }
                _Float16 h0, h1, l0, l1;
                split_pair(v0, v1, one, h0, h1, l0, l1);
                ah[j] = h0; ah[j + 1] = h1; al[j] = l0; al[j + 1] = l1;
            }
            const floatx4 hq = mma32(ah, al, g_hi, g_lo);
            const float hv[4] = {hq[0], hq[1], hq[2], hq[3]};
            half4 bh, bl;
            split4(hv, one, bh, bl);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                xh[q][4 * P + j] = bh[j];
                xl[q][4 * P + j] = bl[j];
            }
        }

        // ---- output block kb - 1: vertical product over blocks kb - 1 and kb, maps
        if (kb >= 1) {
            const int rbk = kb - 1;
            floatx4 acc[NQ];
            {
                const half8 tv_hi = P ? TL->tv_hi[lane] : TL->tw_hi[lane];
                const half8 tv_lo = P ? TL->tv_lo[lane] : TL->tw_lo[lane];
#pragma unroll
                for (int q = 0; q < NQ; q++) acc[q] = mma32(tv_hi, tv_lo, xh[q], xl[q]);
            }
            // centred scaled samples at the outputs: window rows 16rbk + 8 + 4lg + i
            const float* ca;
            const float* cb;
            {
                const int r0 = 8 + 4 * lg;  // < 16 for lg < 2
                const int slot = r0 < 16 ? rbk % kSlots : (rbk + 1) % kSlots;
                const int rr = r0 < 16 ? r0 : r0 - 16;
                ca = sA + slot * kSlotFloats + rr * kRP + 16 * wave + kHalo + ln;
                cb = sB + slot * kSlotFloats + rr * kRP + 16 * wave + kHalo + ln;
            }
            // the maps of the lane's 4 outputs; with every tap of the output
            // block in bounds (the wave's columns and the block's 16 rows) the
            // products are the means as they stand (the band sums to 1),
            // elsewhere they take the relative renormalisation
            auto maps = [&](auto inside) {
                constexpr bool kInside = decltype(inside)::value;
                float rnv[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                if constexpr (!kInside) {
                    const float4 rn = *reinterpret_cast<const float4*>(rowinv + 16 * rbk + 4 * lg);
                    rnv[0] = rn.x; rnv[1] = rn.y; rnv[2] = rn.z; rnv[3] = rn.w;
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
#pragma clang fp contract(off)
                    float m1, m2, s12, dsq;
                    if constexpr (kInside) {
                        m1 = acc[0][i];
                        m2 = acc[1][i];
                        s12 = acc[2 % NQ][i];  // NQ = 2 jobs have no SSIM quantities
                        dsq = acc[3 % NQ][i];
                    } else {
                        const float nrm = inx * rnv[i];
                        m1 = acc[0][i] * nrm;
                        m2 = acc[1][i] * nrm;
                        s12 = acc[2 % NQ][i] * nrm;
                        dsq = acc[3 % NQ][i] * nrm;
                    }
                    if constexpr (NQ == 4) {
                        // 1 - num_m num_s / den_s = u + v - u v with
                        // u = (mu1 - mu2)^2, v = var(A - B) / den_s and
                        // den_s = 2 sigma12 + C2 + var(A - B): formed from the
                        // small quantities, not as a difference from 1. One
                        // reciprocal (1 ulp): the map needs no more, and for
                        // var = 0 (a == b) v is 0 whatever it rounds.
                        const float md = m1 - m2, md2 = md * md;
                        const float u = md2 * inv_s2;
                        const float var = dsq - md2;              // E[(A - B)^2] - (mu1 - mu2)^2
                        const float cov = fmaf(-m1, m2, s12);     // sigma12
                        const float num_s = fmaf(2.0f, cov, c2s2);
                        const float v = var * __builtin_amdgcn_rcpf(num_s + var);
                        const float val = fmaxf(fmaf(-u, v, u + v), 0.0f);
                        const float v2 = val * val;
                        s1m[0] += val;
                        s4m[0] = fmaf(v2, v2, s4m[0]);
                    }
                    {
                        // (|B - M2| + s) / (|A - M1| + s) - 1 as (|B - M2| - |A - M1|) / (|A - M1| + s):
                        // the same ratio without the cancellation, and exactly 0 when a == b
                        const float da = fabsf(ca[i * kRP] - m1), db = fabsf(cb[i * kRP] - m2);
                        const float dd = (db - da) * __builtin_amdgcn_rcpf(da + s);
                        const float art = fmaxf(dd, 0.0f), a2 = art * art;
                        const float det = fmaxf(-dd, 0.0f), d2 = det * det;
                        s1m[1] += art;
                        s4m[1] = fmaf(a2, a2, s4m[1]);
                        s1m[2] += det;
                        s4m[2] = fmaf(d2, d2, s4m[2]);
                    }
                }
            };
            const int oy = g.ty0 + 16 * rbk;  // first output row of the block
            if (cols_inside && oy >= kHalo && oy + 15 + kHalo < g.h) maps(std::true_type{});
            else maps(std::false_type{});
        }
        if (more) stage_store<StoreT>(st, kb + 1, s, cs, sA, sB, mag);
        __syncthreads();
    };
    for (int kb = 0; kb < nwb; kb += 2) {
        step(std::integral_constant<int, 0>{}, kb);
        if (kb + 1 < nwb) step(std::integral_constant<int, 1>{}, kb + 1);
    }

    // ---- workgroup reduction of the 6 sums (and of the first attempt's range check)
    float v6[6] = {s1m[0], s1m[1], s1m[2], s4m[0], s4m[1], s4m[2]};
#pragma unroll
    for (int m = 0; m < 6; m++) v6[m] = wave_sum(v6[m]);
    bool over = false;
    if constexpr (!kExact) over = __builtin_amdgcn_ballot_w64(!(mag <= kFirstScaleMax)) != 0;
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < 6; m++) red[wave * 8 + m] = v6[m];
        red[wave * 8 + 6] = over ? 1.0f : 0.0f;
    }
    __syncthreads();
    if constexpr (!kExact) {
        if (red[6] + red[14] + red[22] + red[30] != 0.0f) return false;  // the same for every thread
    }
    if (t < 6) {
        float tot = red[t];
        tot += red[8 + t];
        tot += red[16 + t];
        tot += red[24 + t];
        partial[t] = tot;
    }
    return true;
}

template <typename StoreT>
// 38.6 KB of LDS per workgroup; 3 waves per SIMD (168 registers per lane), no
// spills for either storage. FP32 planes stage twice the registers and would
// spill inside the block loop at 4 waves per SIMD (measured 19% slower per
// pair); FP16 planes run as fast at 4 waves with 23 registers spilled at item
// start (6.8 MB of scratch writes per 1080p pair), so they take 3 as well.
__global__ __launch_bounds__(kBlurThreads, 3)
void blur_strip_kernel(const MapJob* __restrict__ jobs, const int* __restrict__ counts, BlurTaps taps, int segb,
                       float* __restrict__ partials)
{
    __shared__ __attribute__((aligned(16))) float ring[2 * kSlots * kSlotFloats];
    __shared__ float red[32];
    __shared__ __attribute__((aligned(16))) float rowinv[16 * kMaxSegb];
    __shared__ ToeplitzLds toe;
    __shared__ uint32_t band[48];
    const int total = counts[kCountBlurTiles];
    const int njobs = counts[kCountBlurJobs];
    const int per = (total + kXcds - 1) / kXcds;
    const int xcd = blockIdx.x % kXcds;
    const int slots = gridDim.x / kXcds;
    const int end = min((xcd + 1) * per, total);
    const int first = xcd * per + blockIdx.x / kXcds;
    if (first >= end) return;  // a grid sized for larger batches: nothing to set up
    if (threadIdx.x < 48) band[threadIdx.x] = taps.band[threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 64) toeplitz_to_lds(make_toeplitz(band, threadIdx.x), threadIdx.x, &toe);
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
        float* part = partials + (int64_t)(job.part_start + local) * kPartialStride;
        if (job.flags & kMapSsim) {
            if (!blur_item<StoreT, 4, false>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red, rowinv, part)) {
                __syncthreads();
                blur_item<StoreT, 4, true>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red, rowinv, part);
            }
        } else {
            if (!blur_item<StoreT, 2, false>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red, rowinv, part)) {
                __syncthreads();
                blur_item<StoreT, 2, true>(job, tile_x, tile_y, segb, taps, &toe, sA, sB, red, rowinv, part);
            }
        }
    }
}

}  // namespace

template <typename StoreT>
hipError_t launch_blur_strip(int segb, const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps,
                             float* partials, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    if (grid % kXcds || segb < 1 || segb > kMaxSegb) return hipErrorInvalidValue;
    hipLaunchKernelGGL(blur_strip_kernel<StoreT>, dim3(grid), dim3(kBlurThreads), 0, stream, jobs, counts, taps, segb,
                       partials);
    return hipGetLastError();
}

template hipError_t launch_blur_strip<float>(int, const MapJob*, const int*, int, const BlurTaps&, float*,
                                             hipStream_t);
template hipError_t launch_blur_strip<__half>(int, const MapJob*, const int*, int, const BlurTaps&, float*,
                                              hipStream_t);

}  // namespace vship
