// color_device.hpp - per-pixel colour arithmetic of the front-end kernel.
//
// Float restatements of the reference's colour steps; every function cites the
// reference file:line it follows. Divisions by constants are done as
// multiplications by the reciprocal and powers/cube roots use the hardware
// log2/exp2: each differs from the reference's IEEE-division / libm result by
// a few ulp, far below the |delta score| <= 0.005 bar, and both images of a
// pair go through the same code, so identical inputs still give identical
// planes.
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

enum SampleKind : int32_t { kSampleU8 = 0, kSampleU16 = 1, kSampleHalf = 2, kSampleFloat = 3 };

// Range normalisation of one plane, reference
// src/gpuColorToLinear/rangeToFull.hpp:24-55: v' = (v * scale - offset) * gain
// full range:    scale = 1/(2^b - 1), offset = 0.5 (YUV chroma) or 0, gain = 1
// limited range: scale = 2^-(b-8) (x256 for half/float), offset 16 / 128,
//                gain 1/219 (luma) or 1/224 (chroma)
// folded on the host (in double) into one fused multiply-add
// v' = v * mul + add with mul = scale * gain, add = -offset * gain.
struct RangeXform {
    float mul, add;
};

inline RangeXform make_range(int range, int bits, bool chroma)
{
    double scale, offset, gain;
    if (range == 1) {
        scale = 1.0 / (double)((1 << bits) - 1);
        offset = chroma ? 0.5 : 0.0;
        gain = 1.0;
    } else {
        scale = bits >= 8 ? 1.0 / (double)(1 << (bits - 8)) : 256.0;
        offset = chroma ? 128.0 : 16.0;
        gain = chroma ? 1.0 / 224.0 : 1.0 / 219.0;
    }
    RangeXform r;
    r.mul = (float)(scale * gain);
    r.add = (float)(-offset * gain);
    return r;
}

__device__ inline float apply_range(float v, const RangeXform& r)
{
    return fmaf(v, r.mul, r.add);
}

// Cubic Hermite segment with Catmull-Rom tangents, reference
// src/gpuColorToLinear/resize.hpp:5-24 and chromaUpsample.hpp:37,49.
__device__ inline float catmull_rom(float pm1, float p0, float p1, float p2, float t)
{
    const float m0 = (p1 - pm1) * 0.5f;
    const float m1 = (p2 - p0) * 0.5f;
    const float v1 = 2.0f * p0 + m0 - 2.0f * p1 + m1;
    const float v2 = -3.0f * p0 + 3.0f * p1 - 2.0f * m0 - m1;
    float r = v1;
    r = r * t + v2;
    r = r * t + m0;
    r = r * t + p0;
    return r;
}

// x^g (g > 0) for x >= 0 from the hardware log2/exp2 (v_log_f32, v_exp_f32),
// without a branch: log2(+0) = -inf and exp2(-inf) = +0, so x = 0 gives 0;
// negative and NaN inputs are taken as 0 by the max.
__device__ inline float pow_pos(float x, float g)
{
    return __builtin_amdgcn_exp2f(g * __builtin_amdgcn_logf(fmaxf(x, 0.0f)));
}

// sign(x) * |x|^g, branch-free (the magnitude takes x's sign bit).
__device__ inline float pow_signed(float x, float g)
{
    return copysignf(__builtin_amdgcn_exp2f(g * __builtin_amdgcn_logf(fabsf(x))), x);
}

// Cube root of x >= 0 as exp2(log2(x) / 3) from the hardware log2/exp2
// (each about 1 ulp): relative error of a few 1e-7 over the range the opsin
// mix produces, far below what the score can resolve. Non-positive and NaN
// inputs give 0 (the clamp of makeXYB.hpp:36-38), branch-free as in pow_pos.
__device__ inline float cbrt_pos(float x)
{
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fmaxf(x, 0.0f)) * (1.0f / 3.0f));
}

// Per-component EOTFs, reference src/gpuColorToLinear/transferToLinear.hpp:
// sRGB :73-88, BT.709 :110-118 (pure 2.4 power), BT470_M :128-131,
// BT470_BG/BT601 :133-141, ST240_M :143-158, ST428 :162-165, PQ :169-182.
__device__ inline float eotf(float a, int trc)
{
    switch (trc) {
    case 13: { /* sRGB, sign-symmetric; both branches evaluated, one selected */
        const float ax = fabsf(a);
        const float pw = pow_pos((ax + 0.055f) * (1.0f / 1.055f), 2.4f);
        return copysignf(ax > 0.04045f ? pw : ax * (1.0f / 12.92f), a);
    }
    case 1: /* BT709 */ return pow_signed(a, 2.4f);
    case 4: /* BT470_M */ return pow_signed(a, 2.2f);
    case 5: /* BT470_BG */
    case 6: /* BT601 */ return pow_signed(a, 2.8f);
    case 7: /* ST240_M */
    {
        const float ax = fabsf(a);
        const float pw = fmaf(1.1115f, pow_pos(ax, 0.45f), -0.1115f);
        return copysignf(ax < 0.0228f ? 4.0f * ax : pw, a);
    }
    case 17: /* ST428 */ return pow_signed(a, 2.6f) * (float)(52.37 / 48.0);
    case 16: { /* PQ */
        const float m = 78.84375f, n = 0.1593017578125f;
        const float c1 = 0.8359375f, c2 = 18.8515625f, c3 = 18.6875f;
        a = fmaxf(a, 0.0f);
        a = pow_pos(a, 1.0f / m);
        return 100.0f * pow_pos(fmaxf(a - c1, 0.0f) / (c2 - c3 * a), 1.0f / n);
    }
    default: /* Linear */ return a;
    }
}

// reference transferToLinear.hpp:24-51 (HLG on the triple)
__device__ inline void eotf_hlg(float& r, float& g, float& b)
{
    const float a = 0.17883277f, bb = 1.0f - 4.0f * a, c = 0.5f - a * -0.3350097945111627f;
    float v[3] = {r, g, b};
#pragma unroll
    for (int i = 0; i < 3; i++)
        v[i] = (v[i] <= 0.5f) ? v[i] * v[i] * (1.0f / 3.0f) : (__expf((v[i] - c) * (1.0f / a)) + bb) * (1.0f / 12.0f);
    const float ys = 0.2627f * v[0] + 0.6780f * v[1] + 0.0593f * v[2];
    const float k = pow_pos(ys, 0.2f);
    r = v[0] * k; g = v[1] * k; b = v[2] * k;
}

// 3x3 row-major matrix times (x, y, z), in place; each row as the chain
// fma(m2, z, fma(m1, y, m0 x)), so that a row with unit m0 and zero terms is
// the same float as its chain without them.
__device__ inline void mat3_apply(const float m[9], float& x, float& y, float& z)
{
    const float a = fmaf(m[2], z, fmaf(m[1], y, m[0] * x));
    const float b = fmaf(m[5], z, fmaf(m[4], y, m[3] * x));
    const float c = fmaf(m[8], z, fmaf(m[7], y, m[6] * x));
    x = a; y = b; z = c;
}

// Linear RGB -> positive XYB, reference src/ssimu2/makeXYB.hpp:6-52, with
// the cube-root offset c (-0.15595...) folded into the constants: for cube
// roots l, m, s of the three opsin mixes,
//   X' = 14 * ((l + c) - (m + c)) / 2 + 0.42 = 7 (l - m) + 0.42
//   Y' = ((l + c) + (m + c)) / 2 + 0.01      = (l + m) / 2 + (c + 0.01)
//   B' = (s + c) - (Y' - 0.01) + 0.55        = s - (l + m) / 2 + 0.55
// (the offset cancels in X' and B'). Fewer roundings than the reference's
// float steps (which add the last three offsets in double, :44-46): the
// stored values differ from it by about one ulp.
struct XYB {
    float x, y, b;
};

__device__ inline XYB to_xyb(float r, float g, float b, bool need_b)
{
    const float bias = 0.0037930734f;
    constexpr float kYOffset = -0.1559542025327239f + 0.01f;
    const float l = cbrt_pos(fmaf(0.30f, r, fmaf(0.622f, g, fmaf(0.078f, b, bias))));
    const float m = cbrt_pos(fmaf(0.23f, r, fmaf(0.692f, g, fmaf(0.078f, b, bias))));
    const float h = l + m;
    XYB o;
    o.x = fmaf(l - m, 7.0f, 0.42f);
    o.y = fmaf(h, 0.5f, kYOffset);
    if (need_b) {
        const float s = cbrt_pos(fmaf(0.24342269f, r, fmaf(0.20476745f, g, fmaf(0.55180986f, b, bias))));
        o.b = fmaf(h, -0.5f, s) + 0.55f;
    } else {
        o.b = 0.0f;
    }
    return o;
}

template <typename T> __device__ inline T store_cast(float v);
template <> __device__ inline float store_cast<float>(float v) { return v; }
template <> __device__ inline __half store_cast<__half>(float v) { return __float2half_rn(v); }

template <typename T> __device__ inline float load_cast(T v);
template <> __device__ inline float load_cast<float>(float v) { return v; }
template <> __device__ inline float load_cast<__half>(__half v) { return __half2float(v); }

}  // namespace vship
