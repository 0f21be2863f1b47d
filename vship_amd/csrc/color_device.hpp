// color_device.hpp - per-pixel colour arithmetic of the front-end kernel.
//
// Float restatements of the reference's colour steps; every function cites the
// reference file:line it follows. Divisions by constants are done as
// multiplications by the reciprocal and powers/cube roots use the hardware
// log2/exp2 with one Newton step for the cube root: each differs from the
// reference's IEEE-division / libm result by a few ulp, far below the
// |delta score| <= 0.005 bar, and both images of a pair go through the same
// code, so identical inputs still give identical planes.
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

enum SampleKind : int32_t { kSampleU8 = 0, kSampleU16 = 1, kSampleHalf = 2, kSampleFloat = 3 };

// reference src/gpuColorToLinear/anyDepthToFloat.hpp:24-68 (PickValue):
// little endian, integer samples above 8 bits masked to their bit depth.
__device__ inline float load_sample(const uint8_t* p, int kind, uint32_t mask)
{
    switch (kind) {
    case kSampleU8:
        return (float)p[0];
    case kSampleU16: {
        const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8);
        return (float)(v & mask);
    }
    case kSampleHalf: {
        const uint16_t v = (uint16_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8));
        return __half2float(__ushort_as_half(v));
    }
    default: {
        const uint32_t v = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
        return __uint_as_float(v);
    }
    }
}

// Range normalisation constants of one plane, reference
// src/gpuColorToLinear/rangeToFull.hpp:24-55: v' = (v * scale - offset) * gain
// full range:    scale = 1/(2^b - 1), offset = 0.5 (YUV chroma) or 0, gain = 1
// limited range: scale = 2^-(b-8) (x256 for half/float), offset 16 / 128,
//                gain 1/219 (luma) or 1/224 (chroma)
struct RangeXform {
    float scale, offset, gain;
};

__host__ __device__ inline RangeXform make_range(int range, int bits, bool chroma)
{
    RangeXform r;
    if (range == 1) {
        r.scale = 1.0f / (float)((1 << bits) - 1);
        r.offset = chroma ? 0.5f : 0.0f;
        r.gain = 1.0f;
    } else {
        r.scale = bits >= 8 ? 1.0f / (float)(1 << (bits - 8)) : 256.0f;
        r.offset = chroma ? 128.0f : 16.0f;
        r.gain = chroma ? 1.0f / 224.0f : 1.0f / 219.0f;
    }
    return r;
}

__device__ inline float apply_range(float v, const RangeXform& r)
{
    return (v * r.scale - r.offset) * r.gain;
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

// x^g for x > 0 from the hardware log2/exp2 (v_log_f32, v_exp_f32).
__device__ inline float pow_pos(float x, float g)
{
    return x > 0.0f ? __builtin_amdgcn_exp2f(g * __builtin_amdgcn_logf(x)) : 0.0f;
}

__device__ inline float pow_signed(float x, float g)
{
    return x < 0.0f ? -pow_pos(-x, g) : pow_pos(x, g);
}

// Cube root of x >= 0: exp2(log2(x)/3) refined by one Newton step
// y <- (2y + x/y^2)/3.
__device__ inline float cbrt_pos(float x)
{
    if (!(x > 0.0f)) return 0.0f;
    float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
    const float r = __builtin_amdgcn_rcpf(y * y);
    y = (2.0f * y + x * r) * (1.0f / 3.0f);
    return y;
}

// Per-component EOTFs, reference src/gpuColorToLinear/transferToLinear.hpp:
// sRGB :73-88, BT.709 :110-118 (pure 2.4 power), BT470_M :128-131,
// BT470_BG/BT601 :133-141, ST240_M :143-158, ST428 :162-165, PQ :169-182.
__device__ inline float eotf(float a, int trc)
{
    switch (trc) {
    case 13: /* sRGB */
        if (a < 0.0f) return (a < -0.04045f) ? -pow_pos((-a + 0.055f) * (1.0f / 1.055f), 2.4f) : a * (1.0f / 12.92f);
        return (a > 0.04045f) ? pow_pos((a + 0.055f) * (1.0f / 1.055f), 2.4f) : a * (1.0f / 12.92f);
    case 1: /* BT709 */ return pow_signed(a, 2.4f);
    case 4: /* BT470_M */ return pow_signed(a, 2.2f);
    case 5: /* BT470_BG */
    case 6: /* BT601 */ return pow_signed(a, 2.8f);
    case 7: /* ST240_M */
        if (a >= 0.0f) return (a < 0.0228f) ? 4.0f * a : 1.1115f * pow_pos(a, 0.45f) - 0.1115f;
        return (a > -0.0228f) ? 4.0f * a : -1.1115f * pow_pos(-a, 0.45f) + 0.1115f;
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

// Y'CbCr (y = Y, u = Cb, v = Cr) -> linear RGB, reference
// src/gpuColorToLinear/YUVToLinRGB.hpp:157-163 (gamma-light matrix, EOTF,
// linear-light matrix; per-matrix steps :7-155).
__device__ inline void to_linear_rgb(float y, float u, float v, int matrix, int trc, float& R, float& G, float& B)
{
    float r, g, b;
    switch (matrix) {
    case 1: case 5: case 6: case 9: {
        float kr, kb;
        if (matrix == 1) { kr = 0.2126f; kb = 0.0722f; }
        else if (matrix == 9) { kr = 0.2627f; kb = 0.0593f; }
        else { kr = 0.299f; kb = 0.114f; }
        const float kg = 1.0f - kb - kr;
        const float us = 0.5f / (1.0f - kb), vs = 0.5f / (1.0f - kr);
        r = (v + y * vs) * (1.0f / (vs * kr + 0.5f));
        b = (u + y * us) * (1.0f / (us * kb + 0.5f));
        g = (y - kr * r - kb * b) * (1.0f / kg);
        break;
    }
    case 10: { /* BT2020 CL: {R', Y', B'} */
        const float kr = 0.2627f, kb = 0.0593f;
        const float us = 0.5f / (1.0f - kb), vs = 0.5f / (1.0f - kr);
        r = (v + y * vs) * (1.0f / (vs * kr + 0.5f));
        b = (u + y * us) * (1.0f / (us * kb + 0.5f));
        g = y;
        break;
    }
    case 14: { /* ICtCp -> L'M'S' */
        r = fmaf(u, 0.0086090370379327566f, fmaf(v, 0.11102962500302595655f, y));
        g = fmaf(u, -0.01721807407586551f, fmaf(v, -0.22205925000605192f, r));
        b = fmaf(r, 33.025782688766114f, fmaf(g, -32.025782688766114f, -v * 7.543278084714549f));
        break;
    }
    default: /* RGB */
        r = y; g = u; b = v;
        break;
    }
    if (trc == 18) {
        eotf_hlg(r, g, b);
    } else {
        r = eotf(r, trc);
        g = eotf(g, trc);
        b = eotf(b, trc);
    }
    if (matrix == 10) {
        const float kr = 0.2627f, kb = 0.0593f, kg = 1.0f - kr - kb;
        g = (g - kr * r - kb * b) * (1.0f / kg);
    } else if (matrix == 14) {
        const float l = r, m = g, s = b;
        r = fmaf(l, 3.4366066943330784267f, fmaf(m, -2.5064521186562698975f, s * 0.069845424323191470954f));
        g = fmaf(r, -0.23026480071281397f, fmaf(m, 1.4064527541936944f, -s * 0.2083538389058436f));
        b = fmaf(r, -0.02684381778741865f, fmaf(g, -0.08378524945770065f, s * 1.106879231557486f));
    }
    R = r; G = g; B = b;
}

// Linear RGB -> positive XYB, reference src/ssimu2/makeXYB.hpp:6-52. The last
// three additions use double constants like the reference (:44-46).
struct XYB {
    float x, y, b;
};

__device__ inline XYB to_xyb(float r, float g, float b, bool need_b)
{
    const float bias = 0.0037930734f;
    const float cbias = -0.1559542025327239f;
    float l = fmaf(0.30f, r, fmaf(0.622f, g, fmaf(0.078f, b, bias)));
    float m = fmaf(0.23f, r, fmaf(0.692f, g, fmaf(0.078f, b, bias)));
    l = cbrt_pos(l) + cbias;
    m = cbrt_pos(m) + cbias;
    XYB o;
    const float X = 0.5f * (l - m);
    const float Y = X + m;
    o.x = (float)((double)(X * 14.0f) + 0.42);
    o.y = (float)((double)Y + 0.01);
    if (need_b) {
        float s = fmaf(0.24342269f, r, fmaf(0.20476745f, g, fmaf(0.55180986f, b, bias)));
        s = cbrt_pos(s) + cbias;
        o.b = (float)((double)(s - Y) + 0.55);
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
