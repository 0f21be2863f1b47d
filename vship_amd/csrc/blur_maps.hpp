// blur_maps.hpp - per-pixel SSIM and edge-difference arithmetic shared by the
// blur/map kernels. Compiled with FP contraction off so that identical inputs
// give bit-identical numerator and denominator (score exactly 100).
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

// 1 / (sum of the in-bounds taps) at position x of an axis of length w
// (reference gaussianblur.hpp:52-54,68-70: integral[end] - integral[beg]);
// 0 outside the axis. Away from the borders all 17 taps are in bounds and
// the value is the precomputed taps.inv_full (the same float expression).
__device__ inline float inv_norm(int x, int w, const BlurTaps& taps)
{
    if (x >= kHalo && x + kHalo < w) return taps.inv_full;
    if (x >= w) return 0.0f;
    const int beg = x < kHalo ? kHalo - x : 0;
    const int end = (w - x + kHalo) < 17 ? (w - x + kHalo) : 17;
    return 1.0f / (taps.integral[end] - taps.integral[beg]);
}

// The same renormalisation for blurs with the normalised taps (BlurTaps::band,
// which sum to 1): (sum of all 17 taps) / (sum of the in-bounds taps), exactly
// 1 where all taps are in bounds, 0 outside the axis.
__device__ inline float rel_norm(int x, int w, const BlurTaps& taps)
{
    if (x >= kHalo && x + kHalo < w) return 1.0f;
    if (x >= w) return 0.0f;
    const int beg = x < kHalo ? kHalo - x : 0;
    const int end = (w - x + kHalo) < 17 ? (w - x + kHalo) : 17;
    return (taps.integral[17] - taps.integral[0]) / (taps.integral[end] - taps.integral[beg]);
}

// n / d from the hardware reciprocal and one Newton correction: within an
// ulp of IEEE division, and exactly 1 when n == d (the residual n - d * q is
// exact and the correction lands on 1), which keeps identical inputs at a
// score of exactly 100.
__device__ inline float div_nr(float n, float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    const float q = n * r;
    const float e = fmaf(-d, q, n);
    return fmaf(e, r, q);
}

struct MapSums {
    float s1[3];  // sum of ssim, artifact, detail-loss
    float s4[3];  // sum of their 4th powers
};

// SSIM map value 1 - num_m num_s / den_s, reference src/ssimu2/score.hpp:266-275,
// with s11 + s22 = G(a^2 + b^2) blurred directly instead of
// G((a+b)^2) - 2 G(ab). All inputs may be centred on one constant
// (m = mu - c, s12 = G((a-c)(b-c)), sq = G((a-c)^2 + (b-c)^2)): the formula
// only uses mu1 - mu2 and the (co)variances, which centring leaves unchanged.
// Evaluated as 1 - (1 - u)(1 - v) = u + v - u v with u = (mu1 - mu2)^2 and
// v = var(a - b) / (s11 + s22 + C2), var(a - b) = s11 + s22 - 2 s12: the same
// value, formed from the two small quantities instead of a difference from 1
// that float resolves only to 2^-24 (near-identical images have maps of that
// order). For a == b: sq == 2 * s12 and m11 + m22 == 2 * m12 exactly, so u and
// var(a - b) are 0 and so is the map.
__device__ inline float ssim_value(float m1, float m2, float s12, float sq)
{
#pragma clang fp contract(off)
    const float m11 = m1 * m1;
    const float m22 = m2 * m2;
    const float m12 = m1 * m2;
    const float md = m1 - m2;
    const float u = md * md;
    const float s_sum = sq - (m11 + m22);                // s11 + s22
    const float var = s_sum - 2.0f * (s12 - m12);        // s11 + s22 - 2 s12
    const float v = div_nr(var, s_sum + kC2);
    return fmaxf(fmaf(-u, v, u + v), 0.0f);
}

// Edge difference d = (|b - mu2| + 1) / (|a - mu1| + 1) - 1, artifact =
// max(d, 0), detail loss = max(-d, 0) (reference score.hpp:276-281); a, b,
// mu1, mu2 may all be centred on the same constant.
__device__ inline float edge_value(float a, float b, float m1, float m2)
{
#pragma clang fp contract(off)
    return div_nr(fabsf(b - m2) + 1.0f, fabsf(a - m1) + 1.0f) - 1.0f;
}

__device__ inline float pow4(float x)
{
#pragma clang fp contract(off)
    const float y = x * x;
    return y * y;
}

}  // namespace vship
