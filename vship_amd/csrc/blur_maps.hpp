// blur_maps.hpp - per-pixel SSIM and edge-difference arithmetic shared by the
// blur/map kernels. Compiled with FP contraction off so that identical inputs
// give bit-identical numerator and denominator (score exactly 100).
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

struct MapSums {
    float s1[3];  // sum of ssim, artifact, detail-loss
    float s4[3];  // sum of their 4th powers
};

// SSIM map value, reference src/ssimu2/score.hpp:266-275, with
// s11 + s22 = G(a^2 + b^2) blurred directly instead of G((a+b)^2) - 2 G(ab).
// All inputs may be centred on one constant (m = mu - c, s12 = G((a-c)(b-c)),
// sq = G((a-c)^2 + (b-c)^2)): the formula only uses mu1 - mu2 and the
// (co)variances, which centring leaves unchanged.
// For a == b: sq == 2 * s12 and m11 + m22 == 2 * m12 exactly, so num_s and
// den_s are the same float and the map is exactly 0.
__device__ inline float ssim_value(float m1, float m2, float s12, float sq)
{
#pragma clang fp contract(off)
    const float m11 = m1 * m1;
    const float m22 = m2 * m2;
    const float m12 = m1 * m2;
    const float md = m1 - m2;
    const float num_m = 1.0f - md * md;
    const float num_s = 2.0f * (s12 - m12) + kC2;
    const float den_s = (sq - (m11 + m22)) + kC2;
    return fmaxf(1.0f - (num_m * num_s) / den_s, 0.0f);
}

// Edge difference d = (|b - mu2| + 1) / (|a - mu1| + 1) - 1, artifact =
// max(d, 0), detail loss = max(-d, 0) (reference score.hpp:276-281); a, b,
// mu1, mu2 may all be centred on the same constant.
__device__ inline float edge_value(float a, float b, float m1, float m2)
{
#pragma clang fp contract(off)
    return (fabsf(b - m2) + 1.0f) / (fabsf(a - m1) + 1.0f) - 1.0f;
}

__device__ inline float pow4(float x)
{
#pragma clang fp contract(off)
    const float y = x * x;
    return y * y;
}

}  // namespace vship
