/* ref_float.c - CPU emulation of the reference's own FP32 arithmetic for the
 * SSIMULACRA2 maps and score, to explain where the reference libvship differs
 * from the exact formula that ssimu2_oracle.c evaluates in FP64.
 *
 * TEST INFRASTRUCTURE ONLY, like ssimu2_oracle.c: only tests/ load it.
 *
 * The operation order and rounding follow the reference's sources and the
 * gfx950 assembly of its planescale_map_Kernel (built here by hipcc -O3 from
 * src/ssimu2/score.hpp; every instantiation contracts the same way):
 *   - pyramid: ((s00 + s10) + s01) + s11, then * 0.25f (downsample.hpp:13-17);
 *   - XYB: fmaf opsin rows, cbrtf of the non-negative part + float bias,
 *     X = 0.5f (l - m), Y = X + m (uses the new X), then the offsets 0.55,
 *     0.42, 0.01 added in double and rounded back (makeXYB.hpp:6-52);
 *   - blur: per axis a sequential fmaf chain over the 17 taps in index order
 *     from 0, zero outside the image, then one correctly rounded division by
 *     the float difference of the float tap prefix sums (gaussianblur.hpp:3-77);
 *   - SSIM: s12 = G(fl(a b)), S = G(fl(fl(a + b)^2)), s11 + s22 = fmaf(-2, s12, S),
 *     num_m = fmaf(d, -d, 1), num_s = fmaf(fmaf(-m1, m2, s12), 2, C2),
 *     den = fmaf(-m2, m2, fmaf(-m1, m1, s11 + s22)) + C2, 1 - num_m num_s / den
 *     (score.hpp:232-275; the fused forms are the compiler's contraction);
 *   - edge: (|b - m2| + 1) / (|a - m1| + 1) - 1 (score.hpp:276-281);
 *   - x^4 = (x x)(x x) (util/float3operations.hpp:161-164);
 *   - per 16x16 block: the LDS pointer-jumping tree over thread ids
 *     ty*16 + tx, divided by the float W*H of the scale (score.hpp:284-320);
 *     then sumreduce passes of 1024-wide trees over the row-major block
 *     index until one value is left (score.hpp:145-188,361-378);
 *   - host: L4 = sqrtf(sqrtf(x)), the 108-term fmaf chain over float weights,
 *     then the polynomial and pow in double, returned as float
 *     (score.hpp:422-460).
 * The colour front-end (decode, upsampling, matrix, transfer) is the FP64
 * oracle's, rounded to float: the reference evaluates it in float with the
 * device libm, which this emulation does not reproduce bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/VshipColor.h"

#define ORACLE_EXPORT __attribute__((visibility("default")))

int oracle_linear_rgb(const Vship_Colorspace_t* cs, const uint8_t* const planes[3], const int64_t stride[3],
                      double* out);
void oracle_final_dims(const Vship_Colorspace_t* cs, int64_t* w, int64_t* h);
int oracle_map_active(int plane, int scale, int map);
double oracle_weight(int i);

static int64_t half_up(int64_t n) { return (n - 1) / 2 + 1; }

/* reference src/ssimu2/gaussianblur.hpp:7-14 with SIGMA 1.5f, TAU in double
 * (src/util/preprocessor.hpp:81-84) */
static void taps_float(float g[17], float integral[18])
{
    const float sigma = 1.5f;
    integral[0] = 0.0f;
    for (int i = 0; i < 17; i++) {
        const float e = expf((float)(-(8 - i) * (8 - i)) / (2 * sigma * sigma));
        g[i] = (float)((double)e / sqrt(6.28318530718 * sigma * sigma));
        integral[i + 1] = integral[i] + g[i];
    }
}

/* reference src/ssimu2/downsample.hpp:3-19 */
static void downsample_f(const float* src, int64_t w, int64_t h, float* dst)
{
    const int64_t nw = half_up(w), nh = half_up(h);
    for (int64_t y = 0; y < nh; y++) {
        const int64_t y0 = 2 * y < h - 1 ? 2 * y : h - 1, y1 = 2 * y + 1 < h - 1 ? 2 * y + 1 : h - 1;
        for (int64_t x = 0; x < nw; x++) {
            const int64_t x0 = 2 * x < w - 1 ? 2 * x : w - 1, x1 = 2 * x + 1 < w - 1 ? 2 * x + 1 : w - 1;
            float d = src[y0 * w + x0];
            d += src[y1 * w + x0];
            d += src[y0 * w + x1];
            d += src[y1 * w + x1];
            dst[y * nw + x] = d * 0.25f;
        }
    }
}

/* reference src/ssimu2/makeXYB.hpp:6-52 */
static void xyb_f(float* r, float* g, float* b, int64_t n)
{
    const float bias = 0.0037930734f, cbias = -0.1559542025327239f;
    for (int64_t i = 0; i < n; i++) {
        float l = fmaf(0.30f, r[i], fmaf(0.622f, g[i], fmaf(0.078f, b[i], bias)));
        float m = fmaf(0.23f, r[i], fmaf(0.692f, g[i], fmaf(0.078f, b[i], bias)));
        float s = fmaf(0.24342269f, r[i], fmaf(0.20476745f, g[i], fmaf(0.55180986f, b[i], bias)));
        l = cbrtf(l * (float)(l >= 0)) + cbias;
        m = cbrtf(m * (float)(m >= 0)) + cbias;
        s = cbrtf(s * (float)(s >= 0)) + cbias;
        const float X = 0.5f * (l - m);
        const float Y = X + m;
        const float zy = s - Y;
        b[i] = (float)((double)zy + 0.55);
        r[i] = (float)((double)(X * 14.0f) + 0.42);
        g[i] = (float)((double)Y + 0.01);
    }
}

/* One axis of the reference's blur at every position: a sequential fmaf
 * chain over the 17 taps (zero outside), then the division by the in-bounds
 * tap sum integral[end] - integral[beg] (gaussianblur.hpp:41-77). */
static void blur_axis_f(const float* src, int64_t n, int64_t step, int64_t count, int64_t cstep, const float g[17],
                        const float integral[18], float* dst)
{
    for (int64_t x = 0; x < n; x++) {
        const int64_t lo = x - 8;
        const int beg = (int)((lo > 0 ? lo : 0) - lo);
        const int64_t hi = x + 9 < n ? x + 9 : n;
        const int64_t lo2 = lo < n ? lo : n;
        const int end = (int)(hi - lo2);
        const float tot = integral[end] - integral[beg];
        for (int64_t c = 0; c < count; c++) {
            float out = 0.0f;
            for (int i = 0; i < 17; i++) {
                const int64_t xx = lo + i;
                const float v = (xx >= 0 && xx < n) ? src[c * cstep + xx * step] : 0.0f;
                out = fmaf(v, g[i], out);
            }
            dst[c * cstep + x * step] = out / tot;
        }
    }
}

static void blur_f(const float* src, int64_t w, int64_t h, const float g[17], const float integral[18], float* tmp,
                   float* dst)
{
    blur_axis_f(src, w, 1, h, w, g, integral, tmp);  /* horizontal: rows */
    blur_axis_f(tmp, h, w, w, 1, g, integral, dst);  /* vertical: columns */
}

/* pointer-jumping tree of score.hpp:301-313 / 172-182 over n = 2^k values */
static float tree_sum(float* v, int n)
{
    for (int next = 1; next < n; next *= 2)
        for (int t = 0; t + next < n; t += 2 * next) v[t] += v[t + next];
    return v[0];
}

/* The six block-partial arrays of one (plane, scale) reduced as the reference
 * does: per 16x16 block a 256-wide tree divided by W*H, then 1024-wide trees
 * over the block index until one value is left. */
static void reduce_f(float* const maps[6], int64_t w, int64_t h, float out[6])
{
    const int64_t bx = (w - 1) / 16 + 1, by = (h - 1) / 16 + 1, nb = bx * by;
    const float area = (float)(w * h);
    float* part = (float*)malloc(sizeof(float) * (size_t)nb);
    float buf[1024];
    for (int k = 0; k < 6; k++) {
        for (int64_t j = 0; j < by; j++)
            for (int64_t i = 0; i < bx; i++) {
                for (int t = 0; t < 256; t++) {
                    const int64_t x = i * 16 + (t & 15), y = j * 16 + (t >> 4);
                    buf[t] = (x < w && y < h) ? maps[k][y * w + x] : 0.0f;
                }
                part[j * bx + i] = tree_sum(buf, 256) / area;
            }
        int64_t n = nb;
        while (n != 1) {
            const int64_t groups = (n - 1) / 1024 + 1;
            for (int64_t gi = 0; gi < groups; gi++) {
                for (int t = 0; t < 1024; t++) buf[t] = gi * 1024 + t < n ? part[gi * 1024 + t] : 0.0f;
                part[gi] = tree_sum(buf, 1024);
            }
            n = groups;
        }
        out[k] = part[0];
    }
    free(part);
}

/* The six reduced sums {ssim, art, det, ssim^4, art^4, det^4} of one
 * (plane, scale), reference score.hpp:192-322. */
static void plane_sums_f(const float* a, const float* b, int64_t w, int64_t h, int ssim_on, const float g[17],
                         const float integral[18], float out[6])
{
    const int64_t n = w * h;
    float* buf = (float*)calloc((size_t)(12 * n), sizeof(float));
    float *m1 = buf, *m2 = buf + n, *s12 = buf + 2 * n, *S = buf + 3 * n, *t = buf + 4 * n, *tmp = buf + 5 * n;
    float* maps[6] = {buf + 6 * n, buf + 7 * n, buf + 8 * n, buf + 9 * n, buf + 10 * n, buf + 11 * n};
    blur_f(a, w, h, g, integral, tmp, m1);
    blur_f(b, w, h, g, integral, tmp, m2);
    if (ssim_on) {
        for (int64_t i = 0; i < n; i++) t[i] = a[i] * b[i];
        blur_f(t, w, h, g, integral, tmp, s12);
        for (int64_t i = 0; i < n; i++) {
            const float u = a[i] + b[i];
            t[i] = u * u;
        }
        blur_f(t, w, h, g, integral, tmp, S);
    }
    for (int64_t i = 0; i < n; i++) {
        float ssim = 0.0f;
        if (ssim_on) {
            const float s11_s22 = fmaf(-2.0f, s12[i], S[i]);
            const float md = m1[i] - m2[i];
            const float num_m = fmaf(md, -md, 1.0f);
            const float num_s = fmaf(fmaf(-m1[i], m2[i], s12[i]), 2.0f, 0.0009f);
            const float den = fmaf(-m2[i], m2[i], fmaf(-m1[i], m1[i], s11_s22)) + 0.0009f;
            const float q = (num_m * num_s) / den;
            ssim = fmaxf(1.0f - q, 0.0f);
        }
        const float e = (fabsf(b[i] - m2[i]) + 1.0f) / (fabsf(a[i] - m1[i]) + 1.0f) - 1.0f;
        const float art = fmaxf(e, 0.0f), det = fmaxf(-e, 0.0f);
        const float v[3] = {ssim, art, det};
        for (int k = 0; k < 3; k++) {
            const float y = v[k] * v[k];
            maps[k][i] = v[k];
            maps[3 + k][i] = y * y;
        }
    }
    reduce_f(maps, w, h, out);
    free(buf);
}

/* Moves a pseudo-random fraction `frac` of the values by one float ulp up or
 * down (xorshift64 from `seed`): a model of the reference's device libm
 * (cbrtf, powf), whose results can differ from the host's in the last bit. */
static void perturb_ulp(float* v, int64_t n, uint64_t* state, double frac)
{
    for (int64_t i = 0; i < n; i++) {
        uint64_t x = *state;
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        *state = x;
        const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
        if (u < frac) v[i] = nextafterf(v[i], (x & 1) ? INFINITY : -INFINITY);
    }
}

static double reference_float_impl(const float* rgb1, const float* rgb2, int64_t w, int64_t h, uint64_t seed,
                                   double frac, double features[108]);

/* The reference's score of a pair given both images' linear RGB planes (3
 * planes of w*h floats each), with its pruning; features[108] receives the
 * float features it forms (L4 roots included). */
ORACLE_EXPORT double oracle_reference_float_from_linear(const float* rgb1, const float* rgb2, int64_t w, int64_t h,
                                                        double features[108])
{
    return reference_float_impl(rgb1, rgb2, w, h, 0, 0.0, features);
}

/* The same with a fraction `frac` of both images' XYB values (every pyramid
 * level) moved by one ulp, seeded: the spread of such scores is the
 * reference formula's sensitivity to last-bit differences of its libm. */
ORACLE_EXPORT double oracle_reference_float_perturbed(const float* rgb1, const float* rgb2, int64_t w, int64_t h,
                                                      uint64_t seed, double frac)
{
    return reference_float_impl(rgb1, rgb2, w, h, seed | 1, frac, NULL);
}

static double reference_float_impl(const float* rgb1, const float* rgb2, int64_t w, int64_t h, uint64_t seed,
                                   double frac, double features[108])
{
    int64_t T = 0;
    {
        int64_t cw = w, ch = h;
        for (int s = 0; s < 6; s++) { T += cw * ch; cw = half_up(cw); ch = half_up(ch); }
    }
    float* p[2] = {(float*)malloc(sizeof(float) * (size_t)(3 * T)), (float*)malloc(sizeof(float) * (size_t)(3 * T))};
    const float* src[2] = {rgb1, rgb2};
    for (int k = 0; k < 2; k++) {
        for (int c = 0; c < 3; c++) {
            float* base = p[k] + c * T;
            memcpy(base, src[k] + c * w * h, sizeof(float) * (size_t)(w * h));
            int64_t off = 0, cw = w, ch = h;
            for (int s = 1; s < 6; s++) {
                downsample_f(base + off, cw, ch, base + off + cw * ch);
                off += cw * ch; cw = half_up(cw); ch = half_up(ch);
            }
        }
        xyb_f(p[k], p[k] + T, p[k] + 2 * T, T);
        if (frac > 0.0) perturb_ulp(p[k], 3 * T, &seed, frac);
    }
    float g[17], integral[18];
    taps_float(g, integral);
    float res[108];
    memset(res, 0, sizeof(res));
    int64_t off = 0, cw = w, ch = h;
    for (int s = 0; s < 6; s++) {
        for (int pl = 0; pl < 3; pl++) {
            const int on[3] = {oracle_map_active(pl, s, 0), oracle_map_active(pl, s, 1), oracle_map_active(pl, s, 2)};
            if (!on[0] && !on[1] && !on[2]) continue;
            float sums[6];
            plane_sums_f(p[0] + pl * T + off, p[1] + pl * T + off, cw, ch, on[0], g, integral, sums);
            for (int m = 0; m < 3; m++) {
                if (!on[m]) continue;
                res[pl * 36 + s * 6 + m] = sums[m];
                res[pl * 36 + s * 6 + 3 + m] = sqrtf(sqrtf(sums[3 + m]));
            }
        }
        off += cw * ch; cw = half_up(cw); ch = half_up(ch);
    }
    free(p[0]); free(p[1]);
    double ssim = 0.0;
    for (int i = 0; i < 108; i++) {
        ssim = fmaf((float)oracle_weight(i), res[i], (float)ssim);
        if (features) features[i] = res[i];
    }
    ssim *= 0.9562382616834844;
    ssim = (6.248496625763138e-5 * ssim * ssim) * ssim + 2.326765642916932 * ssim - 0.020884521182843837 * ssim * ssim;
    ssim = ssim > 0.0 ? pow(ssim, 0.6276336467831387) * -10.0 + 100.0 : 100.0;
    return (double)(float)ssim;
}

/* The same on raw planes: the FP64 oracle's colour front-end rounded to
 * float, then the reference's float pyramid, XYB, maps and score. Returns 0,
 * -1 for a mismatched or unsupported colourspace, -2 on allocation failure. */
ORACLE_EXPORT int oracle_reference_float(const Vship_Colorspace_t* cs1, const uint8_t* const planes1[3],
                                         const int64_t stride1[3], const Vship_Colorspace_t* cs2,
                                         const uint8_t* const planes2[3], const int64_t stride2[3],
                                         double features[108], double* score)
{
    int64_t w1, h1, w2, h2;
    oracle_final_dims(cs1, &w1, &h1);
    oracle_final_dims(cs2, &w2, &h2);
    if (w1 != w2 || h1 != h2 || w1 <= 0 || h1 <= 0) return -1;
    const int64_t n = 3 * w1 * h1;
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    float* f1 = (float*)malloc(sizeof(float) * (size_t)n);
    float* f2 = (float*)malloc(sizeof(float) * (size_t)n);
    int rc = (!d || !f1 || !f2) ? -2 : 0;
    if (!rc) rc = oracle_linear_rgb(cs1, planes1, stride1, d);
    if (!rc) for (int64_t i = 0; i < n; i++) f1[i] = (float)d[i];
    if (!rc) rc = oracle_linear_rgb(cs2, planes2, stride2, d);
    if (!rc) for (int64_t i = 0; i < n; i++) f2[i] = (float)d[i];
    if (!rc) *score = oracle_reference_float_from_linear(f1, f2, w1, h1, features);
    free(d); free(f1); free(f2);
    return rc;
}
