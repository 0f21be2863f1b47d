/* ssimu2_oracle.c - scalar FP64 CPU restatement of the reference's SSIMULACRA2
 * path (colour front-end -> linear-RGB pyramid -> XYB -> 17-tap renormalised
 * FIR blur -> SSIM / edge-difference maps -> L1/L4 norms -> 108-weight score).
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
 * against. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference source tree). All arithmetic is double; where the reference
 * rounds to float the oracle does not, so the HIP path is compared with the
 * exact formula. Parity pins (tests/test_oracle_*.py):
 *   - gauss_taps / blur: Appendix-B tap values, scipy.ndimage.correlate with
 *     zero padding divided by the same correlation of an all-ones image,
 *     blur(constant) == constant at every border.
 *   - transfer / range / matrix: closed-form values (sRGB, 2.4 power, limited
 *     range 16/235/128, BT.709 inverse matrix vs numpy.linalg.inv of the
 *     forward matrix).
 *   - chroma upsampling: published Catmull-Rom weights (impulse responses at
 *     t = 0.25, 0.5, 0.75, 0.125 ...), linear ramps reproduced in the interior.
 *   - downsample / XYB / final score: closed forms (grey axis gives X' = 0.42
 *     and B' = 0.55; final_score(0) = 100, (1) = 83.5696784, (10) = 33.6420343).
 *   - whole pipeline: identity gives exactly 100; scores of the reference
 *     libvship built from its own sources (oracle/_ref) on the same frames,
 *     committed as tests/golden fixtures.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/VshipColor.h"

#define ORACLE_EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- weights */

/* The 108 SSIMULACRA2 weights, index plane*36 + scale*6 + norm*3 + map
 * (plane X,Y,B; norm L1,L4; map ssim, artifact, detail-loss).
 * Same table as reference src/ssimu2/score.hpp:5-114. */
static const double k_weights[108] = {
    /* X s0 */ 0.0, 0.0007376606707406586, 0.0, 0.0, 0.0007793481682867309, 0.0,
    /* X s1 */ 0.0, 0.0004371155730107379, 0.0, 1.1041726426657346, 0.00066284834129271, 0.00015231632783718752,
    /* X s2 */ 0.0, 0.0016406437456599754, 0.0, 1.8422455520539298, 11.441172603757666, 0.0,
    /* X s3 */ 0.0007989109436015163, 0.000176816438078653, 0.0, 1.8787594979546387, 10.94906990605142, 0.0,
    /* X s4 */ 0.0007289346991508072, 0.9677937080626833, 0.0, 0.00014003424285435884, 0.9981766977854967, 0.00031949755934435053,
    /* X s5 */ 0.0004550992113792063, 0.0, 0.0, 0.0013648766163243398, 0.0, 0.0,
    /* Y s0 */ 0.0, 0.0, 0.0, 7.466890328078848, 0.0, 17.445833984131262,
    /* Y s1 */ 0.0006235601634041466, 0.0, 0.0, 6.683678146179332, 0.00037724407979611296, 1.027889937768264,
    /* Y s2 */ 225.20515300849274, 0.0, 0.0, 19.213238186143016, 0.0011401524586618361, 0.001237755635509985,
    /* Y s3 */ 176.39317598450694, 0.0, 0.0, 24.43300999870476, 0.28520802612117757, 0.0004485436923833408,
    /* Y s4 */ 0.0, 0.0, 0.0, 34.77906344483772, 44.835625328877896, 0.0,
    /* Y s5 */ 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    /* B s0 */ 0.0, 0.0008680556573291698, 0.0, 0.0, 0.0, 0.0,
    /* B s1 */ 0.0, 0.0005313191874358747, 0.0, 0.00016533814161379112, 0.0, 0.0,
    /* B s2 */ 0.0, 0.0, 0.0, 0.0004179171803251336, 0.0017290828234722833, 0.0,
    /* B s3 */ 0.0020827005846636437, 0.0, 0.0, 8.826982764996862, 23.19243343998926, 0.0,
    /* B s4 */ 95.1080498811086, 0.9863978034400682, 0.9834382792465353, 0.0012286405048278493, 171.2667255897307, 0.9807858872435379,
    /* B s5 */ 0.0, 0.0, 0.0, 0.0005130064588990679, 0.0, 0.00010854057858411537,
};

ORACLE_EXPORT double oracle_weight(int i) { return (i >= 0 && i < 108) ? k_weights[i] : 0.0; }

/* A (plane, scale, map) is computed iff its L1 or its L4 weight exceeds 0.01
 * (reference src/ssimu2/score.hpp:4,122-131). */
ORACLE_EXPORT int oracle_map_active(int plane, int scale, int map)
{
    const double prune = 0.01;
    const double w1 = k_weights[plane * 36 + scale * 6 + 0 * 3 + map];
    const double w4 = k_weights[plane * 36 + scale * 6 + 1 * 3 + map];
    return !(w1 <= prune && w4 <= prune);
}

/* ------------------------------------------------------- sample decoding */

static int sample_bits(Vship_Sample_t s)
{
    /* reference src/gpuColorToLinear/rangeToFull.hpp:5-22 */
    switch (s) {
    case Vship_SampleUINT8: return 8;
    case Vship_SampleUINT9: return 9;
    case Vship_SampleUINT10: return 10;
    case Vship_SampleUINT12: return 12;
    case Vship_SampleUINT14: return 14;
    case Vship_SampleUINT16: return 16;
    default: return 1;
    }
}

static int sample_bytes(Vship_Sample_t s)
{
    /* reference src/gpuColorToLinear/anyDepthToFloat.hpp:7-21 */
    switch (s) {
    case Vship_SampleUINT8: return 1;
    case Vship_SampleFLOAT: return 4;
    default: return 2;
    }
}

static double half_bits_to_double(uint16_t h)
{
    const int sign = (h >> 15) & 1;
    const int expo = (h >> 10) & 0x1f;
    const int mant = h & 0x3ff;
    double v;
    if (expo == 0) v = ldexp((double)mant, -24);
    else if (expo == 31) v = mant ? NAN : INFINITY;
    else v = ldexp((double)(mant | 0x400), expo - 25);
    return sign ? -v : v;
}

/* reference src/gpuColorToLinear/anyDepthToFloat.hpp:24-68 (PickValue):
 * little-endian, integer samples above 8 bits masked to their depth. */
static double load_sample(const uint8_t* p, Vship_Sample_t s)
{
    switch (s) {
    case Vship_SampleUINT8: return (double)p[0];
    case Vship_SampleFLOAT: { float f; memcpy(&f, p, 4); return (double)f; }
    case Vship_SampleHALF: { uint16_t h = (uint16_t)(p[0] | (p[1] << 8)); return half_bits_to_double(h); }
    default: {
        uint16_t v = (uint16_t)(p[0] | (p[1] << 8));
        return (double)(v & ((1u << sample_bits(s)) - 1u));
    }
    }
}

/* reference src/gpuColorToLinear/rangeToFull.hpp:24-55 (FullRange) */
ORACLE_EXPORT double oracle_range_to_full(double a, int sample, int range, int chroma)
{
    const int bd = sample_bits((Vship_Sample_t)sample);
    if (range == Vship_RangeFull) {
        a /= (double)((1 << bd) - 1);
        if (chroma) a -= 0.5;
        return a;
    }
    if (bd >= 8) a /= (double)(1 << (bd - 8));
    else a *= 256.0;
    return chroma ? (a - 128.0) / 224.0 : (a - 16.0) / 219.0;
}

/* ------------------------------------------------------ chroma upsampling */

/* Cubic Hermite segment on [p0, p1] with Catmull-Rom tangents
 * (reference src/gpuColorToLinear/resize.hpp:5-24 and
 * chromaUpsample.hpp:28-50: m0 = (p1 - pm1)/2, m1 = (p2 - p0)/2). */
ORACLE_EXPORT double oracle_catmull_rom(double pm1, double p0, double p1, double p2, double t)
{
    const double m0 = (p1 - pm1) / 2.0;
    const double m1 = (p2 - p0) / 2.0;
    const double v1 = 2.0 * p0 + m0 - 2.0 * p1 + m1;
    const double v2 = -3.0 * p0 + 3.0 * p1 - 2.0 * m0 - m1;
    return ((v1 * t + v2) * t + m0) * t + p0;
}

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Segment [k, k+1] of a 1-D signal of n samples read with `step`, neighbours
 * clamped to [0, n-1] (the edge rules of chromaUpsample.hpp:29-35,41-47
 * reduce to index clamping). */
static double seg(const double* base, int64_t step, int64_t n, int64_t k, double t)
{
    const double pm1 = base[clampi(k - 1, 0, n - 1) * step];
    const double p0 = base[clampi(k, 0, n - 1) * step];
    const double p1 = base[clampi(k + 1, 0, n - 1) * step];
    const double p2 = base[clampi(k + 2, 0, n - 1) * step];
    return oracle_catmull_rom(pm1, p0, p1, p2, t);
}

/* Upsample one line of n_small samples to n_full samples, factor 2^sub, cosited
 * ("left"/"top") or centred. Reference kernels
 * bicubicHorizontal{Left,Center}UpscaleX{2,4}_Kernel (chromaUpsample.hpp:54-117)
 * and bicubicVertical{Center,Top}UpscaleX2_Kernel (:121-149). */
static void upsample_line(const double* src, int64_t sstep, int64_t n_small,
                          double* dst, int64_t dstep, int64_t n_full, int sub, int cosited)
{
    if (sub == 1 && cosited) {
        for (int64_t k = 0; k < n_small; k++) {
            if (2 * k < n_full) dst[(2 * k) * dstep] = src[k * sstep];
            if (2 * k + 1 < n_full) dst[(2 * k + 1) * dstep] = seg(src, sstep, n_small, k, 0.5);
        }
    } else if (sub == 1) {
        for (int64_t x = -1; x < n_small; x++) {
            if (x != -1 && 2 * x + 1 < n_full) dst[(2 * x + 1) * dstep] = seg(src, sstep, n_small, x, 0.25);
            if (2 * x + 2 < n_full) dst[(2 * x + 2) * dstep] = seg(src, sstep, n_small, x, 0.75);
        }
    } else if (sub == 2 && cosited) {
        for (int64_t k = 0; k < n_small; k++) {
            if (4 * k < n_full) dst[(4 * k) * dstep] = src[k * sstep];
            if (4 * k + 1 < n_full) dst[(4 * k + 1) * dstep] = seg(src, sstep, n_small, k, 0.25);
            if (4 * k + 2 < n_full) dst[(4 * k + 2) * dstep] = seg(src, sstep, n_small, k, 0.5);
            if (4 * k + 3 < n_full) dst[(4 * k + 3) * dstep] = seg(src, sstep, n_small, k, 0.75);
        }
    } else if (sub == 2) {
        for (int64_t x = -1; x < n_small; x++) {
            if (x != -1 && 4 * x + 2 < n_full) dst[(4 * x + 2) * dstep] = seg(src, sstep, n_small, x, 0.125);
            if (x != -1 && 4 * x + 3 < n_full) dst[(4 * x + 3) * dstep] = seg(src, sstep, n_small, x, 0.375);
            if (4 * x + 4 < n_full) dst[(4 * x + 4) * dstep] = seg(src, sstep, n_small, x, 0.625);
            if (4 * x + 5 < n_full) dst[(4 * x + 5) * dstep] = seg(src, sstep, n_small, x, 0.875);
        }
    }
}

/* reference src/gpuColorToLinear/chromaUpsample.hpp:153-247 (upsample):
 * horizontal pass first (Left/TopLeft cosited, Center/Top centred), then
 * vertical (Top/TopLeft cosited, Center/Left centred). subw <= 2, subh <= 1. */
ORACLE_EXPORT int oracle_upsample_plane(const double* small, int64_t w, int64_t h, int subw, int subh,
                                        int chroma_loc, double* out)
{
    if (subw < 0 || subw > 2 || subh < 0 || subh > 1) return -1;
    const int64_t sw = ((w - 1) >> subw) + 1;
    const int64_t sh = ((h - 1) >> subh) + 1;
    double* mid = (double*)malloc(sizeof(double) * (size_t)(w * sh));
    if (!mid) return -2;
    const int h_cosited = (chroma_loc == Vship_ChromaLoc_Left || chroma_loc == Vship_ChromaLoc_TopLeft);
    const int v_cosited = (chroma_loc == Vship_ChromaLoc_Top || chroma_loc == Vship_ChromaLoc_TopLeft);
    for (int64_t y = 0; y < sh; y++) {
        if (subw == 0) memcpy(mid + y * w, small + y * sw, sizeof(double) * (size_t)w);
        else upsample_line(small + y * sw, 1, sw, mid + y * w, 1, w, subw, h_cosited);
    }
    for (int64_t x = 0; x < w; x++) {
        if (subh == 0) {
            for (int64_t y = 0; y < h; y++) out[y * w + x] = mid[y * w + x];
        } else {
            upsample_line(mid + x, w, sh, out + x, w, h, subh, v_cosited);
        }
    }
    free(mid);
    return 0;
}

/* ------------------------------------------------- transfer and matrices */

static double spow(double a, double g) { return a < 0 ? -pow(-a, g) : pow(a, g); }

/* Per-component EOTFs, reference src/gpuColorToLinear/transferToLinear.hpp:
 * sRGB :73-88 (sign-symmetric), BT.709 :110-118 (pure 2.4 power), BT470_M
 * :128-131 (2.2), BT470_BG and BT601 :133-141 (2.8), ST240_M :143-158,
 * ST428 :162-165, PQ :169-182. HLG is applied on the triple (:24-51). */
ORACLE_EXPORT double oracle_transfer(double a, int trc)
{
    switch (trc) {
    case Vship_TRC_Linear: return a;
    case Vship_TRC_sRGB:
        if (a < 0) return (a < -0.04045) ? -pow((-a + 0.055) / 1.055, 2.4) : a / 12.92;
        return (a > 0.04045) ? pow((a + 0.055) / 1.055, 2.4) : a / 12.92;
    case Vship_TRC_BT709: return spow(a, 2.4);
    case Vship_TRC_BT470_M: return spow(a, 2.2);
    case Vship_TRC_BT470_BG: return spow(a, 2.8);
    case Vship_TRC_BT601: return spow(a, 2.8);
    case Vship_TRC_ST240_M:
        if (a >= 0) return (a < 0.0228) ? 4.0 * a : 1.1115 * pow(a, 0.45) - 0.1115;
        return (a > -0.0228) ? 4.0 * a : -1.1115 * pow(-a, 0.45) + 0.1115;
    case Vship_TRC_ST428: return spow(a, 2.6) * (52.37 / 48.0);
    case Vship_TRC_PQ: {
        const double Lmax = 100.0, n = 0.15930175781250000, m = 78.843750000000000;
        const double c1 = 0.83593750000000000, c2 = 18.851562500000000, c3 = 18.687500000000000;
        a = a > 0 ? a : 0.0;
        a = pow(a, 1.0 / m);
        const double num = (a - c1) > 0 ? (a - c1) : 0.0;
        return Lmax * pow(num / (c2 - c3 * a), 1.0 / n);
    }
    default: return a;
    }
}

static void transfer3(double v[3], int trc)
{
    if (trc == Vship_TRC_HLG) {
        /* reference transferToLinear.hpp:24-51: inverse OETF then OOTF with
         * system gamma 1.2 on the BT.2020 luminance. */
        const double a = 0.17883277, b = 1.0 - 4.0 * a, c = 0.5 - a * -0.3350097945111627;
        for (int i = 0; i < 3; i++)
            v[i] = (v[i] <= 0.5) ? v[i] * v[i] / 3.0 : (exp((v[i] - c) / a) + b) / 12.0;
        const double ys = 0.2627 * v[0] + 0.6780 * v[1] + 0.0593 * v[2];
        const double k = pow(ys, 1.2 - 1.0);
        for (int i = 0; i < 3; i++) v[i] *= k;
        return;
    }
    for (int i = 0; i < 3; i++) v[i] = oracle_transfer(v[i], trc);
}

/* Non-constant-luminance Y'CbCr -> R'G'B' (reference YUVToLinRGB.hpp:7-18);
 * v = {Y, Cb, Cr}. */
static void ncl(double v[3], double kr, double kb)
{
    const double kg = 1.0 - kb - kr;
    const double us = 0.5 / (1.0 - kb), vs = 0.5 / (1.0 - kr);
    const double r = (v[2] + v[0] * vs) / (vs * kr + 0.5);
    const double b = (v[1] + v[0] * us) / (us * kb + 0.5);
    const double g = (v[0] - kr * r - kb * b) / kg;
    v[0] = r; v[1] = g; v[2] = b;
}

/* reference YUVToLinRGB.hpp:157-163: gamma-light matrix, EOTF, linear-light
 * matrix. Per-matrix steps :26-155. */
ORACLE_EXPORT void oracle_yuv_to_linear(double v[3], int matrix, int trc)
{
    if (trc == Vship_TRC_Linear && matrix == Vship_MATRIX_RGB) return; /* YUVToLinRGB.hpp:219 */
    switch (matrix) {
    case Vship_MATRIX_BT709: ncl(v, 0.2126, 0.0722); break;
    case Vship_MATRIX_BT470_BG:
    case Vship_MATRIX_ST170_M: ncl(v, 0.299, 0.114); break;
    case Vship_MATRIX_BT2020_NCL: ncl(v, 0.2627, 0.0593); break;
    case Vship_MATRIX_BT2020_CL: {
        /* reference YUVToLinRGB.hpp:77-94: {Y', Cb, Cr} -> {R', Y', B'} */
        const double kr = 0.2627, kb = 0.0593;
        const double us = 0.5 / (1.0 - kb), vs = 0.5 / (1.0 - kr);
        const double r = (v[2] + v[0] * vs) / (vs * kr + 0.5);
        const double b = (v[1] + v[0] * us) / (us * kb + 0.5);
        const double y = v[0];
        v[0] = r; v[1] = y; v[2] = b;
        break;
    }
    case Vship_MATRIX_BT2100_ICTCP: {
        /* ICtCp -> L'M'S', reference YUVToLinRGB.hpp:47-63 (order-dependent) */
        const double i = v[0], ct = v[1], cp = v[2];
        const double l = ct * 0.0086090370379327566 + (cp * 0.11102962500302595655 + i);
        const double m = ct * -0.01721807407586551 + (cp * -0.22205925000605192 + l);
        const double s = l * 33.025782688766114 + (m * -32.025782688766114 + -cp * 7.543278084714549);
        v[0] = l; v[1] = m; v[2] = s;
        break;
    }
    default: break;
    }
    transfer3(v, trc);
    if (matrix == Vship_MATRIX_BT2020_CL) {
        /* reference YUVToLinRGB.hpp:65-75: G from the linear luminance */
        const double kr = 0.2627, kb = 0.0593, kg = 1.0 - kr - kb;
        v[1] = (v[1] - kr * v[0] - kb * v[2]) / kg;
    }
    if (matrix == Vship_MATRIX_BT2100_ICTCP) {
        /* LMS -> RGB, reference YUVToLinRGB.hpp:28-43 (each row uses the
         * components already overwritten above it) */
        const double l = v[0], m = v[1], s = v[2];
        const double r = l * 3.4366066943330784267 + (m * -2.5064521186562698975 + s * 0.069845424323191470954);
        const double g = r * -0.23026480071281397 + (m * 1.4064527541936944 + -s * 0.2083538389058436);
        const double b = r * -0.02684381778741865 + (g * -0.08378524945770065 + s * 1.106879231557486);
        v[0] = r; v[1] = g; v[2] = b;
    }
}

/* ------------------------------------------------------------- primaries */

/* RGB -> XYZ and XYZ -> RGB matrices per primaries set, and the direct
 * BT.2020 -> BT.709 matrix; same constants as reference
 * src/gpuColorToLinear/primaries.hpp:17-131. Index 0..6 via prim_index. */
static const double k_to_xyz[6][9] = {
    /* BT709 */ {0.4124, 0.3576, 0.1805, 0.2126, 0.7152, 0.0722, 0.0193, 0.1192, 0.9505},
    /* BT470_M */ {0.6068, 0.1735, 0.2003, 0.2989, 0.5866, 0.1145, 0.0, 0.06609, 1.116},
    /* BT470_BG */ {0.4305, 0.3416, 0.1783, 0.2220, 0.7067, 0.07132, 0.0202, 0.1296, 0.9391},
    /* ST170_M */ {0.39349893, 0.3652766, 0.19162447, 0.2123645, 0.70109542, 0.08654008, 0.01873804, 0.1119396, 0.95812235},
    /* ST240_M */ {0.39349893, 0.3652766, 0.19162447, 0.2123645, 0.70109542, 0.08654008, 0.01873804, 0.1119396, 0.95812235},
    /* BT2020 */ {0.6370, 0.1446, 0.1689, 0.2627, 0.6780, 0.0593, 0.0, 0.0281, 1.0610},
};
static const double k_from_xyz_709[9] = {3.2410, -1.5374, -0.4986, -0.9692, 1.8760, 0.0416, 0.0556, -0.2040, 1.0570};
static const double k_2020_to_709[9] = {1.6605, -0.5876, -0.0728, -0.1246, 1.1329, -0.0083, -0.0182, -0.1006, 1.1187};

static int prim_index(int p)
{
    switch (p) {
    case Vship_PRIMARIES_BT709: return 0;
    case Vship_PRIMARIES_BT470_M: return 1;
    case Vship_PRIMARIES_BT470_BG: return 2;
    case Vship_PRIMARIES_ST170_M: return 3;
    case Vship_PRIMARIES_ST240_M: return 4;
    case Vship_PRIMARIES_BT2020: return 5;
    default: return -1;
    }
}

static void mat3(const double m[9], double v[3])
{
    const double a = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
    const double b = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
    const double c = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
    v[0] = a; v[1] = b; v[2] = c;
}

/* reference primaries.hpp:134-144 with destination BT.709
 * (vshipColor.hpp:119-122): direct matrix for BT.2020, else through XYZ. */
ORACLE_EXPORT void oracle_primaries_to_709(double v[3], int prim)
{
    if (prim == Vship_PRIMARIES_BT709) return;
    if (prim == Vship_PRIMARIES_BT2020) { mat3(k_2020_to_709, v); return; }
    if (prim != Vship_PRIMARIES_INTERNAL) {
        const int i = prim_index(prim);
        if (i >= 0) mat3(k_to_xyz[i], v);
    }
    mat3(k_from_xyz_709, v);
}

/* ---------------------------------------------------------- crop / resize */

/* One separable cubic-Hermite pass that also transposes, reference
 * src/gpuColorToLinear/resize.hpp:26-48: dst[x*sh + y] from row y of src. */
static void resize_transpose(double* dst, const double* src, int64_t sw, int64_t sh, int64_t rw)
{
    for (int64_t y = 0; y < sh; y++) {
        for (int64_t x = 0; x < rw; x++) {
            const double pos = (double)x * (double)sw / (double)rw;
            const int64_t b = (int64_t)pos;
            const double* row = src + y * sw;
            const double v1 = row[b];
            const double v0 = (b - 1 >= 0) ? row[b - 1] : v1;
            const double v2 = (b + 1 < sw) ? row[b + 1] : v1;
            const double v3 = (b + 2 < sw) ? row[b + 2] : v2;
            dst[x * sh + y] = oracle_catmull_rom(v0, v1, v2, v3, pos - (double)b);
        }
    }
}

/* reference resize.hpp:51-65 (resizePlane) */
static int resize_plane(double* dst, const double* src, int64_t sw, int64_t sh, int64_t rw, int64_t rh)
{
    if (sw == rw && sh == rh) { memcpy(dst, src, sizeof(double) * (size_t)(sw * sh)); return 0; }
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(rw * sh));
    if (!tmp) return -2;
    resize_transpose(tmp, src, sw, sh, rw);
    resize_transpose(dst, tmp, sh, rw, rh);
    free(tmp);
    return 0;
}

/* Final (post resize and crop) size, reference vshipColor.hpp:46-61. */
ORACLE_EXPORT void oracle_final_dims(const Vship_Colorspace_t* cs, int64_t* w, int64_t* h)
{
    const int64_t tw = cs->target_width == -1 ? cs->width : cs->target_width;
    const int64_t th = cs->target_height == -1 ? cs->height : cs->target_height;
    *w = tw - cs->crop.left - cs->crop.right;
    *h = th - cs->crop.top - cs->crop.bottom;
}

/* ------------------------------------------------------ colour front-end */

/* Raw planes -> linear BT.709 RGB, 3 planes of the final size, in the order
 * of reference src/gpuColorToLinear/vshipColor.hpp:62-146: decode + range per
 * plane, chroma upsampling, matrix + EOTF, primaries, resize then crop. */
ORACLE_EXPORT int oracle_linear_rgb(const Vship_Colorspace_t* cs, const uint8_t* const planes[3],
                                    const int64_t stride[3], double* out)
{
    const int64_t W = cs->width, H = cs->height;
    const int subw = cs->subsampling.subw, subh = cs->subsampling.subh;
    if (W <= 0 || H <= 0 || subw < 0 || subw > 2 || subh < 0 || subh > 1) return -1;
    const int64_t pw[3] = {W, ((W - 1) >> subw) + 1, ((W - 1) >> subw) + 1};
    const int64_t ph[3] = {H, ((H - 1) >> subh) + 1, ((H - 1) >> subh) + 1};
    const int bytes = sample_bytes(cs->sample);
    double* full = (double*)malloc(sizeof(double) * (size_t)(3 * W * H));
    double* small = (double*)malloc(sizeof(double) * (size_t)(pw[1] * ph[1]));
    if (!full || !small) { free(full); free(small); return -2; }
    for (int p = 0; p < 3; p++) {
        const int chroma = (p != 0) && cs->colorFamily == Vship_ColorYUV;
        double* dst = (p == 0 || (subw == 0 && subh == 0)) ? full + p * W * H : small;
        for (int64_t y = 0; y < ph[p]; y++)
            for (int64_t x = 0; x < pw[p]; x++)
                dst[y * pw[p] + x] = oracle_range_to_full(
                    load_sample(planes[p] + y * stride[p] + x * bytes, cs->sample),
                    cs->sample, cs->range, chroma);
        if (p != 0 && !(subw == 0 && subh == 0)) {
            if (oracle_upsample_plane(small, W, H, subw, subh, cs->chromaLocation, full + p * W * H) != 0) {
                free(full); free(small); return -1;
            }
        }
    }
    free(small);
    for (int64_t i = 0; i < W * H; i++) {
        double v[3] = {full[i], full[W * H + i], full[2 * W * H + i]};
        oracle_yuv_to_linear(v, cs->YUVMatrix, cs->transferFunction);
        oracle_primaries_to_709(v, cs->primaries);
        full[i] = v[0]; full[W * H + i] = v[1]; full[2 * W * H + i] = v[2];
    }
    const int cropped = cs->crop.top || cs->crop.bottom || cs->crop.left || cs->crop.right;
    const int64_t tw = cs->target_width == -1 ? W : cs->target_width;
    const int64_t th = cs->target_height == -1 ? H : cs->target_height;
    int64_t fw, fh;
    oracle_final_dims(cs, &fw, &fh);
    if (!cropped && tw == W && th == H) {
        memcpy(out, full, sizeof(double) * (size_t)(3 * W * H));
    } else if (!cropped) {
        for (int p = 0; p < 3; p++) resize_plane(out + p * tw * th, full + p * W * H, W, H, tw, th);
    } else {
        double* rs = (double*)malloc(sizeof(double) * (size_t)(tw * th));
        if (!rs) { free(full); return -2; }
        for (int p = 0; p < 3; p++) {
            resize_plane(rs, full + p * W * H, W, H, tw, th);
            for (int64_t y = 0; y < fh; y++)
                for (int64_t x = 0; x < fw; x++)
                    out[p * fw * fh + y * fw + x] = rs[(cs->crop.top + y) * tw + cs->crop.left + x];
        }
        free(rs);
    }
    free(full);
    return 0;
}

/* ------------------------------------------------------- pyramid and XYB */

static int64_t half_up(int64_t n) { return (n - 1) / 2 + 1; }

/* 2x2 box mean with the edge clamp min(2y+1, h-1), min(2x+1, w-1)
 * (reference src/ssimu2/downsample.hpp:3-19). */
ORACLE_EXPORT void oracle_downsample(const double* src, int64_t w, int64_t h, double* dst)
{
    const int64_t nw = half_up(w), nh = half_up(h);
    for (int64_t y = 0; y < nh; y++) {
        const int64_t y0 = 2 * y, y1 = (2 * y + 1 < h) ? 2 * y + 1 : h - 1;
        for (int64_t x = 0; x < nw; x++) {
            const int64_t x0 = 2 * x, x1 = (2 * x + 1 < w) ? 2 * x + 1 : w - 1;
            dst[y * nw + x] = (src[y0 * w + x0] + src[y1 * w + x0] + src[y0 * w + x1] + src[y1 * w + x1]) * 0.25;
        }
    }
}

/* Linear RGB -> positive XYB in place (reference src/ssimu2/makeXYB.hpp:6-52):
 * opsin mix + bias, cbrt of the non-negative part minus the bias' cube root,
 * X = (l - m)/2, Y = (l + m)/2, then B' = (B - Y) + 0.55, X' = 14 X + 0.42,
 * Y' = Y + 0.01. */
ORACLE_EXPORT void oracle_xyb(double* r, double* g, double* b, int64_t n)
{
    const double bias = 0.0037930734, cbias = -0.1559542025327239;
    for (int64_t i = 0; i < n; i++) {
        double l = 0.30 * r[i] + 0.622 * g[i] + 0.078 * b[i] + bias;
        double m = 0.23 * r[i] + 0.692 * g[i] + 0.078 * b[i] + bias;
        double s = 0.24342269 * r[i] + 0.20476745 * g[i] + 0.55180986 * b[i] + bias;
        l = cbrt(l > 0 ? l : 0.0) + cbias;
        m = cbrt(m > 0 ? m : 0.0) + cbias;
        s = cbrt(s > 0 ? s : 0.0) + cbias;
        const double X = 0.5 * (l - m), Y = 0.5 * (l + m), B = s;
        r[i] = X * 14.0 + 0.42;
        g[i] = Y + 0.01;
        b[i] = (B - Y) + 0.55;
    }
}

/* Pixel count of the 6-level pyramid (reference src/ssimu2/main.hpp:16-24) */
ORACLE_EXPORT int64_t oracle_total_scale_size(int64_t w, int64_t h)
{
    int64_t t = 0;
    for (int s = 0; s < 6; s++) { t += w * h; w = half_up(w); h = half_up(h); }
    return t;
}

/* Linear RGB (3 planes w*h) -> XYB pyramid, 3 planes of total_scale_size
 * each, levels back to back (reference src/ssimu2/main.hpp:30-56). */
ORACLE_EXPORT void oracle_xyb_pyramid(const double* rgb, int64_t w, int64_t h, double* out)
{
    const int64_t T = oracle_total_scale_size(w, h);
    for (int p = 0; p < 3; p++) {
        memcpy(out + p * T, rgb + p * w * h, sizeof(double) * (size_t)(w * h));
        int64_t off = 0, cw = w, ch = h;
        for (int s = 1; s < 6; s++) {
            oracle_downsample(out + p * T + off, cw, ch, out + p * T + off + cw * ch);
            off += cw * ch; cw = half_up(cw); ch = half_up(ch);
        }
    }
    oracle_xyb(out, out + T, out + 2 * T, T);
}

/* --------------------------------------------------------------- blur */

/* g_i = exp(-(8-i)^2 / (2 sigma^2)) / sqrt(2 pi sigma^2), sigma 1.5, i = 0..16
 * (reference src/ssimu2/gaussianblur.hpp:7-14, src/util/preprocessor.hpp:81-84) */
ORACLE_EXPORT void oracle_gauss_taps(double taps[17])
{
    const double sigma = 1.5;
    for (int i = 0; i < 17; i++)
        taps[i] = exp(-(double)((8 - i) * (8 - i)) / (2.0 * sigma * sigma)) / sqrt(2.0 * M_PI * sigma * sigma);
}

/* Separable 17-tap FIR, zero outside the image, each axis divided by the sum
 * of its in-bounds taps (reference src/ssimu2/gaussianblur.hpp:24-77). */
ORACLE_EXPORT void oracle_blur(const double* src, int64_t w, int64_t h, double* dst)
{
    double g[17];
    oracle_gauss_taps(g);
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(w * h));
    for (int64_t y = 0; y < h; y++) {
        for (int64_t x = 0; x < w; x++) {
            double acc = 0.0, nrm = 0.0;
            for (int i = 0; i < 17; i++) {
                const int64_t xx = x + i - 8;
                if (xx < 0 || xx >= w) continue;
                acc += g[i] * src[y * w + xx];
                nrm += g[i];
            }
            tmp[y * w + x] = acc / nrm;
        }
    }
    for (int64_t y = 0; y < h; y++) {
        for (int64_t x = 0; x < w; x++) {
            double acc = 0.0, nrm = 0.0;
            for (int j = 0; j < 17; j++) {
                const int64_t yy = y + j - 8;
                if (yy < 0 || yy >= h) continue;
                acc += g[j] * tmp[yy * w + x];
                nrm += g[j];
            }
            dst[y * w + x] = acc / nrm;
        }
    }
    free(tmp);
}

/* ------------------------------------------------------ maps and norms */

/* SSIM, artifact and detail-loss maps of one (plane, scale), reference
 * src/ssimu2/score.hpp:232-282: s11 + s22 = G(a^2 + b^2) (the reference gets
 * it as G((a+b)^2) - 2 G(ab), equal in exact arithmetic), C2 = 0.0009. Any
 * of the three output pointers may be NULL. */
ORACLE_EXPORT void oracle_maps(const double* a, const double* b, int64_t w, int64_t h,
                               double* ssim, double* art, double* det)
{
    const int64_t n = w * h;
    double* buf = (double*)malloc(sizeof(double) * (size_t)(6 * n));
    double *m1 = buf, *m2 = buf + n, *s12 = buf + 2 * n, *sq = buf + 3 * n, *t1 = buf + 4 * n, *t2 = buf + 5 * n;
    oracle_blur(a, w, h, m1);
    oracle_blur(b, w, h, m2);
    for (int64_t i = 0; i < n; i++) { t1[i] = a[i] * b[i]; t2[i] = a[i] * a[i] + b[i] * b[i]; }
    oracle_blur(t1, w, h, s12);
    oracle_blur(t2, w, h, sq);
    const double C2 = 0.0009;
    for (int64_t i = 0; i < n; i++) {
        if (ssim) {
            const double md = m1[i] - m2[i];
            const double num_m = 1.0 - md * md;
            /* evaluated so that a == b gives num_s == den_s bit for bit */
            const double num_s = 2.0 * (s12[i] - m1[i] * m2[i]) + C2;
            const double den_s = (sq[i] - (m1[i] * m1[i] + m2[i] * m2[i])) + C2;
            const double d = 1.0 - num_m * num_s / den_s;
            ssim[i] = d > 0 ? d : 0.0;
        }
        const double e = (fabs(b[i] - m2[i]) + 1.0) / (fabs(a[i] - m1[i]) + 1.0) - 1.0;
        if (art) art[i] = e > 0 ? e : 0.0;
        if (det) det[i] = -e > 0 ? -e : 0.0;
    }
    free(buf);
}

/* {L1 ssim, L1 art, L1 det, L4 ssim, L4 art, L4 det} of one (plane, scale):
 * L1 = mean, L4 = mean(x^4)^(1/4) (reference score.hpp:284-320,422-438). */
ORACLE_EXPORT void oracle_plane_features(const double* a, const double* b, int64_t w, int64_t h, double out[6])
{
    const int64_t n = w * h;
    double* m = (double*)malloc(sizeof(double) * (size_t)(3 * n));
    oracle_maps(a, b, w, h, m, m + n, m + 2 * n);
    for (int k = 0; k < 3; k++) {
        double s1 = 0.0, s4 = 0.0;
        for (int64_t i = 0; i < n; i++) { const double v = m[k * n + i]; const double v2 = v * v; s1 += v; s4 += v2 * v2; }
        out[k] = s1 / (double)n;
        out[3 + k] = sqrt(sqrt(s4 / (double)n));
    }
    free(m);
}

/* ------------------------------------------------------------ final score */

/* reference src/ssimu2/score.hpp:442-460 (in double throughout) */
ORACLE_EXPORT double oracle_final_score(const double features[108])
{
    double s = 0.0;
    for (int i = 0; i < 108; i++) s += k_weights[i] * features[i];
    s *= 0.9562382616834844;
    s = (6.248496625763138e-5 * s * s) * s + 2.326765642916932 * s - 0.020884521182843837 * s * s;
    return (s > 0.0) ? pow(s, 0.6276336467831387) * -10.0 + 100.0 : 100.0;
}

/* Features of two linear-RGB images of equal size. prune = 1 computes only
 * the maps the reference computes (others stay 0); prune = 0 all 108. */
ORACLE_EXPORT int oracle_features_from_linear(const double* rgb1, const double* rgb2, int64_t w, int64_t h,
                                              int prune, double features[108])
{
    const int64_t T = oracle_total_scale_size(w, h);
    double* p1 = (double*)malloc(sizeof(double) * (size_t)(3 * T));
    double* p2 = (double*)malloc(sizeof(double) * (size_t)(3 * T));
    if (!p1 || !p2) { free(p1); free(p2); return -2; }
    oracle_xyb_pyramid(rgb1, w, h, p1);
    oracle_xyb_pyramid(rgb2, w, h, p2);
    memset(features, 0, sizeof(double) * 108);
    int64_t off = 0, cw = w, ch = h;
    for (int s = 0; s < 6; s++) {
        for (int p = 0; p < 3; p++) {
            const int any = oracle_map_active(p, s, 0) || oracle_map_active(p, s, 1) || oracle_map_active(p, s, 2);
            if (prune && !any) continue;
            double f[6];
            oracle_plane_features(p1 + p * T + off, p2 + p * T + off, cw, ch, f);
            for (int norm = 0; norm < 2; norm++)
                for (int m = 0; m < 3; m++)
                    if (!prune || oracle_map_active(p, s, m))
                        features[p * 36 + s * 6 + norm * 3 + m] = f[norm * 3 + m];
        }
        off += cw * ch; cw = half_up(cw); ch = half_up(ch);
    }
    free(p1); free(p2);
    return 0;
}

/* Whole path on raw planes: the oracle of Vship_ComputeSSIMU2. Returns 0, -1
 * for an unsupported or mismatched colourspace, -2 on allocation failure. */
ORACLE_EXPORT int oracle_ssimu2(const Vship_Colorspace_t* cs1, const uint8_t* const planes1[3], const int64_t stride1[3],
                                const Vship_Colorspace_t* cs2, const uint8_t* const planes2[3], const int64_t stride2[3],
                                int prune, double features[108], double* score)
{
    int64_t w1, h1, w2, h2;
    oracle_final_dims(cs1, &w1, &h1);
    oracle_final_dims(cs2, &w2, &h2);
    if (w1 != w2 || h1 != h2 || w1 <= 0 || h1 <= 0) return -1;
    double* rgb1 = (double*)malloc(sizeof(double) * (size_t)(3 * w1 * h1));
    double* rgb2 = (double*)malloc(sizeof(double) * (size_t)(3 * w1 * h1));
    int rc = (!rgb1 || !rgb2) ? -2 : 0;
    if (!rc) rc = oracle_linear_rgb(cs1, planes1, stride1, rgb1);
    if (!rc) rc = oracle_linear_rgb(cs2, planes2, stride2, rgb2);
    double f[108];
    if (!rc) rc = oracle_features_from_linear(rgb1, rgb2, w1, h1, prune, f);
    if (!rc) {
        if (features) memcpy(features, f, sizeof(f));
        *score = oracle_final_score(f);
    }
    free(rgb1); free(rgb2);
    return rc;
}
