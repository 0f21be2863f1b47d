// engine.cpp - host side of the plain (one pair per call) SSIMULACRA2 path.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>

#include "errors.hpp"
#include "kernels.hpp"
#include "color_device.hpp"
#include "weights.hpp"

namespace vship {

namespace {

const double kWeights[108] = VSHIP_SSIMU2_WEIGHTS;

bool map_active(int plane, int scale, int map)
{
    const double w1 = kWeights[plane * 36 + scale * 6 + map];
    const double w4 = kWeights[plane * 36 + scale * 6 + 3 + map];
    return !(w1 <= kWeightPruning && w4 <= kWeightPruning);
}

int sample_bits(int s)
{
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

// RGB -> XYZ per primaries set and XYZ -> BT.709; same constants as reference
// src/gpuColorToLinear/primaries.hpp:17-131.
const double kToXyz709[9] = {0.4124, 0.3576, 0.1805, 0.2126, 0.7152, 0.0722, 0.0193, 0.1192, 0.9505};
const double kToXyz470M[9] = {0.6068, 0.1735, 0.2003, 0.2989, 0.5866, 0.1145, 0.0, 0.06609, 1.116};
const double kToXyz470BG[9] = {0.4305, 0.3416, 0.1783, 0.2220, 0.7067, 0.07132, 0.0202, 0.1296, 0.9391};
const double kToXyz170M[9] = {0.39349893, 0.3652766, 0.19162447, 0.2123645, 0.70109542, 0.08654008,
                              0.01873804, 0.1119396, 0.95812235};
const double kFromXyz709[9] = {3.2410, -1.5374, -0.4986, -0.9692, 1.8760, 0.0416, 0.0556, -0.2040, 1.0570};
const double k2020To709[9] = {1.6605, -0.5876, -0.0728, -0.1246, 1.1329, -0.0083, -0.0182, -0.1006, 1.1187};
const double kToXyz2020[9] = {0.6370, 0.1446, 0.1689, 0.2627, 0.6780, 0.0593, 0.0, 0.0281, 1.0610};

void matmul3(const double a[9], const double b[9], double out[9])
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) out[r * 3 + c] = a[r * 3 + 0] * b[0 * 3 + c] + a[r * 3 + 1] * b[1 * 3 + c] + a[r * 3 + 2] * b[2 * 3 + c];
}

// Source primaries -> BT.709 as one matrix (primaries.hpp:134-144,195-220);
// identity for BT.709.
void primaries_matrix(int prim, double out[9])
{
    const double* to = nullptr;
    switch (prim) {
    case Vship_PRIMARIES_BT2020:
        for (int i = 0; i < 9; i++) out[i] = k2020To709[i];
        return;
    case Vship_PRIMARIES_INTERNAL:
        for (int i = 0; i < 9; i++) out[i] = kFromXyz709[i];
        return;
    case Vship_PRIMARIES_BT470_M: to = kToXyz470M; break;
    case Vship_PRIMARIES_BT470_BG: to = kToXyz470BG; break;
    case Vship_PRIMARIES_ST170_M:
    case Vship_PRIMARIES_ST240_M: to = kToXyz170M; break;
    default:
        for (int i = 0; i < 9; i++) out[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    (void)kToXyz709;
    (void)kToXyz2020;
    matmul3(kFromXyz709, to, out);
}

// Rows of R' and B' from (Y', Cb, Cr) for constants Kr, Kb, written as the
// reference evaluates them (YUVToLinRGB.hpp:7-18): R = (Cr + Y vs)/(vs Kr + 1/2),
// B = (Cb + Y us)/(us Kb + 1/2).
void ncl_rows(double kr, double kb, double R[3], double B[3])
{
    const double us = 0.5 / (1.0 - kb), vs = 0.5 / (1.0 - kr);
    const double dr = vs * kr + 0.5, db = us * kb + 0.5;
    R[0] = vs / dr; R[1] = 0.0; R[2] = 1.0 / dr;
    B[0] = us / db; B[1] = 1.0 / db; B[2] = 0.0;
}

// The matrix steps of YUVToLinRGB.hpp:157-163 as pre (before the EOTF) and
// post (after it) 3x3 matrices; post also carries the primaries conversion.
void colour_matrices(int matrix, int prim, double pre[9], double post[9], bool* use_pre)
{
    double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double lin[9];
    for (int i = 0; i < 9; i++) { pre[i] = id[i]; lin[i] = id[i]; }
    *use_pre = true;
    switch (matrix) {
    case Vship_MATRIX_BT709: case Vship_MATRIX_BT470_BG: case Vship_MATRIX_ST170_M: case Vship_MATRIX_BT2020_NCL: {
        double kr = 0.2126, kb = 0.0722;
        if (matrix == Vship_MATRIX_BT470_BG || matrix == Vship_MATRIX_ST170_M) { kr = 0.299; kb = 0.114; }
        if (matrix == Vship_MATRIX_BT2020_NCL) { kr = 0.2627; kb = 0.0593; }
        const double kg = 1.0 - kb - kr;
        double R[3], B[3];
        ncl_rows(kr, kb, R, B);
        for (int c = 0; c < 3; c++) {
            pre[0 * 3 + c] = R[c];
            pre[2 * 3 + c] = B[c];
            pre[1 * 3 + c] = ((c == 0 ? 1.0 : 0.0) - kr * R[c] - kb * B[c]) / kg;
        }
        break;
    }
    case Vship_MATRIX_BT2020_CL: {
        // YUVToLinRGB.hpp:65-94: {R', Y', B'}, EOTF, then G from the linear luminance
        const double kr = 0.2627, kb = 0.0593, kg = 1.0 - kr - kb;
        double R[3], B[3];
        ncl_rows(kr, kb, R, B);
        for (int c = 0; c < 3; c++) {
            pre[0 * 3 + c] = R[c];
            pre[1 * 3 + c] = (c == 0) ? 1.0 : 0.0;
            pre[2 * 3 + c] = B[c];
        }
        const double g[9] = {1, 0, 0, -kr / kg, 1.0 / kg, -kb / kg, 0, 0, 1};
        for (int i = 0; i < 9; i++) lin[i] = g[i];
        break;
    }
    case Vship_MATRIX_BT2100_ICTCP: {
        // YUVToLinRGB.hpp:26-63: ICtCp -> L'M'S' (rows expanded), EOTF, LMS -> RGB
        const double L[3] = {1.0, 0.0086090370379327566, 0.11102962500302595655};
        const double M[3] = {L[0], L[1] - 0.01721807407586551, L[2] - 0.22205925000605192};
        double S3[3];
        for (int c = 0; c < 3; c++)
            S3[c] = 33.025782688766114 * L[c] - 32.025782688766114 * M[c] - (c == 2 ? 7.543278084714549 : 0.0);
        for (int c = 0; c < 3; c++) { pre[c] = L[c]; pre[3 + c] = M[c]; pre[6 + c] = S3[c]; }
        const double Rr[3] = {3.4366066943330784267, -2.5064521186562698975, 0.069845424323191470954};
        double Gr[3], Br[3];
        for (int c = 0; c < 3; c++)
            Gr[c] = -0.23026480071281397 * Rr[c] + (c == 1 ? 1.4064527541936944 : 0.0) + (c == 2 ? -0.2083538389058436 : 0.0);
        for (int c = 0; c < 3; c++)
            Br[c] = -0.02684381778741865 * Rr[c] - 0.08378524945770065 * Gr[c] + (c == 2 ? 1.106879231557486 : 0.0);
        for (int c = 0; c < 3; c++) { lin[c] = Rr[c]; lin[3 + c] = Gr[c]; lin[6 + c] = Br[c]; }
        break;
    }
    default:
        *use_pre = false;
        break;
    }
    double prm[9];
    primaries_matrix(prim, prm);
    matmul3(prm, lin, post);
}

int64_t half_up(int64_t n) { return (n - 1) / 2 + 1; }

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

// Catmull-Rom weights (the cubic Hermite of resize.hpp:5-24 with tangents
// (p[i+1] - p[i-1]) / 2) of the two output phases of a 2x-upsampled chroma
// axis, over the window of chroma samples m-2..m+2 around m = X >> 1 for
// output X = 2m + phase. Cosited samples sit on even X (phase 0 copies,
// phase 1 is t = 0.5 between m and m+1, chromaUpsample.hpp:28-52); centred
// samples sit on X = 2k + 0.5 (phase 0 is t = 0.75 between m-1 and m, phase
// 1 is t = 0.25 between m and m+1, chromaUpsample.hpp:54-149).
static void phase_weights(bool cosited, float w[2][5])
{
    for (int ph = 0; ph < 2; ph++) {
        for (int q = 0; q < 5; q++) w[ph][q] = 0.0f;
        const int k = cosited ? 0 : (ph == 0 ? -1 : 0);  // segment start, relative to m
        if (cosited && ph == 0) {
            w[ph][k + 2] = 1.0f;
            continue;
        }
        const float t = cosited ? 0.5f : (ph == 0 ? 0.75f : 0.25f);
        const float t2 = t * t, t3 = t2 * t;
        // taps at chroma m+k-1 .. m+k+2 = window entries k+1 .. k+4
        w[ph][k + 1] = 0.5f * (-t3 + 2.0f * t2 - t);
        w[ph][k + 2] = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
        w[ph][k + 3] = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
        w[ph][k + 4] = 0.5f * (t3 - t2);
    }
}

ImageSpec make_image_spec(const Vship_Colorspace_t& cs)
{
    ImageSpec s;
    std::memset(&s, 0, sizeof(s));
    switch (cs.sample) {
    case Vship_SampleFLOAT: case Vship_SampleHALF: case Vship_SampleUINT8: case Vship_SampleUINT9:
    case Vship_SampleUINT10: case Vship_SampleUINT12: case Vship_SampleUINT14: case Vship_SampleUINT16: break;
    default: VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_Sample_t value");
    }
    if (cs.range != Vship_RangeFull && cs.range != Vship_RangeLimited)
        VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_Range_t value");
    if (cs.width <= 0 || cs.height <= 0 || cs.width > (1 << 20) || cs.height > (1 << 20))
        VSHIP_THROW(Vship_NonRGBSInput, "width and height must be in [1, 2^20]");
    if (cs.subsampling.subw < 0 || cs.subsampling.subw > 2 || cs.subsampling.subh < 0 || cs.subsampling.subh > 1)
        VSHIP_THROW(Vship_NonRGBSInput, "chroma subsampling supported: subw in [0, 2], subh in [0, 1]");
    if (cs.chromaLocation < Vship_ChromaLoc_Left || cs.chromaLocation > Vship_ChromaLoc_Top)
        VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_ChromaLocation_t value");
    if (cs.colorFamily != Vship_ColorYUV && cs.colorFamily != Vship_ColorRGB)
        VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_ColorFamily_t value");
    switch (cs.YUVMatrix) {
    case Vship_MATRIX_RGB: case Vship_MATRIX_BT709: case Vship_MATRIX_BT470_BG: case Vship_MATRIX_ST170_M:
    case Vship_MATRIX_BT2020_NCL: case Vship_MATRIX_BT2020_CL: case Vship_MATRIX_BT2100_ICTCP: break;
    default: VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_YUVMatrix_t value");
    }
    switch (cs.transferFunction) {
    case Vship_TRC_BT709: case Vship_TRC_BT470_M: case Vship_TRC_BT470_BG: case Vship_TRC_BT601:
    case Vship_TRC_ST240_M: case Vship_TRC_Linear: case Vship_TRC_sRGB: case Vship_TRC_PQ:
    case Vship_TRC_ST428: case Vship_TRC_HLG: break;
    default: VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_TransferFunction_t value");
    }
    switch (cs.primaries) {
    case Vship_PRIMARIES_INTERNAL: case Vship_PRIMARIES_BT709: case Vship_PRIMARIES_BT470_M:
    case Vship_PRIMARIES_BT470_BG: case Vship_PRIMARIES_ST170_M: case Vship_PRIMARIES_ST240_M:
    case Vship_PRIMARIES_BT2020: break;
    default: VSHIP_THROW(Vship_NonRGBSInput, "unknown Vship_Primaries_t value");
    }
    // resize then crop, reference vshipColor.hpp:46-61,87-103
    if ((cs.target_width != -1 && (cs.target_width < 1 || cs.target_width > (1 << 20))) ||
        (cs.target_height != -1 && (cs.target_height < 1 || cs.target_height > (1 << 20))))
        VSHIP_THROW(Vship_NonRGBSInput, "target_width and target_height must be -1 or in [1, 2^20]");
    const int64_t tw = cs.target_width == -1 ? cs.width : cs.target_width;
    const int64_t th = cs.target_height == -1 ? cs.height : cs.target_height;
    if (cs.crop.top < 0 || cs.crop.bottom < 0 || cs.crop.left < 0 || cs.crop.right < 0 ||
        cs.crop.left + cs.crop.right >= tw || cs.crop.top + cs.crop.bottom >= th)
        VSHIP_THROW(Vship_NonRGBSInput, "crop must be non-negative and leave at least one pixel");
    const bool cropped = cs.crop.top || cs.crop.bottom || cs.crop.left || cs.crop.right;
    s.resample = cropped || tw != cs.width || th != cs.height;
    s.target_w = tw;
    s.target_h = th;
    s.crop_left = cs.crop.left;
    s.crop_top = cs.crop.top;
    s.final_w = tw - cs.crop.left - cs.crop.right;
    s.final_h = th - cs.crop.top - cs.crop.bottom;

    FrontColor& fc = s.fc;
    const int bits = sample_bits(cs.sample);
    fc.sample = cs.sample == Vship_SampleUINT8 ? 0 : (cs.sample == Vship_SampleHALF ? 2 : (cs.sample == Vship_SampleFLOAT ? 3 : 1));
    fc.mask = bits >= 16 ? 0xffffu : ((1u << bits) - 1u);
    s.bytes = cs.sample == Vship_SampleUINT8 ? 1 : (cs.sample == Vship_SampleFLOAT ? 4 : 2);
    fc.subw = cs.subsampling.subw;
    fc.subh = cs.subsampling.subh;
    fc.h_cosited = (cs.chromaLocation == Vship_ChromaLoc_Left || cs.chromaLocation == Vship_ChromaLoc_TopLeft);
    fc.v_cosited = (cs.chromaLocation == Vship_ChromaLoc_Top || cs.chromaLocation == Vship_ChromaLoc_TopLeft);
    fc.transfer = cs.transferFunction;
    phase_weights(fc.h_cosited != 0, fc.hw);
    phase_weights(fc.v_cosited != 0, fc.vw);
    const RangeXform r0 = make_range(cs.range, bits, false);
    const RangeXform r12 = make_range(cs.range, bits, cs.colorFamily == Vship_ColorYUV);
    fc.range0[0] = r0.mul; fc.range0[1] = r0.add;
    fc.range12[0] = r12.mul; fc.range12[1] = r12.add;
    // the reference skips the whole matrix/EOTF step for linear RGB (YUVToLinRGB.hpp:219)
    const int matrix = (cs.transferFunction == Vship_TRC_Linear && cs.YUVMatrix == Vship_MATRIX_RGB) ? Vship_MATRIX_RGB
                                                                                                      : cs.YUVMatrix;
    double pre[9], post[9];
    bool use_pre = false;
    colour_matrices(matrix, cs.primaries, pre, post, &use_pre);
    fc.use_pre = use_pre ? 1 : 0;
    bool ident = true;
    for (int i = 0; i < 9; i++) {
        fc.pre[i] = (float)pre[i];
        fc.post[i] = (float)post[i];
        ident = ident && post[i] == ((i % 4 == 0) ? 1.0 : 0.0);
    }
    fc.use_post = ident ? 0 : 1;
    s.width = cs.width;
    s.height = cs.height;
    s.plane_w[0] = cs.width;
    s.plane_h[0] = cs.height;
    for (int p = 1; p < 3; p++) {
        s.plane_w[p] = ((cs.width - 1) >> fc.subw) + 1;
        s.plane_h[p] = ((cs.height - 1) >> fc.subh) + 1;
    }
    return s;
}

FrontColor linear_rgb_color()
{
    FrontColor fc;
    std::memset(&fc, 0, sizeof(fc));
    fc.sample = kSampleFloat;
    fc.transfer = Vship_TRC_Linear;
    fc.range0[0] = fc.range12[0] = 1.0f;  // v * 1 + 0
    fc.range0[1] = fc.range12[1] = 0.0f;
    return fc;
}

BlurTaps make_blur_taps()
{
    BlurTaps t;
    const float sigma = 1.5f;
    const double tau = 6.28318530718;
    t.integral[0] = 0.0f;
    for (int i = 0; i < 17; i++) {
        const float e = std::exp((float)(-(8 - i) * (8 - i)) / (2 * sigma * sigma));
        t.g[i] = (float)((double)e / std::sqrt(tau * sigma * sigma));
        t.integral[i + 1] = t.integral[i] + t.g[i];
    }
    // the same float expression as the kernels' border case with beg = 0, end = 17
    t.inv_full = 1.0f / (t.integral[17] - t.integral[0]);
    for (int k = -16; k < 32; k++) {
        uint32_t v = 0;
        if (k >= 0 && k < 17) {
            const float gn = t.g[k] * t.inv_full;  // normalised: the band sums to 1
            const _Float16 hi = (_Float16)gn;      // IEEE round to nearest, as on the device
            const _Float16 lo = (_Float16)(gn - (float)hi);
            v = (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
        }
        t.band[k + 16] = v;
    }
    return t;
}

void MapJobList::push_back(const MapJob& m)
{
    if (count >= kMaxJobsPerPair) VSHIP_THROW(Vship_BadErrorType, "more map jobs than a pair has");
    items[count++] = m;
}

PairLayout make_pair_layout(int w, int h, int blur_config)
{
    const BlurTile bt = blur_tile_shape(blur_config);
    PairLayout L;
    std::memset(&L.geom, 0, sizeof(L.geom));
    int64_t cw = w, ch = h, off = 0;
    for (int s = 0; s < kUsedScales; s++) {
        L.geom.w[s] = (int32_t)cw;
        L.geom.h[s] = (int32_t)ch;
        L.geom.pitch[s] = (int32_t)round_up(cw, 64);
        for (int p = 0; p < 3; p++) {
            const bool any = map_active(p, s, 0) || map_active(p, s, 1) || map_active(p, s, 2);
            if (any) {
                L.geom.off[s][p] = off;
                off += (int64_t)L.geom.pitch[s] * ch;
            } else {
                L.geom.off[s][p] = -1;
            }
        }
        cw = half_up(cw);
        ch = half_up(ch);
    }
    L.geom.image_elems = off;
    // the kernels address an image's stored planes with 32-bit byte offsets
    if (off >= ((int64_t)1 << 30))
        VSHIP_THROW(Vship_OutOfVRAM, "image too large: its stored XYB planes exceed 2^30 elements");
    for (int sc = 0; sc < kUsedScales; sc++)
        for (int p = 0; p < 3; p++)
            if ((L.geom.off[sc][p] >= 0) != kStored[sc][p])
                VSHIP_THROW(Vship_BadErrorType, "stored plane set differs from the kernels' table");
    L.front_tiles_x = (w + kFrontTileW - 1) / kFrontTileW;
    L.front_tiles_per_image = L.front_tiles_x * ((h + kFrontTileH - 1) / kFrontTileH);
    int tiles = 0;
    for (int s = 0; s < kUsedScales; s++) {
        for (int p = 0; p < 3; p++) {
            if (L.geom.off[s][p] < 0) continue;
            MapJob m;
            std::memset(&m, 0, sizeof(m));
            m.a = reinterpret_cast<const void*>(L.geom.off[s][p]);
            m.b = reinterpret_cast<const void*>(L.geom.off[s][p]);
            m.w = L.geom.w[s];
            m.h = L.geom.h[s];
            m.pitch = L.geom.pitch[s];
            m.flags = (map_active(p, s, 0) ? kMapSsim : 0) | (map_active(p, s, 1) ? kMapArtifact : 0) |
                      (map_active(p, s, 2) ? kMapDetail : 0);
            m.tiles_x = (m.w + bt.tw - 1) / bt.tw;
            m.ntiles = m.tiles_x * ((m.h + bt.th - 1) / bt.th);
            m.tile_start = tiles;
            m.part_start = tiles;
            m.pair = 0;
            m.feature_base = p * 36 + s * 6;
            tiles += m.ntiles;
            L.maps.push_back(m);
        }
    }
    L.blur_tiles = tiles;
    L.blur_config = blur_config;
    return L;
}

Ssimu2Handler::Ssimu2Handler(const Vship_Colorspace_t& src, const Vship_Colorspace_t& dis, int device)
    : device_(device)
{
    spec_[0] = make_image_spec(src);
    spec_[1] = make_image_spec(dis);
    if (spec_[0].final_w != spec_[1].final_w || spec_[0].final_h != spec_[1].final_h)
        VSHIP_THROW(Vship_DifferingInputType, "reference and distorted sizes differ");
    width_ = (int)spec_[0].final_w;
    height_ = (int)spec_[0].final_h;
    layout_ = make_pair_layout(width_, height_);
    taps_ = make_blur_taps();
    const FrontColor lin_fc = linear_rgb_color();
    front_lds_ = std::max(std::max(frontend_lds_bytes(spec_[0].fc), frontend_lds_bytes(spec_[1].fc)),
                          frontend_lds_bytes(lin_fc));
    // resampled images enter the pyramid as linear RGB: generic front-end
    front_spec_ = (spec_[0].resample || spec_[1].resample) ? kFrontGeneric : front_spec(spec_[0].fc, spec_[1].fc);

    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    VSHIP_HIP_CHECK(hipEventCreateWithFlags(&gather_done_, hipEventDisableTiming));
    auto dmalloc = [](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) VSHIP_THROW(Vship_OutOfVRAM, "hipMalloc failed in SSIMU2 init");
    };
    const size_t elems = (size_t)layout_.geom.image_elems;
    for (int i = 0; i < 2; i++) dmalloc(&xyb_[i], elems * sizeof(float));
    dmalloc((void**)&d_colors_, 3 * sizeof(FrontColor));
    dmalloc((void**)&d_geom_, sizeof(PairGeom));
    dmalloc((void**)&d_maps_, layout_.maps.size() * sizeof(MapJob));
    dmalloc((void**)&d_counts_, kCountSlots * sizeof(int));
    if (hipHostMalloc((void**)&h_partials_, (size_t)layout_.blur_tiles * kPartialStride * sizeof(float),
                      hipHostMallocDefault) != hipSuccess)
        VSHIP_THROW(Vship_OutOfRAM, "hipHostMalloc of the partial sums failed in SSIMU2 init");
    if (hipHostMalloc((void**)&h_front_jobs_, sizeof(FrontJob), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_score_, sizeof(double), hipHostMallocDefault) != hipSuccess)
        VSHIP_THROW(Vship_OutOfRAM, "hipHostMalloc failed in SSIMU2 init");

    // static tables
    MapJobList maps = layout_.maps;
    for (auto& m : maps) {
        const int64_t off = reinterpret_cast<int64_t>(m.a);
        m.a = static_cast<const float*>(xyb_[0]) + off;
        m.b = static_cast<const float*>(xyb_[1]) + off;
    }
    FrontColor colors[3] = {spec_[0].fc, spec_[1].fc, lin_fc};
    int counts[kCountSlots] = {};
    counts[kCountFrontTiles] = layout_.front_tiles_per_image;
    counts[kCountFrontJobs] = 1;
    counts[kCountBlurTiles] = layout_.blur_tiles;
    counts[kCountBlurJobs] = (int)maps.size();
    counts[kCountPairs] = 1;
    VSHIP_HIP_CHECK(hipMemcpy(d_maps_, maps.data(), maps.size() * sizeof(MapJob), hipMemcpyHostToDevice));
    VSHIP_HIP_CHECK(hipMemcpy(d_colors_, colors, sizeof(colors), hipMemcpyHostToDevice));
    VSHIP_HIP_CHECK(hipMemcpy(d_geom_, &layout_.geom, sizeof(PairGeom), hipMemcpyHostToDevice));
    VSHIP_HIP_CHECK(hipMemcpy(d_counts_, counts, sizeof(counts), hipMemcpyHostToDevice));
    FrontJob& j = *h_front_jobs_;
    std::memset(&j, 0, sizeof(j));
    for (int i = 0; i < 2; i++) {
        j.out[i] = xyb_[i];
        j.color[i] = i;
    }
    j.width = width_;
    j.height = height_;
    set_front_tiles(j, layout_.front_tiles_x);
    j.tile_start = 0;
    j.pair = 0;

    // crop/resize path: linear RGB at the input size, resized, then the
    // pyramid kernel reads the cropped window as float linear RGB
    if (spec_[0].resample || spec_[1].resample) {
        dmalloc((void**)&d_lin_jobs_, 2 * sizeof(FrontJob));
        if (hipHostMalloc((void**)&h_lin_jobs_, 2 * sizeof(FrontJob), hipHostMallocDefault) != hipSuccess)
            VSHIP_THROW(Vship_OutOfRAM, "hipHostMalloc failed in SSIMU2 init");
        std::memset(h_lin_jobs_, 0, 2 * sizeof(FrontJob));
    }
    for (int i = 0; i < 2; i++) {
        const ImageSpec& sp = spec_[i];
        if (!sp.resample) continue;
        const int64_t W = sp.width, H = sp.height, tw = sp.target_w, th = sp.target_h;
        dmalloc((void**)&lin_[i], (size_t)(3 * W * H) * sizeof(float));
        const float* cur = lin_[i];
        if (tw != W) {
            dmalloc((void**)&rsz_h_[i], (size_t)(3 * tw * H) * sizeof(float));
            cur = rsz_h_[i];
        }
        if (th != H) {
            dmalloc((void**)&rsz_[i], (size_t)(3 * tw * th) * sizeof(float));
            cur = rsz_[i];
        }
        FrontJob& lj = h_lin_jobs_[i];
        lj.color[0] = i;
        lj.width = (int32_t)W;
        lj.height = (int32_t)H;
        set_front_tiles(lj, (int)((W + kFrontTileW - 1) / kFrontTileW));
        for (int p = 0; p < 3; p++) {
            j.plane[i][p] = reinterpret_cast<const uint8_t*>(cur + p * tw * th + sp.crop_top * tw + sp.crop_left);
            j.stride[i][p] = tw * (int64_t)sizeof(float);
        }
        j.color[i] = 2;
    }
}

Ssimu2Handler::~Ssimu2Handler()
{
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (gather_done_) (void)hipEventDestroy(gather_done_);
    for (int i = 0; i < 2; i++) {
        (void)hipFree(xyb_[i]);
        (void)hipFree(lin_[i]);
        (void)hipFree(rsz_h_[i]);
        (void)hipFree(rsz_[i]);
        for (int p = 0; p < 3; p++) (void)hipFree(staging_[i][p]);
    }
    (void)hipFree(d_lin_jobs_);
    if (h_lin_jobs_) (void)hipHostFree(h_lin_jobs_);
    (void)hipFree(d_colors_);
    (void)hipFree(d_geom_);
    (void)hipFree(d_maps_);
    (void)hipFree(d_counts_);
    (void)hipHostFree(h_front_jobs_);
    (void)hipHostFree(h_score_);
    if (h_partials_) (void)hipHostFree(h_partials_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Ssimu2Handler::ensure_staging(int image, const int64_t ls[3])
{
    const ImageSpec& s = spec_[image];
    for (int p = 0; p < 3; p++) {
        const int64_t row = s.plane_w[p] * s.bytes;
        if (ls[p] < row) VSHIP_THROW(Vship_BadPointer, "lineSize smaller than one row of samples");
        const size_t need = (size_t)(ls[p] * s.plane_h[p]);
        if (need > staging_cap_[image][p]) {
            (void)hipFree(staging_[image][p]);
            staging_[image][p] = nullptr;
            staging_cap_[image][p] = 0;
            if (hipMalloc((void**)&staging_[image][p], need) != hipSuccess)
                VSHIP_THROW(Vship_OutOfVRAM, "hipMalloc of the input staging buffer failed");
            staging_cap_[image][p] = need;
        }
    }
}

int Ssimu2Handler::call_spec() const
{
    if (front_spec_ == kFrontGeneric) return kFrontGeneric;
    return front_offsets_fit(*h_front_jobs_, spec_[0].fc, spec_[1].fc) ? front_spec_ : kFrontGeneric;
}

// FP64 completion of one pair's score from its blur items' partial sums
// (reference src/ssimu2/score.hpp:315-320,422-460), on the host, with the
// operations and order of the GPU's score kernel: per map job, lane k of a
// wave sums items k, k + 64, ... and the lanes combine by the xor butterfly of
// its wave reductions (lane m's result is feature m's); the weighted sum takes
// features k and k + 64 per lane, then the butterfly; the polynomial has no
// contraction on either side. The features are therefore the same doubles
// as the batched path's (whose score kernel runs this on the GPU).
// Lane `lane`'s value after K stages (offsets 32, 16, ... of the 64-lane xor
// butterfly v[i] += v[i ^ off]): the same additions as the wave reduction,
// evaluated for one lane only.
template <int K>
static double butterfly_lane(const double v[64], int lane)
{
    if constexpr (K == 0) {
        return v[lane];
    } else {
        constexpr int off = 64 >> K;
        return butterfly_lane<K - 1>(v, lane) + butterfly_lane<K - 1>(v, lane ^ off);
    }
}

double score_from_partials(const MapJob* maps, int nmaps, const float* partials, double feat[108])
{
#pragma clang fp contract(off)
    static const double kWeights[108] = VSHIP_SSIMU2_WEIGHTS;
    for (int i = 0; i < 108; i++) feat[i] = 0.0;
    for (int jn = 0; jn < nmaps; jn++) {
        const MapJob& job = maps[jn];
        double acc[6][64] = {};
        for (int k = 0; k < job.ntiles; k++) {
            const float* p = partials + (size_t)(job.part_start + k) * kPartialStride;
            for (int m = 0; m < 6; m++) acc[m][k & 63] += (double)p[m];
        }
        const double n = (double)job.w * (double)job.h;
        for (int m = 0; m < 3; m++) {
            const int flag = m == 0 ? kMapSsim : (m == 1 ? kMapArtifact : kMapDetail);
            if (!(job.flags & flag)) continue;
            feat[job.feature_base + m] = butterfly_lane<6>(acc[m], m) / n;
            feat[job.feature_base + 3 + m] = std::sqrt(std::sqrt(butterfly_lane<6>(acc[3 + m], m) / n));
        }
    }
    double s[64];
    for (int l = 0; l < 64; l++) {
        s[l] = l < 108 ? kWeights[l] * feat[l] : 0.0;
        if (l + 64 < 108) s[l] = std::fma(kWeights[l + 64], feat[l + 64], s[l]);
    }
    double x = butterfly_lane<6>(s, 0) * 0.9562382616834844;
    x = (6.248496625763138e-5 * x * x) * x + 2.326765642916932 * x - 0.020884521182843837 * x * x;
    return (x > 0.0) ? std::pow(x, 0.6276336467831387) * -10.0 + 100.0 : 100.0;
}

void Ssimu2Handler::launch_all()
{
    if (d_lin_jobs_) {
        VSHIP_HIP_CHECK(
            hipMemcpyAsync(d_lin_jobs_, h_lin_jobs_, 2 * sizeof(FrontJob), hipMemcpyHostToDevice, stream_));
        for (int i = 0; i < 2; i++) {
            const ImageSpec& sp = spec_[i];
            if (!sp.resample) continue;
            const int W = (int)sp.width, H = (int)sp.height, tw = (int)sp.target_w, th = (int)sp.target_h;
            const int tiles = h_lin_jobs_[i].tiles_x * ((H + kFrontTileH - 1) / kFrontTileH);
            VSHIP_HIP_CHECK(launch_linear_rgb(d_lin_jobs_ + i, tiles, d_colors_, lin_[i], W, front_lds_, stream_));
            const float* cur = lin_[i];
            if (tw != W) {
                VSHIP_HIP_CHECK(launch_resize_h(rsz_h_[i], cur, W, H, tw, stream_));
                cur = rsz_h_[i];
            }
            if (th != H) VSHIP_HIP_CHECK(launch_resize_v(rsz_[i], cur, tw, H, th, stream_));
        }
    }
    // the job (plane pointers of this call) travels as a kernel argument
    VSHIP_HIP_CHECK(launch_frontend_single<float>(*h_front_jobs_, layout_.front_tiles_per_image, d_colors_, d_geom_,
                                                  call_spec(), stream_));
    // the items' partial sums go straight to pinned host memory; run() forms
    // the score from them (score_from_partials), so a call is two kernel launches
    VSHIP_HIP_CHECK(launch_blur<float>(layout_.blur_config, d_maps_, d_counts_, blur_grid_for(layout_.blur_tiles),
                                       taps_, h_partials_, stream_));
}

void Ssimu2Handler::run(const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3],
                        const int64_t ls2[3])
{
    enqueue(p1, p2, ls1, ls2);
    complete();
}

void Ssimu2Handler::enqueue(const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3],
                            const int64_t ls2[3])
{
    const uint8_t* const* planes[2] = {p1, p2};
    const int64_t* strides[2] = {ls1, ls2};
    for (int i = 0; i < 2; i++) {
        // resampled images enter through the linear-RGB job; the pyramid job
        // keeps its fixed pointers to the resized planes
        FrontJob& dst = spec_[i].resample ? h_lin_jobs_[i] : *h_front_jobs_;
        const int k = spec_[i].resample ? 0 : i;
        for (int p = 0; p < 3; p++) {
            dst.plane[k][p] = planes[i][p];
            dst.stride[k][p] = strides[i][p];
        }
    }
    launch_all();
}

void Ssimu2Handler::complete()
{
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    *h_score_ = score_from_partials(layout_.maps.data(), layout_.maps.count, h_partials_, h_features_);
    have_run_ = true;
}

namespace {

// Host-plane reads in flight per device: a call holds a slot from its gather
// launch until the gather has completed (its later kernels run without it).
// More concurrent readers of pinned host memory share the one PCIe link and
// lower its rate (8 handlers read 2-5% slower than 4), so handlers beyond
// kMaxConcurrentGathers wait for a slot before reading.
constexpr int kMaxConcurrentGathers = 4;
constexpr int kGateDevices = 64;
std::mutex g_gather_mutex;
std::condition_variable g_gather_cv;
int g_gathers[kGateDevices] = {};

class GatherSlot {
public:
    explicit GatherSlot(int device) : device_(device >= 0 && device < kGateDevices ? device : -1)
    {
        if (device_ < 0) return;
        std::unique_lock<std::mutex> lock(g_gather_mutex);
        g_gather_cv.wait(lock, [&] { return g_gathers[device_] < kMaxConcurrentGathers; });
        g_gathers[device_]++;
    }
    ~GatherSlot()
    {
        if (device_ < 0) return;
        {
            std::lock_guard<std::mutex> lock(g_gather_mutex);
            g_gathers[device_]--;
        }
        g_gather_cv.notify_all();
    }
    GatherSlot(const GatherSlot&) = delete;
    GatherSlot& operator=(const GatherSlot&) = delete;

private:
    int device_;
};

// Device address of host bytes [p, p + bytes) when they lie in pinned host
// memory the device can read (both ends registered, one contiguous mapping,
// 16-byte aligned); null otherwise (pageable memory: the copy engines take
// it). A failed query leaves no error for the next HIP call.
const uint8_t* pinned_device_view(const uint8_t* p, size_t bytes)
{
    if (bytes == 0) return nullptr;
    hipPointerAttribute_t a, b;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || a.devicePointer == nullptr) return nullptr;
    if (hipPointerGetAttributes(&b, p + bytes - 1) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    const uint8_t* d = static_cast<const uint8_t*>(a.devicePointer);
    if (b.type != hipMemoryTypeHost || static_cast<const uint8_t*>(b.devicePointer) != d + bytes - 1) return nullptr;
    if (reinterpret_cast<uintptr_t>(d) % 16 != 0) return nullptr;
    return d;
}

}  // namespace

double Ssimu2Handler::compute_host(const uint8_t* const src1[3], const uint8_t* const src2[3], const int64_t ls1[3],
                                   const int64_t ls2[3])
{
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    const uint8_t* const* srcs[2] = {src1, src2};
    const int64_t* lss[2] = {ls1, ls2};
    const uint8_t* dev[2][3];
    size_t bytes[2][3];
    // planes in pinned host memory: one gather kernel reads all six over
    // PCIe; otherwise one stream-ordered copy per plane
    GatherPlanes g;
    std::memset(&g, 0, sizeof(g));
    bool direct = true;
    for (int i = 0; i < 2; i++) {
        ensure_staging(i, lss[i]);
        const ImageSpec& s = spec_[i];
        for (int p = 0; p < 3; p++) {
            if (srcs[i][p] == nullptr) VSHIP_THROW(Vship_BadPointer, "null plane pointer");
            // same extent as the reference's upload (vshipColor.hpp:96)
            bytes[i][p] = (size_t)(lss[i][p] * s.plane_h[p] - (lss[i][p] - s.bytes * s.plane_w[p]));
            dev[i][p] = staging_[i][p];
            const uint8_t* view = direct ? pinned_device_view(srcs[i][p], bytes[i][p]) : nullptr;
            if (view == nullptr) {
                direct = false;
            } else {
                g.src[g.n] = view;
                g.dst[g.n] = staging_[i][p];
                g.bytes[g.n] = (int64_t)bytes[i][p];
                g.n++;
            }
        }
    }
    if (direct) {
        GatherSlot slot(device_);
        VSHIP_HIP_CHECK(launch_gather(g, stream_));
        VSHIP_HIP_CHECK(hipEventRecord(gather_done_, stream_));
        enqueue(dev[0], dev[1], ls1, ls2);
        VSHIP_HIP_CHECK(hipEventSynchronize(gather_done_));
    } else {
        for (int i = 0; i < 2; i++)
            for (int p = 0; p < 3; p++)
                VSHIP_HIP_CHECK(
                    hipMemcpyAsync(staging_[i][p], srcs[i][p], bytes[i][p], hipMemcpyHostToDevice, stream_));
        enqueue(dev[0], dev[1], ls1, ls2);
    }
    complete();
    return *h_score_;
}

double Ssimu2Handler::compute_device(const uint8_t* const src1[3], const uint8_t* const src2[3], const int64_t ls1[3],
                                     const int64_t ls2[3])
{
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    for (int p = 0; p < 3; p++) {
        if (src1[p] == nullptr || src2[p] == nullptr) VSHIP_THROW(Vship_BadPointer, "null plane pointer");
        if (ls1[p] < spec_[0].plane_w[p] * spec_[0].bytes || ls2[p] < spec_[1].plane_w[p] * spec_[1].bytes)
            VSHIP_THROW(Vship_BadPointer, "lineSize smaller than one row of samples");
    }
    run(src1, src2, ls1, ls2);
    return *h_score_;
}

void Ssimu2Handler::last_features(double out[108])
{
    std::memcpy(out, h_features_, 108 * sizeof(double));
}

void Ssimu2Handler::debug_plane(int image, int plane, int scale, float* out, int64_t* w, int64_t* h)
{
    if (image < 0 || image > 1 || plane < 0 || plane > 2 || scale < 0 || scale >= kUsedScales ||
        layout_.geom.off[scale][plane] < 0)
        VSHIP_THROW(Vship_BadPointer, "no such stored XYB plane");
    *w = layout_.geom.w[scale];
    *h = layout_.geom.h[scale];
    if (!out) return;
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    const float* src = static_cast<const float*>(xyb_[image]) + layout_.geom.off[scale][plane];
    VSHIP_HIP_CHECK(hipMemcpy2D(out, *w * sizeof(float), src, layout_.geom.pitch[scale] * sizeof(float),
                                *w * sizeof(float), *h, hipMemcpyDeviceToHost));
}

void Ssimu2Handler::profile(int iters, double* ms_front, double* ms_blur, double* ms_score)
{
    if (!have_run_) VSHIP_THROW(Vship_BadPointer, "profile needs one compute call first");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    hipEvent_t ev[3];
    for (auto& e : ev) VSHIP_HIP_CHECK(hipEventCreate(&e));
    double acc[3] = {0, 0, 0};
    double feat[108];
    for (int it = 0; it < iters; it++) {
        VSHIP_HIP_CHECK(hipEventRecord(ev[0], stream_));
        VSHIP_HIP_CHECK(launch_frontend_single<float>(*h_front_jobs_, layout_.front_tiles_per_image, d_colors_,
                                                      d_geom_, call_spec(), stream_));
        VSHIP_HIP_CHECK(hipEventRecord(ev[1], stream_));
        VSHIP_HIP_CHECK(launch_blur<float>(layout_.blur_config, d_maps_, d_counts_, blur_grid_for(layout_.blur_tiles),
                                           taps_, h_partials_, stream_));
        VSHIP_HIP_CHECK(hipEventRecord(ev[2], stream_));
        VSHIP_HIP_CHECK(hipEventSynchronize(ev[2]));
        // the score's FP64 completion on the host, as in run()
        const auto t0 = std::chrono::steady_clock::now();
        (void)score_from_partials(layout_.maps.data(), layout_.maps.count, h_partials_, feat);
        acc[2] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int k = 0; k < 2; k++) {
            float ms = 0.0f;
            VSHIP_HIP_CHECK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_front = acc[0] / iters;
    *ms_blur = acc[1] / iters;
    *ms_score = acc[2] / iters;
}

void Ssimu2Handler::work(double* bytes, double* flops) const
{
    double in_bytes = 0.0;
    for (int i = 0; i < 2; i++)
        for (int p = 0; p < 3; p++) in_bytes += (double)spec_[i].plane_w[p] * spec_[i].plane_h[p] * spec_[i].bytes;
    double active = 0.0, fl = 0.0, blocks = 0.0;
    for (const auto& m : layout_.maps) {
        const double px = (double)m.w * m.h;
        active += px;
        const bool ssim = m.flags & kMapSsim;
        const bool edge = m.flags & (kMapArtifact | kMapDetail);
        const int n_edge = ((m.flags & kMapArtifact) ? 1 : 0) + ((m.flags & kMapDetail) ? 1 : 0);
        fl += px * (70.0 * (ssim ? 4 : 2) + 23.0 * ssim + 10.0 * edge + 4.0 * n_edge);
        blocks += std::ceil(m.w / 16.0) * std::ceil(m.h / 16.0);
    }
    // 2 * in_bytes + 4 * B * A + 48 * nblocks with B = 4 (float planes)
    *bytes = in_bytes + 4.0 * 4.0 * active + 48.0 * blocks;
    *flops = fl;
}

}  // namespace vship
