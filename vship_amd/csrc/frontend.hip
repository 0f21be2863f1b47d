// frontend.hip - fused colour front-end + 5-level pyramid + XYB for gfx950.
//
// One workgroup of 512 threads owns a 64x32 level-0 tile of both images of a
// pair; thread (tx, ty) owns the 2x2 block at (2tx, 2ty). Per tile and image:
//   raw samples -> range (rangeToFull.hpp:24-55) -> chroma Catmull-Rom
//   upsampling through LDS (chromaUpsample.hpp:28-247) -> matrix, EOTF,
//   linear-light matrix and primaries (YUVToLinRGB.hpp:157-181,
//   primaries.hpp:134-158; the matrices are folded on the host into one 3x3
//   before and one after the EOTF) -> linear-RGB 2x2 box pyramid with the
//   edge clamp (downsample.hpp:3-19): level 1 in registers, 2 to 4 through
//   LDS -> positive XYB (makeXYB.hpp:6-64) of exactly the planes the
//   score uses.
// A 64x32 level-0 tile maps onto whole 4x2 level-4 blocks, and the edge clamp
// of every level only reads pixels inside the same tile, so tiles are
// independent. Both images' XYB values stay in registers until the tile is
// stored, which the FP16 layout needs (distorted = fp16(b - a) of the FP32
// values). Only the raw-sample loads are specialised per sample type and
// the EOTF switch sits outside the 16-pixel loop, which keeps the kernel small
// enough for the instruction cache.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "kernels.hpp"

#include <type_traits>

#define VSHIP_INLINE __device__ __attribute__((always_inline)) inline

namespace vship {

namespace {

VSHIP_INLINE int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }  // lo <= hi

// Thread block of level-0 pixels: kBW columns x kBH rows per thread.
constexpr int kBW = 2, kBH = 2;
constexpr int kTX = kFrontTileW / kBW;  // threads along x (32)
static_assert(kTX * (kFrontTileH / kBH) == kFrontThreads, "front-end tile and block shape");

// Catmull-Rom weights of the four samples around segment parameter t: the
// cubic Hermite form of resize.hpp:5-24 expanded (same polynomial).
struct CrWeights {
    float w0, w1, w2, w3;
};

VSHIP_INLINE CrWeights cr_weights(float t)
{
    const float t2 = t * t, t3 = t2 * t;
    CrWeights w;
    w.w0 = 0.5f * (-t3 + 2.0f * t2 - t);
    w.w1 = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
    w.w2 = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
    w.w3 = 0.5f * (t3 - t2);
    return w;
}

VSHIP_INLINE float cr_apply(const CrWeights& w, float pm1, float p0, float p1, float p2)
{
    return fmaf(w.w0, pm1, fmaf(w.w1, p0, fmaf(w.w2, p1, w.w3 * p2)));
}

// Sample p0 itself on a cosited position (copy), else the interpolation.
// The four taps are read unconditionally (they lie inside the staged halo
// either way) and the result selected, so no lane diverges.
VSHIP_INLINE float cr_pick(bool copy, const CrWeights& w, const float* p, int pitch)
{
    const float pm1 = p[-pitch], p0 = p[0], p1 = p[pitch], p2 = p[2 * pitch];
    const float v = cr_apply(w, pm1, p0, p1, p2);
    return copy ? p0 : v;
}

// Segment of full-resolution coordinate X for decimation 2^sub (sub >= 1):
// X maps to segment [k, k+1] at parameter t. Cosited: sample k sits on
// X = k * 2^sub (kernels *LeftUpscale*, *TopUpscale*); centred: on
// X = k * 2^sub + 2^(sub-1) - 0.5 (kernels *CenterUpscale*),
// chromaUpsample.hpp:54-149.
VSHIP_INLINE void seg_of(int X, int sub, bool cosited, int& k, float& t, bool& copy)
{
    const int mask = (1 << sub) - 1;
    const float inv = sub == 1 ? 0.5f : 0.25f;
    if (cosited) {
        k = X >> sub;
        const int r = X & mask;
        copy = (r == 0);
        t = (float)r * inv;
    } else {
        const int off = 1 << (sub - 1);
        k = (X - off) >> sub;  // arithmetic shift: X < off gives k = -1
        const int r = (X - off) & mask;
        copy = false;
        t = ((float)r + 0.5f) * inv;
    }
}

typedef const __attribute__((address_space(1))) uint8_t GByte;
typedef const __attribute__((address_space(1))) uint16_t GU16;
typedef const __attribute__((address_space(1))) uint32_t GU32;

// One raw sample as float. Plane pointers come from job tables, so they are
// read through the global address space (flat loads would also count against
// the LDS wait counter). Samples of planes whose base and row stride are
// multiples of the sample size are read with one load, others byte by byte.
template <int KIND>
VSHIP_INLINE uint32_t raw_kind(const uint8_t* p8, bool aligned)
{
    GByte* p = (GByte*)p8;
    if constexpr (KIND == kSampleU8) {
        return (uint32_t)p[0];
    } else if constexpr (KIND == kSampleU16 || KIND == kSampleHalf) {
        return aligned ? (uint32_t)*(GU16*)p : ((uint32_t)p[0] | ((uint32_t)p[1] << 8));
    } else {
        return aligned ? *(GU32*)p
                       : ((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
    }
}

template <int KIND>
VSHIP_INLINE float convert_kind(uint32_t v, uint32_t mask)
{
    if constexpr (KIND == kSampleU8) return (float)v;
    else if constexpr (KIND == kSampleU16) return (float)(v & mask);
    else if constexpr (KIND == kSampleHalf) return __half2float(__ushort_as_half((uint16_t)v));
    else return __uint_as_float(v);
}

template <int KIND>
VSHIP_INLINE float load_kind(const uint8_t* p8, uint32_t mask, bool aligned)
{
    return convert_kind<KIND>(raw_kind<KIND>(p8, aligned), mask);
}

template <int KIND>
constexpr int kind_bytes();

// Whether every sample of a plane (base + y * stride + x * bytes) is aligned
// to its size.
template <int KIND>
VSHIP_INLINE bool plane_aligned(const uint8_t* base, int64_t stride)
{
    return (((uintptr_t)base | (uintptr_t)stride) & (uintptr_t)(kind_bytes<KIND>() - 1)) == 0;
}

template <int KIND>
constexpr int kind_bytes() { return KIND == kSampleU8 ? 1 : (KIND == kSampleFloat ? 4 : 2); }

// Address of the sample `xb` bytes into row `row` of a plane. In general a
// 64-bit product; with O32 the host has checked (front_offsets_fit) that the
// stride is below 2^24 and that every row of the plane starts within 2^32
// bytes of its base, so the whole offset is one full-rate 24-bit
// multiply-add and the load takes the plane base as its scalar address (the
// 64-bit product takes several quarter-rate instructions).
template <bool O32>
VSHIP_INLINE const uint8_t* sample_ptr(const uint8_t* base, int row, int64_t stride, int xb)
{
    if constexpr (O32) return base + (__umul24((uint32_t)row, (uint32_t)stride) + (uint32_t)xb);
    else return base + (int64_t)row * stride + xb;
}

// kBH x kBW block of one plane at full resolution (clamped at the image edge).
template <int KIND, bool O32 = false>
VSHIP_INLINE void load_block(const uint8_t* base, int64_t stride, int x0, int y0, int W, int H, uint32_t mask,
                             const RangeXform& rx, float out[kBH][kBW])
{
    const bool aligned = plane_aligned<KIND>(base, stride);
#pragma unroll
    for (int j = 0; j < kBH; j++) {
#pragma unroll
        for (int i = 0; i < kBW; i++)
            out[j][i] = apply_range(load_kind<KIND>(sample_ptr<O32>(base, min(y0 + j, H - 1), stride,
                                                                    min(x0 + i, W - 1) * kind_bytes<KIND>()),
                                                    mask, aligned),
                                    rx);
    }
}

// Both chroma planes of the tile (with halo) into LDS, clamped coordinates.
// Every thread issues all of its loads (at most kChromaIters per plane: the
// largest staged plane is 68 x 20 or 36 x 36 samples) before the first one is
// used, so the tile waits for one memory round trip instead of one per row.
constexpr int kChromaIters = 3;
static_assert(kChromaIters * kFrontThreads >= 68 * 20 && kChromaIters * kFrontThreads >= 36 * 36,
              "chroma staging items per thread");

template <int KIND, bool O32 = false>
VSHIP_INLINE void stage_chroma(const uint8_t* const* plane, const int64_t* pstride, uint32_t mask,
                               const RangeXform& rx, float* s_chroma, int ccols, int crows, int cx0, int cy0, int cw,
                               int ch)
{
    const int t = threadIdx.x;
    const int plane_elems = ccols * crows;
    // row = e / ccols from a float reciprocal: (e + 0.5) / ccols lies at least
    // 0.5 / ccols (>= 0.007) from an integer, far beyond the float error
    const float inv_cols = 1.0f / (float)ccols;
    // every load is unconditional (items past the plane reload its last
    // sample) and the raw words are converted only after all loads are
    // issued, so no load waits for an earlier one
    uint32_t v[2][kChromaIters];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const uint8_t* base = plane[1 + p];
        const int64_t stride = pstride[1 + p];
        const bool aligned = plane_aligned<KIND>(base, stride);
#pragma unroll
        for (int it = 0; it < kChromaIters; it++) {
            if (it * kFrontThreads >= plane_elems) break;  // uniform (constant in the layout instantiations)
            const int e = min(t + it * kFrontThreads, plane_elems - 1);
            const int r = (int)(((float)e + 0.5f) * inv_cols);
            const int c = e - r * ccols;
            v[p][it] = raw_kind<KIND>(
                sample_ptr<O32>(base, clampi(cy0 + r, 0, ch - 1), stride, clampi(cx0 + c, 0, cw - 1) * kind_bytes<KIND>()),
                aligned);
        }
    }
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int it = 0; it < kChromaIters; it++) {
            const int e = t + it * kFrontThreads;
            if (e < plane_elems) s_chroma[p * plane_elems + e] = apply_range(convert_kind<KIND>(v[p][it], mask), rx);
        }
}

// Raw samples of one image of a tile, loaded ahead of their use (the batched
// front-end issues an image's loads while the previous image is computed):
// r[0..3] the thread's 2x2 luma block; then either the 2x2 blocks of planes 1
// and 2 (no chroma subsampling) or the thread's chroma staging samples
// (r[4 + 3p + it] = item it of chroma plane p).
constexpr int kRawWords = 12;
struct RawTile {
    uint32_t w[kRawWords];
};

template <int KIND, bool O32>
VSHIP_INLINE void raw_tile_kind(const FrontJob& job, int k, const FrontColor& fc, int tile_x, int tile_y,
                                RawTile& rt)
{
    uint32_t* r = rt.w;
    const uint8_t* const* plane = job.plane[k];
    const int64_t* pstride = job.stride[k];
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int W = job.width, H = job.height;
    const int x0 = tile_x * kFrontTileW + kBW * tx;
    const int y0 = tile_y * kFrontTileH + kBH * ty;
    const bool sub = fc.subw || fc.subh;
    // the thread's 2x2 block of plane p into r[4p .. 4p + 3] (p is a constant
    // at every call, so the words stay in registers)
    auto block = [&](const int p) {
        const bool aligned = plane_aligned<KIND>(plane[p], pstride[p]);
#pragma unroll
        for (int j = 0; j < kBH; j++) {
#pragma unroll
            for (int i = 0; i < kBW; i++)
                r[4 * p + j * kBW + i] = raw_kind<KIND>(
                    sample_ptr<O32>(plane[p], min(y0 + j, H - 1), pstride[p], min(x0 + i, W - 1) * kind_bytes<KIND>()),
                    aligned);
        }
    };
    block(0);
    if (!sub) {
        block(1);
        block(2);
    }
    if (sub) {
        const int ccols = (kFrontTileW >> fc.subw) + 4, crows = (kFrontTileH >> fc.subh) + 4;
        const int cx0 = ((tile_x * kFrontTileW) >> fc.subw) - 2, cy0 = ((tile_y * kFrontTileH) >> fc.subh) - 2;
        const int cw = ((W - 1) >> fc.subw) + 1, ch = ((H - 1) >> fc.subh) + 1;
        const int plane_elems = ccols * crows;
        const float inv_cols = 1.0f / (float)ccols;
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const bool aligned = plane_aligned<KIND>(plane[1 + p], pstride[1 + p]);
#pragma unroll
            for (int it = 0; it < kChromaIters; it++) {
                // every item of the thread's rounds with items in the plane
                // loads (items past the plane reload its last sample)
                if (it * kFrontThreads >= plane_elems) break;
                const int e = min(t + it * kFrontThreads, plane_elems - 1);
                const int rr = (int)(((float)e + 0.5f) * inv_cols);
                const int c = e - rr * ccols;
                r[4 + 3 * p + it] = raw_kind<KIND>(sample_ptr<O32>(plane[1 + p], clampi(cy0 + rr, 0, ch - 1),
                                                                   pstride[1 + p],
                                                                   clampi(cx0 + c, 0, cw - 1) * kind_bytes<KIND>()),
                                                   aligned);
            }
        }
        // both branches store words 4..11 in the same order, so that merged
        // stores keep constant word indices and the words stay in registers
        r[10] = 0;
        r[11] = 0;
    }
}

template <bool O32>
VSHIP_INLINE RawTile raw_tile(const FrontJob& job, int k, const FrontColor& fc, int tile_x, int tile_y)
{
    RawTile r;
    switch (fc.sample) {
    case kSampleU8: raw_tile_kind<kSampleU8, O32>(job, k, fc, tile_x, tile_y, r); break;
    case kSampleU16: raw_tile_kind<kSampleU16, O32>(job, k, fc, tile_x, tile_y, r); break;
    case kSampleHalf: raw_tile_kind<kSampleHalf, O32>(job, k, fc, tile_x, tile_y, r); break;
    default: raw_tile_kind<kSampleFloat, O32>(job, k, fc, tile_x, tile_y, r); break;
    }
    return r;
}

// Plane values of the thread's block (range applied) and the staged chroma
// samples in LDS, from raw samples.
template <int KIND>
VSHIP_INLINE void unpack_raw_kind(const FrontColor& fc, const RawTile& rt, float* s_chroma, float P[3][kBH][kBW])
{
    const uint32_t* r = rt.w;
    const RangeXform r0 = {fc.range0[0], fc.range0[1]};
    const RangeXform r12 = {fc.range12[0], fc.range12[1]};
    const bool sub = fc.subw || fc.subh;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++)
                if (p == 0 || !sub)
                    P[p][j][i] = apply_range(convert_kind<KIND>(r[4 * p + j * kBW + i], fc.mask), p == 0 ? r0 : r12);
    if (sub) {
        const int t = threadIdx.x;
        const int plane_elems = ((kFrontTileW >> fc.subw) + 4) * ((kFrontTileH >> fc.subh) + 4);
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int it = 0; it < kChromaIters; it++) {
                const int e = t + it * kFrontThreads;
                if (e < plane_elems)
                    s_chroma[p * plane_elems + e] = apply_range(convert_kind<KIND>(r[4 + 3 * p + it], fc.mask), r12);
            }
    }
}

VSHIP_INLINE void unpack_raw(const FrontColor& fc, const RawTile& r, float* s_chroma, float P[3][kBH][kBW])
{
    switch (fc.sample) {
    case kSampleU8: unpack_raw_kind<kSampleU8>(fc, r, s_chroma, P); break;
    case kSampleU16: unpack_raw_kind<kSampleU16>(fc, r, s_chroma, P); break;
    case kSampleHalf: unpack_raw_kind<kSampleHalf>(fc, r, s_chroma, P); break;
    default: unpack_raw_kind<kSampleFloat>(fc, r, s_chroma, P); break;
    }
}

// EOTF of the thread's block, one code path per transfer.
template <int TRC>
VSHIP_INLINE void eotf_block(float P[3][kBH][kBW])
{
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++) P[c][j][i] = eotf(P[c][j][i], TRC);
}

VSHIP_INLINE void apply_transfer(int trc, float P[3][kBH][kBW])
{
    switch (trc) {
    case 1: eotf_block<1>(P); break;
    case 4: eotf_block<4>(P); break;
    case 5: case 6: eotf_block<5>(P); break;
    case 7: eotf_block<7>(P); break;
    case 13: eotf_block<13>(P); break;
    case 16: eotf_block<16>(P); break;
    case 17: eotf_block<17>(P); break;
    case 18:
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++) eotf_hlg(P[0][j][i], P[1][j][i], P[2][j][i]);
        break;
    default: break;  // linear
    }
}

// XYB values of one image's tile in the thread's ownership (stored planes only).
struct TileXyb {
    float y0[kBH][kBW];              // level 0: Y of the thread's block
    float x1[kBH / 2], y1[kBH / 2];  // level 1: X, Y of the block's parents
    float x2[2] = {}, y2[2] = {};    // level 2: X, Y of two adjacent pixels (wave 0)
    float l3[3] = {};                // level 3 or 4 (some lanes of wave 0): X, Y, B
};

// Flat field index of a TileXyb value.
constexpr int kFieldY0 = 0;
constexpr int kFieldX1 = kFieldY0 + kBH * kBW;
constexpr int kFieldY1 = kFieldX1 + kBH / 2;
constexpr int kFieldX2 = kFieldY1 + kBH / 2;
constexpr int kFieldY2 = kFieldX2 + 2;
constexpr int kFieldL3 = kFieldY2 + 2;
constexpr int kFields = kFieldL3 + 3;
// The FP16 path's LDS stash is field-major, one float per thread and field;
// the second level-2 pixel of thread t < 64 takes thread t + 64's slot of the
// same row (threads from 64 on hold no level-2 values), so the stash has two
// rows fewer than there are fields.
constexpr int kStashRows = kFields - 2;
static_assert(kStashRows * kFrontThreads == kFrontStashFloats, "front-end stash size");
static_assert(kFrontThreads >= 128, "level-2 pairs in the stash");

VSHIP_INLINE int stash_slot(int f, int t)
{
    if (f < kFieldX2) return f * kFrontThreads + t;
    if (f < kFieldL3) return (kFieldX2 + ((f - kFieldX2) >> 1)) * kFrontThreads + t + 64 * ((f - kFieldX2) & 1);
    return (f - 2) * kFrontThreads + t;
}

// Whether thread t holds field f (the rest are not stored).
VSHIP_INLINE bool holds_field(int f, int t) { return f < kFieldX2 || t < 64; }

VSHIP_INLINE float field(const TileXyb& r, int f)
{
    if (f < kFieldX1) return r.y0[(f - kFieldY0) / kBW][(f - kFieldY0) % kBW];
    if (f < kFieldY1) return r.x1[f - kFieldX1];
    if (f < kFieldX2) return r.y1[f - kFieldY1];
    if (f < kFieldY2) return r.x2[f - kFieldX2];
    if (f < kFieldL3) return r.y2[f - kFieldY2];
    return r.l3[f - kFieldL3];
}

// Chroma of the thread's 2x2 block when no axis is subsampled more than 2x:
// both output phases of a 2x axis have tile-independent Catmull-Rom weights
// (FrontColor::hw, vw) over the 5 chroma samples around the block's own
// chroma sample, so the upsampling is a 5-tap dot product per output and no
// lane computes segment positions. With both axes subsampled the horizontal
// pass runs over every staged row first (like the reference,
// chromaUpsample.hpp:165-202), into s_mid with row pitch kMidPitch.
constexpr int kMidPitch = kFrontTileW + 2;

// LDS floats of the chroma staging (both planes with halo, plus the
// horizontally upsampled rows when both axes are subsampled) of one layout,
// the largest over the supported layouts, and of one image's level 1.
constexpr int chroma_lds_floats(int subw, int subh)
{
    return (subw == 0 && subh == 0)
               ? 0
               : 2 * ((kFrontTileW >> subw) + 4) * ((kFrontTileH >> subh) + 4) +
                     ((subw && subh) ? 2 * ((kFrontTileH >> subh) + 4) * (subw <= 1 ? kMidPitch : kFrontTileW) : 0);
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kChromaLdsMax =
    cmax(cmax(cmax(chroma_lds_floats(1, 1), chroma_lds_floats(1, 0)), cmax(chroma_lds_floats(2, 0), chroma_lds_floats(2, 1))),
         chroma_lds_floats(0, 1));
constexpr int kL1Floats = 3 * 16 * 32;  // level 1 of one image: [3][16 rows][32]

VSHIP_INLINE float dot5(const float w[5], const float v[5])
{
    return fmaf(w[0], v[0], fmaf(w[1], v[1], fmaf(w[2], v[2], fmaf(w[3], v[3], w[4] * v[4]))));
}

// 4:2:0 with chroma horizontally co-sited and vertically centred (the layout
// instantiations): horizontal phase 0 copies the sample (weights 0 0 1 0 0),
// horizontal phase 1 and vertical phase 1 have w[0] = 0, vertical phase 0 has
// w[4] = 0. Dropping those terms from dot5's chain gives the same floats
// (fmaf(0, v, x) == x and fmaf(w, v, 0 * u) == w * v for finite samples).
// The horizontal items (2 planes x 20 staged rows, 32 column pairs each) go
// to the threads in a loop of fixed trip count.
VSHIP_INLINE void chroma_420_left(const FrontColor& fc, const float* s_chroma, float* s_mid, float P[3][kBH][kBW])
{
    constexpr int ccols = (kFrontTileW >> 1) + 4, crows = (kFrontTileH >> 1) + 4;
    constexpr int kRowStep = kFrontThreads / 32;
    constexpr int kIters = (2 * crows + kRowStep - 1) / kRowStep;
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int m = t & 31;
    // staged row e = p * crows + r of plane p is row e of both the staged
    // planes and s_mid (planes back to back), so item it of the thread reads
    // and writes at fixed offsets from its first one (unrolled: the offsets
    // are the LDS instructions' immediates, and the last item's condition is
    // uniform per wave)
    const float* src0 = s_chroma + (t >> 5) * ccols + m;
    float* dst0 = s_mid + (t >> 5) * kMidPitch + 2 * m;
#pragma unroll
    for (int it = 0; it < kIters; it++) {
        if ((t >> 5) + kRowStep * it >= 2 * crows) break;
        const float* src = src0 + kRowStep * it * ccols;
        const float v1 = src[1], v2 = src[2], v3 = src[3], v4 = src[4];
        const float h1 = fmaf(fc.hw[1][1], v1, fmaf(fc.hw[1][2], v2, fmaf(fc.hw[1][3], v3, fc.hw[1][4] * v4)));
        *reinterpret_cast<float2*>(dst0 + kRowStep * it * kMidPitch) = make_float2(v2, h1);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 2; p++) {
        float2 v2[5];
#pragma unroll
        for (int q = 0; q < 5; q++)
            v2[q] = *reinterpret_cast<const float2*>(s_mid + (p * crows + ty + q) * kMidPitch + 2 * tx);
        const float* w0 = fc.vw[0];
        const float* w1 = fc.vw[1];
        P[1 + p][0][0] = fmaf(w0[0], v2[0].x, fmaf(w0[1], v2[1].x, fmaf(w0[2], v2[2].x, w0[3] * v2[3].x)));
        P[1 + p][0][1] = fmaf(w0[0], v2[0].y, fmaf(w0[1], v2[1].y, fmaf(w0[2], v2[2].y, w0[3] * v2[3].y)));
        P[1 + p][1][0] = fmaf(w1[1], v2[1].x, fmaf(w1[2], v2[2].x, fmaf(w1[3], v2[3].x, w1[4] * v2[4].x)));
        P[1 + p][1][1] = fmaf(w1[1], v2[1].y, fmaf(w1[2], v2[2].y, fmaf(w1[3], v2[3].y, w1[4] * v2[4].y)));
    }
}

VSHIP_INLINE void chroma_2x(const FrontColor& fc, const float* s_chroma, float* s_mid, int ccols, int crows,
                            bool both, float P[3][kBH][kBW])
{
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int plane_elems = ccols * crows;
    if (both) {
        // lane pair (2m, 2m + 1) of tile columns from staged columns m .. m + 4
        const int m = t & 31;
        for (int e = t >> 5; e < 2 * crows; e += kFrontThreads / 32) {
            const int p = e >= crows ? 1 : 0, r = e - p * crows;
            const float* src = s_chroma + p * plane_elems + r * ccols + m;
            float v[5];
#pragma unroll
            for (int q = 0; q < 5; q++) v[q] = src[q];
            *reinterpret_cast<float2*>(s_mid + (p * crows + r) * kMidPitch + 2 * m) =
                make_float2(dot5(fc.hw[0], v), dot5(fc.hw[1], v));
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 2; p++) {
            float2 v2[5];
#pragma unroll
            for (int q = 0; q < 5; q++)
                v2[q] = *reinterpret_cast<const float2*>(s_mid + (p * crows + ty + q) * kMidPitch + 2 * tx);
            const float v0[5] = {v2[0].x, v2[1].x, v2[2].x, v2[3].x, v2[4].x};
            const float v1[5] = {v2[0].y, v2[1].y, v2[2].y, v2[3].y, v2[4].y};
#pragma unroll
            for (int j = 0; j < kBH; j++) {
                P[1 + p][j][0] = dot5(fc.vw[j], v0);
                P[1 + p][j][1] = dot5(fc.vw[j], v1);
            }
        }
    } else if (fc.subw) {
        // 2x horizontal only: rows y0 + j are staged rows 2ty + j + 2
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int j = 0; j < kBH; j++) {
                const float* src = s_chroma + p * plane_elems + (2 * ty + j + 2) * ccols + tx;
                float v[5];
#pragma unroll
                for (int q = 0; q < 5; q++) v[q] = src[q];
                P[1 + p][j][0] = dot5(fc.hw[0], v);
                P[1 + p][j][1] = dot5(fc.hw[1], v);
            }
    } else {
        // 2x vertical only: columns x0 + i are staged columns 2tx + i + 2
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int i = 0; i < kBW; i++) {
                const float* src = s_chroma + p * plane_elems + ty * ccols + 2 * tx + i + 2;
                float v[5];
#pragma unroll
                for (int q = 0; q < 5; q++) v[q] = src[q * ccols];
                P[1 + p][0][i] = dot5(fc.vw[0], v);
                P[1 + p][1][i] = dot5(fc.vw[1], v);
            }
    }
}

// Linear BT.709 RGB of the thread's 2x2 block of image k: sample decode,
// range, chroma upsampling (LDS), matrix, EOTF, linear-light matrix and
// primaries. Ends with the LDS free for reuse only after the caller syncs.
// preloaded: P's plane values and the staged chroma samples are already in
// place (unpack_raw of samples loaded ahead by raw_tile); otherwise they are
// loaded here.
template <int SPEC = kFrontGeneric>
VSHIP_INLINE void to_linear(const FrontJob& job, int k, const FrontColor& fc, int tile_x, int tile_y, float* lds,
                            float P[3][kBH][kBW], bool preloaded = false)
{
    constexpr bool O32 = SPEC != kFrontGeneric;  // 32-bit sample offsets (front_offsets_fit)
    const uint8_t* const* plane = job.plane[k];
    const int64_t* pstride = job.stride[k];
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int W = job.width, H = job.height;
    const int x0 = tile_x * kFrontTileW + kBW * tx;
    const int y0 = tile_y * kFrontTileH + kBH * ty;
    const RangeXform r0 = {fc.range0[0], fc.range0[1]};
    const RangeXform r12 = {fc.range12[0], fc.range12[1]};

    // P: plane values, then linear R, G, B of the block
    const bool sub = fc.subw || fc.subh;
    if (!preloaded) {
        switch (fc.sample) {
        case kSampleU8:
            load_block<kSampleU8, O32>(plane[0], pstride[0], x0, y0, W, H, fc.mask, r0, P[0]);
            if (!sub) {
                load_block<kSampleU8, O32>(plane[1], pstride[1], x0, y0, W, H, fc.mask, r12, P[1]);
                load_block<kSampleU8, O32>(plane[2], pstride[2], x0, y0, W, H, fc.mask, r12, P[2]);
            }
            break;
        case kSampleU16:
            load_block<kSampleU16, O32>(plane[0], pstride[0], x0, y0, W, H, fc.mask, r0, P[0]);
            if (!sub) {
                load_block<kSampleU16, O32>(plane[1], pstride[1], x0, y0, W, H, fc.mask, r12, P[1]);
                load_block<kSampleU16, O32>(plane[2], pstride[2], x0, y0, W, H, fc.mask, r12, P[2]);
            }
            break;
        case kSampleHalf:
            load_block<kSampleHalf, O32>(plane[0], pstride[0], x0, y0, W, H, fc.mask, r0, P[0]);
            if (!sub) {
                load_block<kSampleHalf, O32>(plane[1], pstride[1], x0, y0, W, H, fc.mask, r12, P[1]);
                load_block<kSampleHalf, O32>(plane[2], pstride[2], x0, y0, W, H, fc.mask, r12, P[2]);
            }
            break;
        default:
            load_block<kSampleFloat, O32>(plane[0], pstride[0], x0, y0, W, H, fc.mask, r0, P[0]);
            if (!sub) {
                load_block<kSampleFloat, O32>(plane[1], pstride[1], x0, y0, W, H, fc.mask, r12, P[1]);
                load_block<kSampleFloat, O32>(plane[2], pstride[2], x0, y0, W, H, fc.mask, r12, P[2]);
            }
            break;
        }

        if (sub) {
            // both chroma planes staged in LDS with a 2-sample halo; clamped
            // coordinates reproduce the reference's edge rules
            // (chromaUpsample.hpp:29-35,41-47)
            const int ccols = (kFrontTileW >> fc.subw) + 4;
            const int crows = (kFrontTileH >> fc.subh) + 4;
            const int cx0 = ((tile_x * kFrontTileW) >> fc.subw) - 2;
            const int cy0 = ((tile_y * kFrontTileH) >> fc.subh) - 2;
            const int cw = ((W - 1) >> fc.subw) + 1;
            const int ch = ((H - 1) >> fc.subh) + 1;
            switch (fc.sample) {
            case kSampleU8: stage_chroma<kSampleU8, O32>(plane, pstride, fc.mask, r12, lds, ccols, crows, cx0, cy0, cw, ch); break;
            case kSampleU16: stage_chroma<kSampleU16, O32>(plane, pstride, fc.mask, r12, lds, ccols, crows, cx0, cy0, cw, ch); break;
            case kSampleHalf: stage_chroma<kSampleHalf, O32>(plane, pstride, fc.mask, r12, lds, ccols, crows, cx0, cy0, cw, ch); break;
            default: stage_chroma<kSampleFloat, O32>(plane, pstride, fc.mask, r12, lds, ccols, crows, cx0, cy0, cw, ch); break;
            }
        }
    }

    if (sub) {
        // chroma upsampling from the staged planes
        const int ccols = (kFrontTileW >> fc.subw) + 4;
        const int crows = (kFrontTileH >> fc.subh) + 4;
        const int cx0 = ((tile_x * kFrontTileW) >> fc.subw) - 2;
        const int cy0 = ((tile_y * kFrontTileH) >> fc.subh) - 2;
        const int plane_elems = ccols * crows;
        float* s_chroma = lds;
        float* s_mid = lds + 2 * plane_elems;
        __syncthreads();
        const bool both = fc.subw && fc.subh;
        if constexpr (SPEC != kFrontGeneric) {
            chroma_420_left(fc, s_chroma, s_mid, P);
        } else if (fc.subw <= 1 && fc.subh <= 1) {
            chroma_2x(fc, s_chroma, s_mid, ccols, crows, both, P);
        } else if (both) {
            // horizontal pass over every staged chroma row first, like the
            // reference (chromaUpsample.hpp:165-202)
            const int X = t & (kFrontTileW - 1);
            int kk; float tt; bool copy;
            seg_of(tile_x * kFrontTileW + X, fc.subw, fc.h_cosited != 0, kk, tt, copy);
            const int c = kk - cx0;
            const CrWeights cw4 = cr_weights(tt);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                for (int r = t / kFrontTileW; r < crows; r += kFrontThreads / kFrontTileW) {
                    const float* row = s_chroma + p * plane_elems + r * ccols;
                    s_mid[(p * crows + r) * kFrontTileW + X] = cr_pick(copy, cw4, row + c, 1);
                }
            }
            __syncthreads();
        }
        if (fc.subw > 1) {
            // 4:1:x: segment parameters of the thread's columns and rows
            int kx[kBW], ky[kBH];
            CrWeights wx[kBW], wy[kBH];
            bool cpx[kBW], cpy[kBH];
#pragma unroll
            for (int i = 0; i < kBW; i++) {
                float tt = 0.0f;
                if (fc.subw) seg_of(x0 + i, fc.subw, fc.h_cosited != 0, kx[i], tt, cpx[i]);
                else { kx[i] = x0 + i; cpx[i] = true; }
                wx[i] = cr_weights(tt);
            }
#pragma unroll
            for (int j = 0; j < kBH; j++) {
                float tt = 0.0f;
                if (fc.subh) seg_of(y0 + j, fc.subh, fc.v_cosited != 0, ky[j], tt, cpy[j]);
                else { ky[j] = y0 + j; cpy[j] = true; }
                wy[j] = cr_weights(tt);
            }
#pragma unroll
            for (int p = 0; p < 2; p++) {
#pragma unroll
                for (int j = 0; j < kBH; j++) {
#pragma unroll
                    for (int i = 0; i < kBW; i++) {
                        float val;
                        if (fc.subh == 0) {
                            const float* row = s_chroma + p * plane_elems + (y0 + j - cy0) * ccols;
                            const int c = kx[i] - cx0;
                            val = cr_pick(cpx[i], wx[i], row + c, 1);
                        } else {
                            const int r = ky[j] - cy0;
                            const float* colbase;
                            int pitch;
                            if (both) {
                                colbase = s_mid + p * crows * kFrontTileW + (x0 + i - tile_x * kFrontTileW);
                                pitch = kFrontTileW;
                            } else {
                                colbase = s_chroma + p * plane_elems + (x0 + i - cx0);
                                pitch = ccols;
                            }
                            val = cr_pick(cpy[j], wy[j], colbase + r * pitch, pitch);
                        }
                        P[1 + p][j][i] = val;
                    }
                }
        }
        }
    }

    // matrix, EOTF, linear-light matrix with primaries (P becomes linear RGB)
    if constexpr (SPEC != kFrontGeneric) {
        // YCbCr matrix with unit luma column and R(Cb) = B(Cr) = 0 (front_spec
        // checks): mat3_apply's chains without the terms fmaf(0, v, t) == t
        // and 1 * Y == Y, so the same floats
        const float* m = fc.pre;
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++) {
                const float y = P[0][j][i], u = P[1][j][i], v = P[2][j][i];
                P[0][j][i] = fmaf(m[2], v, y);
                P[1][j][i] = fmaf(m[5], v, fmaf(m[4], u, y));
                P[2][j][i] = fmaf(m[7], u, y);
            }
    } else if (fc.use_pre) {
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++) mat3_apply(fc.pre, P[0][j][i], P[1][j][i], P[2][j][i]);
    }
    apply_transfer(fc.transfer, P);
    if (fc.use_post) {
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++) mat3_apply(fc.post, P[0][j][i], P[1][j][i], P[2][j][i]);
    }
}

// Cross-lane partners of the level-3 and level-4 sums: lane ^ 1 and lane ^ 8
// (a row rotation by 8 within each 16-lane row) through DPP, lane ^ 16
// through the LDS crossbar's swizzle (no LDS memory).
VSHIP_INLINE float lane_xor1(float v) { return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf, 0xf, false)); }
VSHIP_INLINE float lane_ror8(float v) { return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xf, 0xf, false)); }
VSHIP_INLINE float lane_xor16(float v) { return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F)); }

// P holds the image's unpacked plane values when preloaded (see to_linear).
// after_level0(res) runs once the level-0 values are in res (the FP32 layouts
// issue their stores there, so that they drain while levels 1 to 4 are
// computed).
struct NoLevel0Hook {
    VSHIP_INLINE void operator()(const TileXyb&) const {}
};

template <int SPEC = kFrontGeneric, typename AfterLevel0 = NoLevel0Hook>
VSHIP_INLINE void compute_tile(const FrontJob& job, int k, const FrontColor& fc, const PairGeom& geo, int tile_x,
                               int tile_y, float* lds, float* lds_l1, TileXyb& res, float P[3][kBH][kBW],
                               bool preloaded = false, AfterLevel0 after_level0 = {})
{
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int W = job.width, H = job.height;
    const int x0 = tile_x * kFrontTileW + kBW * tx;
    const int y0 = tile_y * kFrontTileH + kBH * ty;
    to_linear<SPEC>(job, k, fc, tile_x, tile_y, lds, P, preloaded);

    // ---- level 0: Y only
#pragma unroll
    for (int j = 0; j < kBH; j++)
#pragma unroll
        for (int i = 0; i < kBW; i++) res.y0[j][i] = to_xyb(P[0][j][i], P[1][j][i], P[2][j][i], false).y;
    after_level0(res);

    // ---- level 1 (kBH/2 rows of one column per thread), edge clamp min(2x+1, w-1);
    // a tile inside the image needs no clamp (the same sums)
    float L1[kBH / 2][3];
    auto level1 = [&](auto clamped) {
        const bool xlast = decltype(clamped)::value && !(x0 + 1 < W);
#pragma unroll
        for (int j = 0; j < kBH / 2; j++) {
            const bool ylast = decltype(clamped)::value && !(y0 + 2 * j + 1 < H);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float p00 = P[c][2 * j][0];
                const float p10 = ylast ? p00 : P[c][2 * j + 1][0];
                const float p01 = xlast ? p00 : P[c][2 * j][1];
                const float p11 = ylast ? p01 : (xlast ? p10 : P[c][2 * j + 1][1]);
                float s = p00;
                s += p10;
                s += p01;
                s += p11;
                L1[j][c] = s * 0.25f;
            }
        }
    };
    if ((tile_x + 1) * kFrontTileW <= W && (tile_y + 1) * kFrontTileH <= H) level1(std::false_type{});
    else level1(std::true_type{});
#pragma unroll
    for (int j = 0; j < kBH / 2; j++) {
        const XYB q = to_xyb(L1[j][0], L1[j][1], L1[j][2], false);
        res.x1[j] = q.x;
        res.y1[j] = q.y;
    }

    // ---- levels 2, 3, 4 (32x16 -> 16x8 -> 8x4 -> 4x2 per tile): level 1
    // through LDS (its own region per image, so no barrier protects it from
    // the next image), then wave 0 alone: lane 8r + c holds the level-2
    // pixels (2c, r) and (2c + 1, r); level 3 adds the lane's pair and the
    // next row's pair (lane + 8) in lanes of even r, level 4 the level-3
    // pixels of lanes ^ 1, ^ 16 and ^ 17 in lanes 0, 2, 4, 6, 32, ..., 38.
    // Every sum is in the reference's order with its edge clamp
    // (downsample.hpp:13-17): p(x, y) + p(x, y + 1) + p(x + 1, y) +
    // p(x + 1, y + 1), a parent beyond the level's edge replaced by the one
    // inside.
    float* s_l1 = lds_l1;  // [3][16][32]
#pragma unroll
    for (int j = 0; j < kBH / 2; j++)
#pragma unroll
        for (int c = 0; c < 3; c++) s_l1[(c * 16 + ty * (kBH / 2) + j) * 32 + tx] = L1[j][c];
    __syncthreads();
    if (t < 64) {
        const int c = t & 7, r = t >> 3;
        // level-1 parents: columns 4c .. 4c + 3 of rows 2r, 2r + 1
        const int gx1 = tile_x * 32 + 4 * c, gy1 = tile_y * 16 + 2 * r;
        const bool x1in = gx1 + 1 < geo.w[1], x3in = gx1 + 3 < geo.w[1], y1in = gy1 + 1 < geo.h[1];
        float L2[2][3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float4 a = *reinterpret_cast<const float4*>(s_l1 + (ch * 16 + 2 * r) * 32 + 4 * c);
            const float4 b0 = *reinterpret_cast<const float4*>(s_l1 + (ch * 16 + 2 * r + 1) * 32 + 4 * c);
            const float4 b = y1in ? b0 : a;
            float s0 = a.x;
            s0 += b.x;
            s0 += x1in ? a.y : a.x;
            s0 += x1in ? b.y : b.x;
            L2[0][ch] = s0 * 0.25f;
            float s1 = a.z;
            s1 += b.z;
            s1 += x3in ? a.w : a.z;
            s1 += x3in ? b.w : b.z;
            L2[1][ch] = s1 * 0.25f;
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const XYB q = to_xyb(L2[i][0], L2[i][1], L2[i][2], false);
            res.x2[i] = q.x;
            res.y2[i] = q.y;
        }
        // level 3 pixel (c, r / 2) in lanes of even r
        const bool x2in = tile_x * 16 + 2 * c + 1 < geo.w[2], y2in = tile_y * 8 + r + 1 < geo.h[2];
        float L3[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float a0 = L2[0][ch], a1 = x2in ? L2[1][ch] : a0;
            const float n0 = lane_ror8(L2[0][ch]), n1 = lane_ror8(L2[1][ch]);
            const float b0 = y2in ? n0 : a0, b1 = y2in ? (x2in ? n1 : n0) : a1;
            float v = a0;
            v += b0;
            v += a1;
            v += b1;
            L3[ch] = v * 0.25f;
        }
        // level 4 pixel (c / 2, r / 4) in lanes of even c and r % 4 == 0
        const bool x3in4 = tile_x * 8 + c + 1 < geo.w[3], y3in = tile_y * 4 + (r >> 1) + 1 < geo.h[3];
        float L4[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float p00 = L3[ch];
            const float py = lane_xor16(p00);
            const float px = lane_xor1(p00), pxy = lane_xor1(py);
            const float p01 = y3in ? py : p00;
            const float p10 = x3in4 ? px : p00;
            const float p11 = y3in ? (x3in4 ? pxy : py) : p10;
            float v = p00;
            v += p01;
            v += p10;
            v += p11;
            L4[ch] = v * 0.25f;
        }
        // one XYB pass: lanes of odd r take the level-3 pixel of lane - 8,
        // the level-4 lanes keep theirs
        float G[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float l3 = lane_ror8(L3[ch]);
            G[ch] = (r & 1) ? l3 : L4[ch];
        }
        const XYB q = to_xyb(G[0], G[1], G[2], true);
        res.l3[0] = q.x; res.l3[1] = q.y; res.l3[2] = q.b;
    }
}

// Writes the tile's stored planes. emit(plane, element offset, k) writes
// plane element `off` (32-bit: an image buffer holds fewer than 2^31
// elements) from field value k of the thread's TileXyb; emit2(plane, off, k)
// writes fields k and k + 1 to elements off and off + 1 (level-0 and level-2
// pixel pairs: plane offsets, row pitches and their first columns are even,
// so the pair is aligned).
template <typename Emit, typename Emit2>
VSHIP_INLINE void store_level0(const FrontJob& job, const PairGeom& geo, int tile_x, int tile_y, Emit emit,
                               Emit2 emit2)
{
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    const int W = job.width, H = job.height;
    const int x0 = tile_x * kFrontTileW + kBW * tx;
    const int y0 = tile_y * kFrontTileH + kBH * ty;
    static_assert(kBW == 2, "level-0 row pairs");
#pragma unroll
    for (int j = 0; j < kBH; j++) {
        if (y0 + j < H) {
            const uint32_t o = (uint32_t)geo.off[0][1] + __umul24((uint32_t)(y0 + j), (uint32_t)geo.pitch[0]) + (uint32_t)x0;
            if (x0 + 1 < W) {
                emit2(1, o, kFieldY0 + j * kBW);
            } else if (x0 < W) {
                emit(1, o, kFieldY0 + j * kBW);
            }
        }
    }
}

// kLevel0 false: levels 1 to 4 only (level 0 went out through
// store_level0).
template <bool kLevel0 = true, typename Emit, typename Emit2>
VSHIP_INLINE void store_tile(const FrontJob& job, const PairGeom& geo, int tile_x, int tile_y, Emit emit,
                             Emit2 emit2)
{
    const int t = threadIdx.x;
    const int tx = t % kTX, ty = t / kTX;
    if constexpr (kLevel0) {
        // store_level0's rows, written out: calling it from here moves the
        // FP16 instantiations' instruction order enough to spill registers
        const int W = job.width, H = job.height;
        const int x0 = tile_x * kFrontTileW + kBW * tx;
        const int y0 = tile_y * kFrontTileH + kBH * ty;
#pragma unroll
        for (int j = 0; j < kBH; j++) {
            if (y0 + j < H) {
                const uint32_t o =
                    (uint32_t)geo.off[0][1] + __umul24((uint32_t)(y0 + j), (uint32_t)geo.pitch[0]) + (uint32_t)x0;
                if (x0 + 1 < W) {
                    emit2(1, o, kFieldY0 + j * kBW);
                } else if (x0 < W) {
                    emit(1, o, kFieldY0 + j * kBW);
                }
            }
        }
    }
    const int x1 = tile_x * 32 + tx;
#pragma unroll
    for (int j = 0; j < kBH / 2; j++) {
        const int y1 = tile_y * 16 + ty * (kBH / 2) + j;
        if (y1 < geo.h[1] && x1 < geo.w[1]) {
            const uint32_t o = __umul24((uint32_t)y1, (uint32_t)geo.pitch[1]) + (uint32_t)x1;
            emit(0, (uint32_t)geo.off[1][0] + o, kFieldX1 + j);
            emit(1, (uint32_t)geo.off[1][1] + o, kFieldY1 + j);
        }
    }
    // levels 2 to 4: wave 0 (lane 8r + c). Level 2: pixels (2c, r) and
    // (2c + 1, r), a pair of consecutive elements (even column). Levels 3
    // and 4: lanes of odd r hold level-3 pixel (c, r / 2), lanes of even c
    // and r % 4 == 0 level-4 pixel (c / 2, r / 4). Both levels' geometry is
    // read as uniform values and selected per lane: indexing geo by a
    // lane-dependent level would be per-lane memory loads issued after this
    // tile's plane stores, and waiting for them (loads and stores share one
    // in-order counter) would wait for every store to complete.
    if (t < 64) {
        const int c = t & 7, r = t >> 3;
        const int x2 = tile_x * 16 + 2 * c, y2 = tile_y * 8 + r;
        if (y2 < geo.h[2] && x2 < geo.w[2]) {
            const uint32_t o = __umul24((uint32_t)y2, (uint32_t)geo.pitch[2]) + (uint32_t)x2;
            if (x2 + 1 < geo.w[2]) {
                emit2(0, (uint32_t)geo.off[2][0] + o, kFieldX2);
                emit2(1, (uint32_t)geo.off[2][1] + o, kFieldY2);
            } else {
                emit(0, (uint32_t)geo.off[2][0] + o, kFieldX2);
                emit(1, (uint32_t)geo.off[2][1] + o, kFieldY2);
            }
        }
        // readfirstlane: the values are uniform; as opaque scalars the
        // compiler cannot turn their per-lane selection back into addresses
        auto u = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
        const int w3 = u(geo.w[3]), h3 = u(geo.h[3]), p3 = u(geo.pitch[3]);
        const int w4 = u(geo.w[4]), h4 = u(geo.h[4]), p4 = u(geo.pitch[4]);
        uint32_t o3[3], o4[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            o3[ch] = (uint32_t)u((int)geo.off[3][ch]);
            o4[ch] = (uint32_t)u((int)geo.off[4][ch]);
        }
        const bool l3 = (r & 1) != 0;
        const bool l4 = (r & 3) == 0 && (c & 1) == 0;
        const int x = l3 ? tile_x * 8 + c : tile_x * 4 + (c >> 1);
        const int y = l3 ? tile_y * 4 + (r >> 1) : tile_y * 2 + (r >> 2);
        if ((l3 || l4) && x < (l3 ? w3 : w4) && y < (l3 ? h3 : h4)) {
            const uint32_t o = __umul24((uint32_t)y, (uint32_t)(l3 ? p3 : p4)) + (uint32_t)x;
#pragma unroll
            for (int ch = 0; ch < 3; ch++) emit(ch, (l3 ? o3[ch] : o4[ch]) + o, kFieldL3 + ch);
        }
    }
}

// Global-memory stores and loads of one or two consecutive plane elements at
// element offset `off` of an image's plane buffer. The byte offset is formed
// in 32 bits (the host keeps an image's stored elements below 2^30), so the
// access takes the buffer base as its scalar address and the offset as a
// 32-bit lane register, with no per-lane 64-bit address arithmetic.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

template <typename T>
VSHIP_INLINE __attribute__((address_space(1))) T* elem_ptr(const void* base, uint32_t off)
{
    return (__attribute__((address_space(1))) T*)((__attribute__((address_space(1))) uint8_t*)base +
                                                  off * (uint32_t)sizeof(T));
}

template <typename StoreT>
VSHIP_INLINE void gstore1(void* base, uint32_t off, float v);
template <>
VSHIP_INLINE void gstore1<float>(void* base, uint32_t off, float v)
{
    *elem_ptr<float>(base, off) = v;
}
template <>
VSHIP_INLINE void gstore1<__half>(void* base, uint32_t off, float v)
{
    *elem_ptr<_Float16>(base, off) = (_Float16)v;
}

template <typename StoreT>
VSHIP_INLINE void gstore2(void* base, uint32_t off, float v0, float v1);
template <>
VSHIP_INLINE void gstore2<float>(void* base, uint32_t off, float v0, float v1)
{
    *(__attribute__((address_space(1))) f32x2_t*)elem_ptr<float>(base, off) = f32x2_t{v0, v1};
}
template <>
VSHIP_INLINE void gstore2<__half>(void* base, uint32_t off, float v0, float v1)
{
    *(__attribute__((address_space(1))) f16x2_t*)elem_ptr<_Float16>(base, off) = f16x2_t{(_Float16)v0, (_Float16)v1};
}

// FP32 elements of a plane layout (the FP32 values of an image computed by an
// earlier job).
VSHIP_INLINE float gload1(const float* base, uint32_t off)
{
    return *elem_ptr<const float>(base, off);
}

VSHIP_INLINE void gload2(const float* base, uint32_t off, float& v0, float& v1)
{
    const f32x2_t v = *(const __attribute__((address_space(1))) f32x2_t*)elem_ptr<const float>(base, off);
    v0 = v[0];
    v1 = v[1];
}

// Colour layouts with their own instantiation of the front-end: with the
// sample type, subsampling, chroma siting, transfer and matrix steps known at
// compile time the per-format switches disappear, the chroma pass drops its
// zero taps and the kernel needs fewer registers. The descriptor's values
// (ranges, matrices, weights) stay run-time.
template <int SPEC>
VSHIP_INLINE FrontColor front_color(const FrontColor& in)
{
    FrontColor fc = in;
    fc.sample = SPEC == kFrontU16Yuv420Bt709 ? kSampleU16 : kSampleU8;
    fc.subw = 1;
    fc.subh = 1;
    fc.transfer = 1;
    fc.use_pre = 1;
    fc.use_post = 0;
    return fc;
}

// One tile of both images of a pair. FP32 layout: each image is written as
// soon as it is computed. FP16 layout: the reference plane holds
// fp16(a - offset) and the distorted plane fp16(b - a) of the FP32 values, so
// the reference image's values wait in an LDS stash (field-major, one float
// per thread and field) while the distorted image is computed.
template <typename StoreT, int SPEC = kFrontGeneric>
VSHIP_INLINE void frontend_tile(const FrontJob& job, const FrontColor* __restrict__ colors, const PairGeom& geo,
                                int tile_x, int tile_y, float* lds)
{
    constexpr bool kHalf = sizeof(StoreT) == 2;
    float* stash = lds;
    float* work = lds + (kHalf ? kFrontStashFloats : 0);
    float* l1_area = work + kChromaLdsMax;  // [2 images][kL1Floats]
    const int t = threadIdx.x;
    // FP32 layout: an image's level-0 plane is stored as soon as it is
    // computed (the largest part of its stores, which then drains while
    // levels 1 to 4 are computed), the other levels at the end
    auto level0_to = [&](void* out) {
        return [&, out](const TileXyb& r) {
            store_level0(
                job, geo, tile_x, tile_y, [&](int, uint32_t off, int f) { gstore1<StoreT>(out, off, field(r, f)); },
                [&](int, uint32_t off, int f) { gstore2<StoreT>(out, off, field(r, f), field(r, f + 1)); });
        };
    };
    auto levels1to4_to = [&](void* out, const TileXyb& r) {
        store_tile<false>(
            job, geo, tile_x, tile_y, [&](int, uint32_t off, int f) { gstore1<StoreT>(out, off, field(r, f)); },
            [&](int, uint32_t off, int f) { gstore2<StoreT>(out, off, field(r, f), field(r, f + 1)); });
    };
    if constexpr (!kHalf) {
        if (job.shared) {  // image 0 is stored already: image 1 only
            TileXyb r;
            float P[3][kBH][kBW];
            void* o1 = job.out[1];
            if constexpr (SPEC == kFrontGeneric) {
                compute_tile(job, 1, colors[job.color[1]], geo, tile_x, tile_y, work, l1_area + kL1Floats, r, P,
                             false, level0_to(o1));
            } else {
                const FrontColor c1 = front_color<SPEC>(colors[job.color[1]]);
                compute_tile<SPEC>(job, 1, c1, geo, tile_x, tile_y, work, l1_area + kL1Floats, r, P, false,
                                   level0_to(o1));
            }
            levels1to4_to(o1, r);
            return;
        }
    }
    if constexpr (!kHalf && SPEC == kFrontGeneric) {
#pragma unroll 1
        for (int k = 0; k < 2; k++) {
            TileXyb r;
            float P[3][kBH][kBW];
            void* out = job.out[k];
            compute_tile(job, k, colors[job.color[k]], geo, tile_x, tile_y, work, l1_area + k * kL1Floats, r, P,
                         false, level0_to(out));
            levels1to4_to(out, r);
        }
    } else if constexpr (!kHalf) {
        // layout instantiation: image 1's raw samples load while image 0 is
        // computed (the chroma staging area is free again once compute_tile
        // has passed its level-2 barrier)
        TileXyb r;
        float P[3][kBH][kBW];
        const FrontColor c0 = front_color<SPEC>(colors[job.color[0]]);
        const FrontColor c1 = front_color<SPEC>(colors[job.color[1]]);
        RawTile raw = raw_tile<true>(job, 0, c0, tile_x, tile_y);
        unpack_raw(c0, raw, work, P);
        raw = raw_tile<true>(job, 1, c1, tile_x, tile_y);
        void* o0 = job.out[0];
        void* o1 = job.out[1];
        // (image 1's samples were loaded before image 0's level-0 stores, so
        // waiting for them does not wait for the stores: loads and stores
        // share one in-order wait counter)
        compute_tile<SPEC>(job, 0, c0, geo, tile_x, tile_y, work, l1_area, r, P, true, level0_to(o0));
        unpack_raw(c1, raw, work, P);
        levels1to4_to(o0, r);
        compute_tile<SPEC>(job, 1, c1, geo, tile_x, tile_y, work, l1_area + kL1Floats, r, P, true, level0_to(o1));
        levels1to4_to(o1, r);
    } else {
        TileXyb a, b;
        float P[3][kBH][kBW];
        compute_tile(job, 0, colors[job.color[0]], geo, tile_x, tile_y, work, l1_area, a, P);
#pragma unroll
        for (int f = 0; f < kFields; f++)
            if (holds_field(f, t)) stash[stash_slot(f, t)] = field(a, f);
        compute_tile(job, 1, colors[job.color[1]], geo, tile_x, tile_y, work, l1_area + kL1Floats, b, P);
        void* o0 = job.out[0];
        void* o1 = job.out[1];
        store_tile(
            job, geo, tile_x, tile_y,
            [&](int plane, uint32_t off, int f) {
                const float va = stash[stash_slot(f, t)];
                gstore1<StoreT>(o0, off, va - kHalfOffset[plane]);
                gstore1<StoreT>(o1, off, field(b, f) - va);
            },
            [&](int plane, uint32_t off, int f) {
                const float va0 = stash[stash_slot(f, t)], va1 = stash[stash_slot(f + 1, t)];
                gstore2<StoreT>(o0, off, va0 - kHalfOffset[plane], va1 - kHalfOffset[plane]);
                gstore2<StoreT>(o1, off, field(b, f) - va0, field(b, f + 1) - va1);
            });
    }
}

// Persistent front-end over the job table (batches). Workgroups go
// round-robin over the XCDs: each XCD takes a contiguous range of tiles, so
// horizontally adjacent tiles, which share the 128-byte lines of their input
// rows, are fetched through the same L2.
//
// FP16 layout: the raw samples of each image are loaded while the previous
// image is computed (image 1 of a tile during image 0, image 0 of the next
// tile during image 1), so the memory latency of a tile's inputs overlaps
// work. The generic FP32 instantiation has no registers to spare for the 12
// words in flight and loads each image when it starts; the FP32 layout
// instantiations load image 1 of a tile during image 0 (frontend_tile).
VSHIP_INLINE int front_job_of(const FrontJob* __restrict__ jobs, int njobs, int item)
{
    // largest j with tile_start <= item
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_start <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The layout instantiations need fewer registers: 80 give 6 waves per SIMD,
// i.e. three 8-wave workgroups per CU (LDS allows three in both layouts).
//
// Each launch covers the tiles of one job group (CountSlot kCountFrontGroup):
// the jobs of layout SPEC that compute both images (PHASE 0) or whose image 0
// a phase-0 job stored (PHASE 1, launched after every phase-0 group on the
// same stream). Separate instantiations keep the image-1-only code out of the
// register allocation of the common one.
template <typename StoreT, int SPEC, int PHASE>
__global__ __launch_bounds__(kFrontThreads, SPEC == kFrontGeneric ? 4 : 6)
void frontend_kernel(const FrontJob* __restrict__ jobs, const int* __restrict__ counts,
                     const FrontColor* __restrict__ colors, const PairGeom* __restrict__ geoms)
{
    extern __shared__ float lds[];
    const int lo = counts[kCountFrontGroup + PHASE * 3 + SPEC];
    const int hi = counts[kCountFrontGroup + PHASE * 3 + SPEC + 1];
    const int njobs = counts[kCountFrontJobs];
    const int per = (hi - lo + kXcds - 1) / kXcds;
    const int xcd = blockIdx.x % kXcds;
    const int end = min(lo + (xcd + 1) * per, hi);
    const int step = gridDim.x / kXcds;
    int item = lo + xcd * per + blockIdx.x / kXcds;
    if constexpr (sizeof(StoreT) == 4) {
        for (; item < end; item += step) {
            const FrontJob& job = jobs[front_job_of(jobs, njobs, item)];
            const int local = item - job.tile_start;
            int tile_x, tile_y;
            front_tile_xy(job, local, tile_x, tile_y);
            frontend_tile<StoreT, SPEC>(job, colors, geoms[job.pair], tile_x, tile_y, lds);
            __syncthreads();  // LDS is reused by the next item
        }
    } else {
        float* stash = lds;
        float* work = lds + kFrontStashFloats;
        float* l1_area = work + kChromaLdsMax;
        const int t = threadIdx.x;
        constexpr bool kO32 = SPEC != kFrontGeneric;  // layout instantiations: 32-bit sample offsets
        // descriptor of image color index i, with the layout's constants folded in
        auto color = [&](int i) {
            if constexpr (SPEC == kFrontGeneric) return colors[i];
            else return front_color<SPEC>(colors[i]);
        };
        // raw samples of a tile's first image: image 1 when image 0 is shared
        auto first_raw = [&](const FrontJob& j, int local) {
            int tile_x, tile_y;
            front_tile_xy(j, local, tile_x, tile_y);
            return raw_tile<kO32>(j, PHASE, color(j.color[PHASE]), tile_x, tile_y);
        };
        RawTile raw = {};
        float P[3][kBH][kBW];
        if (item < end) {
            const FrontJob& job = jobs[front_job_of(jobs, njobs, item)];
            raw = first_raw(job, item - job.tile_start);
            // phase 0 unpacks each tile's first image before the loop reaches
            // it: this one here, the next ones before the current tile's
            // stores (loads and stores share one in-order wait counter, so
            // waiting for prefetched samples after the stores would wait for
            // the stores to complete)
            if constexpr (PHASE == 0) unpack_raw(color(job.color[0]), raw, work, P);
        }
        for (; item < end; item += step) {
            const FrontJob& job = jobs[front_job_of(jobs, njobs, item)];
            const PairGeom& geo = geoms[job.pair];
            const int local = item - job.tile_start;
            int tile_x, tile_y;
            front_tile_xy(job, local, tile_x, tile_y);
            const FrontColor c1 = color(job.color[1]);
            TileXyb a, b;
            const int next = item + step;
            if constexpr (PHASE == 1) {
                // image 0's FP32 values, stored by the job that computed it:
                // those of levels 0 and 1 in flight (registers) while image 1
                // is computed, those of levels 2 to 4 (wave 0) loaded after
                const float* e0 = job.exact0;
                float ev[kFields];
                store_tile(
                    job, geo, tile_x, tile_y,
                    [&](int, uint32_t off, int f) {
                        if (f < kFieldX2) ev[f] = gload1(e0, off);
                    },
                    [&](int, uint32_t off, int f) {
                        if (f < kFieldX2) gload2(e0, off, ev[f], ev[f + 1]);
                    });
                unpack_raw(c1, raw, work, P);
                if (next < end) {
                    const FrontJob& nj = jobs[front_job_of(jobs, njobs, next)];
                    raw = first_raw(nj, next - nj.tile_start);
                }
                compute_tile<SPEC>(job, 1, c1, geo, tile_x, tile_y, work, l1_area + kL1Floats, b, P, true);
                store_tile(
                    job, geo, tile_x, tile_y,
                    [&](int, uint32_t off, int f) {
                        if (f >= kFieldX2) ev[f] = gload1(e0, off);
                    },
                    [&](int, uint32_t off, int f) {
                        if (f >= kFieldX2) gload2(e0, off, ev[f], ev[f + 1]);
                    });
                void* o1 = job.out[1];
                store_tile(
                    job, geo, tile_x, tile_y,
                    [&](int, uint32_t off, int f) { gstore1<StoreT>(o1, off, field(b, f) - ev[f]); },
                    [&](int, uint32_t off, int f) {
                        gstore2<StoreT>(o1, off, field(b, f) - ev[f], field(b, f + 1) - ev[f + 1]);
                    });
                __syncthreads();  // LDS is reused by the next item
            } else {
                const FrontColor c0 = color(job.color[0]);  // P and the staged chroma hold image 0's samples
                raw = raw_tile<kO32>(job, 1, c1, tile_x, tile_y);  // in flight during image 0
                compute_tile<SPEC>(job, 0, c0, geo, tile_x, tile_y, work, l1_area, a, P, true);
#pragma unroll
                for (int f = 0; f < kFields; f++)
                    if (holds_field(f, t)) stash[stash_slot(f, t)] = field(a, f);
                unpack_raw(c1, raw, work, P);
                int nxt_color = 0;
                if (next < end) {  // the next tile's first image, in flight during image 1
                    const FrontJob& nj = jobs[front_job_of(jobs, njobs, next)];
                    raw = first_raw(nj, next - nj.tile_start);
                    nxt_color = nj.color[0];
                }
                compute_tile<SPEC>(job, 1, c1, geo, tile_x, tile_y, work, l1_area + kL1Floats, b, P, true);
                // compute_tile has passed its level-2 barrier: the chroma staging is free
                if (next < end) unpack_raw(color(nxt_color), raw, work, P);
                void* o0 = job.out[0];
                void* o1 = job.out[1];
                float* e0 = job.exact0;  // image 0 shared with later jobs: its FP32 values too
                store_tile(
                    job, geo, tile_x, tile_y,
                    [&](int plane, uint32_t off, int f) {
                        const float va = stash[stash_slot(f, t)];
                        gstore1<StoreT>(o0, off, va - kHalfOffset[plane]);
                        gstore1<StoreT>(o1, off, field(b, f) - va);
                        if (e0) gstore1<float>(e0, off, va);
                    },
                    [&](int plane, uint32_t off, int f) {
                        const float va0 = stash[stash_slot(f, t)], va1 = stash[stash_slot(f + 1, t)];
                        gstore2<StoreT>(o0, off, va0 - kHalfOffset[plane], va1 - kHalfOffset[plane]);
                        gstore2<StoreT>(o1, off, field(b, f) - va0, field(b, f + 1) - va1);
                        if (e0) gstore2<float>(e0, off, va0, va1);
                    });
                __syncthreads();  // LDS is reused by the next item
            }
        }
    }
}

// One pair whose job travels as a kernel argument (the single-pair handler:
// no upload of the job before each call). One workgroup per tile, in the
// same XCD-contiguous order.
template <typename StoreT, int SPEC>
__global__ __launch_bounds__(kFrontThreads, SPEC == kFrontGeneric ? 4 : 6)
void frontend_single_kernel(FrontJob job, int tiles, const FrontColor* __restrict__ colors,
                            const PairGeom* __restrict__ geoms)
{
    extern __shared__ float lds[];
    const int per = (tiles + kXcds - 1) / kXcds;
    const int item = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if (item >= tiles || (int)(blockIdx.x / kXcds) >= per) return;
    int tile_x, tile_y;
    front_tile_xy(job, item, tile_x, tile_y);
    frontend_tile<StoreT, SPEC>(job, colors, geoms[job.pair], tile_x, tile_y, lds);
}

// Linear RGB of image 0 of one job, stored as three W x H planes with row
// pitch `pitch` elements (the input of the crop/resize path; reference
// vshipColor.hpp:62-86 before its resize step).
__global__ __launch_bounds__(kFrontThreads)
void linear_rgb_kernel(const FrontJob* __restrict__ job_ptr, const FrontColor* __restrict__ colors,
                       float* __restrict__ out, int64_t pitch)
{
    extern __shared__ float lds[];
    const FrontJob& job = *job_ptr;
    const int tile_x = blockIdx.x % job.tiles_x, tile_y = blockIdx.x / job.tiles_x;
    float P[3][kBH][kBW];
    to_linear(job, 0, colors[job.color[0]], tile_x, tile_y, lds, P);
    const int t = threadIdx.x;
    const int x0 = tile_x * kFrontTileW + kBW * (t % kTX);
    const int y0 = tile_y * kFrontTileH + kBH * (t / kTX);
    const int64_t plane_elems = pitch * job.height;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int j = 0; j < kBH; j++)
#pragma unroll
            for (int i = 0; i < kBW; i++)
                if (x0 + i < job.width && y0 + j < job.height)
                    out[c * plane_elems + (int64_t)(y0 + j) * pitch + x0 + i] = P[c][j][i];
}

}  // namespace

hipError_t launch_linear_rgb(const FrontJob* job, int tiles, const FrontColor* colors, float* out, int64_t pitch,
                             size_t lds_bytes, hipStream_t stream)
{
    if (tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(linear_rgb_kernel, dim3(tiles), dim3(kFrontThreads), lds_bytes, stream, job, colors, out,
                       pitch);
    return hipGetLastError();
}

size_t frontend_lds_bytes(const FrontColor& fc)
{
    return sizeof(float) * (size_t)chroma_lds_floats(fc.subw, fc.subh);
}

size_t frontend_max_lds_bytes()
{
    size_t m = 0;
    for (int subw = 0; subw <= 2; subw++)
        for (int subh = 0; subh <= 1; subh++) {
            FrontColor fc = {};
            fc.subw = subw;
            fc.subh = subh;
            const size_t b = frontend_lds_bytes(fc);
            m = b > m ? b : m;
        }
    return m;
}

template <typename StoreT>
hipError_t launch_frontend(const FrontJob* jobs, const int* counts, int grid, const FrontColor* colors,
                           const PairGeom* geoms, size_t lds_bytes, int spec, int phase, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    grid = (grid + kXcds - 1) / kXcds * kXcds;  // the item mapping needs whole XCD rounds
    // chroma staging (largest layout), level 1 of each image, FP16 stash
    (void)lds_bytes;
    const size_t bytes = sizeof(float) * ((size_t)kChromaLdsMax + 2 * kL1Floats +
                                          (sizeof(StoreT) == 2 ? (size_t)kFrontStashFloats : 0));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kFrontThreads), bytes, stream, jobs, counts, colors, geoms);
    };
    switch (spec) {
    case kFrontU8Yuv420Bt709:
        if (phase == 0) go(frontend_kernel<StoreT, kFrontU8Yuv420Bt709, 0>);
        else go(frontend_kernel<StoreT, kFrontU8Yuv420Bt709, 1>);
        break;
    case kFrontU16Yuv420Bt709:
        if (phase == 0) go(frontend_kernel<StoreT, kFrontU16Yuv420Bt709, 0>);
        else go(frontend_kernel<StoreT, kFrontU16Yuv420Bt709, 1>);
        break;
    default:
        if (phase == 0) go(frontend_kernel<StoreT, kFrontGeneric, 0>);
        else go(frontend_kernel<StoreT, kFrontGeneric, 1>);
        break;
    }
    return hipGetLastError();
}

int front_spec(const FrontColor& a, const FrontColor& b)
{
    auto fixed = [](const FrontColor& c) {
        return c.subw == 1 && c.subh == 1 && c.h_cosited == 1 && c.v_cosited == 0 && c.transfer == 1 &&
               c.use_pre == 1 && c.use_post == 0 && c.pre[0] == 1.0f && c.pre[1] == 0.0f && c.pre[3] == 1.0f &&
               c.pre[6] == 1.0f && c.pre[8] == 0.0f;
    };
    if (!fixed(a) || !fixed(b) || a.sample != b.sample) return kFrontGeneric;
    if (a.sample == kSampleU8) return kFrontU8Yuv420Bt709;
    if (a.sample == kSampleU16) return kFrontU16Yuv420Bt709;
    return kFrontGeneric;
}

bool front_offsets_fit(const FrontJob& job, const FrontColor& c0, const FrontColor& c1)
{
    const FrontColor* c[2] = {&c0, &c1};
    for (int k = 0; k < 2; k++)
        for (int p = 0; p < 3; p++) {
            const int64_t s = job.stride[k][p];
            const int64_t rows = p == 0 ? job.height : ((job.height - 1) >> c[k]->subh) + 1;
            if (s <= 0 || s >= (int64_t(1) << 24) || s * rows > int64_t(0xFFFFFFFF)) return false;
        }
    return true;
}

template <typename StoreT>
hipError_t launch_frontend_single(const FrontJob& job, int tiles, const FrontColor* colors, const PairGeom* geoms,
                                  int spec, hipStream_t stream)
{
    if (tiles <= 0) return hipSuccess;
    const int grid = (tiles + kXcds - 1) / kXcds * kXcds;
    const size_t bytes = sizeof(float) * ((size_t)kChromaLdsMax + 2 * kL1Floats +
                                          (sizeof(StoreT) == 2 ? (size_t)kFrontStashFloats : 0));
    switch (spec) {
    case kFrontU8Yuv420Bt709:
        hipLaunchKernelGGL((frontend_single_kernel<StoreT, kFrontU8Yuv420Bt709>), dim3(grid), dim3(kFrontThreads),
                           bytes, stream, job, tiles, colors, geoms);
        break;
    case kFrontU16Yuv420Bt709:
        hipLaunchKernelGGL((frontend_single_kernel<StoreT, kFrontU16Yuv420Bt709>), dim3(grid), dim3(kFrontThreads),
                           bytes, stream, job, tiles, colors, geoms);
        break;
    default:
        hipLaunchKernelGGL((frontend_single_kernel<StoreT, kFrontGeneric>), dim3(grid), dim3(kFrontThreads), bytes,
                           stream, job, tiles, colors, geoms);
        break;
    }
    return hipGetLastError();
}

template hipError_t launch_frontend_single<float>(const FrontJob&, int, const FrontColor*, const PairGeom*, int,
                                                  hipStream_t);

template hipError_t launch_frontend<float>(const FrontJob*, const int*, int, const FrontColor*, const PairGeom*, size_t,
                                           int, int, hipStream_t);
template hipError_t launch_frontend<__half>(const FrontJob*, const int*, int, const FrontColor*, const PairGeom*, size_t,
                                            int, int, hipStream_t);

}  // namespace vship
