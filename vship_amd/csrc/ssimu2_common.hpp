// ssimu2_common.hpp - types shared by the host engine and the gfx950 kernels of
// the SSIMULACRA2 path.
//
// Data layout in HBM (per image of a pair): only the 11 (plane, scale) XYB
// planes the score uses are stored, each a dense row-major W_s x H_s array
// with a row pitch rounded up to 64 elements, back to back in one buffer.
// The linear-RGB pyramid of the reference (src/ssimu2/main.hpp:30-52) never
// reaches HBM: the front-end kernel builds levels 0..4 of a 64x64 level-0 tile
// in registers/LDS and writes XYB directly.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "tile_div.hpp"

namespace vship {

constexpr int kScales = 6;        // pyramid levels of the reference (main.hpp:18)
constexpr int kUsedScales = 5;    // scale 5 carries no active map
constexpr int kMaxJobsPerPair = 11;
constexpr float kC2 = 0.0009f;    // SSIM constant (score.hpp:272-273)

// Front-end tile: 64x32 level-0 pixels (whole 4x2 blocks at level 4), 512
// threads, each owning a 2x2 block.
constexpr int kXcds = 8;  // MI355X: workgroups are dispatched round-robin over 8 XCDs, each with its own L2
// Blur tile configs (blur_tile_shape): the single-pair handler uses 64x128
// strips (7) and the batch 64x256 strips (8) with 4 workgroups per CU, the
// fastest measured on MI355X for one 1080p pair and for 240 of them.
constexpr int kDefaultBlurConfig = 7;
constexpr int kDefaultBatchBlurConfig = 8;
// Batched grids, in workgroups per CU: more than fit on a CU at once (4 blur,
// 3 front-end), so the hardware dispatcher, not the static item stride,
// evens out items of unequal cost (measured on 240 1080p pairs: blur 3.76 ms
// at 4 per CU, 3.32 at 32; front-end 4.20 ms at 4, 3.55 at 32). 64 per CU
// (the tuning maximum) was no faster within run-to-run spread on 240 pairs
// and 1.5% slower on 8.
constexpr int kDefaultBatchBlurWaves = 32;
constexpr int kDefaultBatchFrontWaves = 32;
constexpr int kFrontTileW = 64;
constexpr int kFrontTileH = 32;
constexpr int kFrontThreads = 512;
constexpr int kFrontStashFloats = 11 * kFrontThreads;  // FP16 path: reference image values per tile

// Blur/map tiles: 256 threads, output tile shape chosen per launch (blur.hip),
// halo 8 on every side.
constexpr int kBlurThreads = 256;
constexpr int kHalo = 8;

// Partial sums written per blur tile: {sum ssim, sum art, sum det,
// sum ssim^4, sum art^4, sum det^4}, padded to 8 floats.
constexpr int kPartialStride = 8;

// Device-resident item counts read by the persistent kernels, so that one
// captured launch sequence serves any batch composition.
enum CountSlot : int32_t {
    kCountFrontTiles = 0,
    kCountFrontJobs = 1,
    kCountBlurTiles = 2,
    kCountBlurJobs = 3,
    kCountPairs = 4,
    // front-end groups g = phase * 3 + FrontSpec (phase 1: the jobs whose
    // image 0 a phase-0 job stores): group g covers front-end tiles
    // [counts[kCountFrontGroup + g], counts[kCountFrontGroup + g + 1])
    kCountFrontGroup = 8,
    kCountSlots = 16,
};
constexpr int kFrontGroups = 6;

enum MapFlags : int32_t {
    kMapSsim = 1,
    kMapArtifact = 2,
    kMapDetail = 4,
};

// Colour front-end description of one image: the device-side form of
// Vship_Colorspace_t (reference src/VshipColor.h:81-96) after validation. The
// reference's colour steps are folded on the host into
//   v' = (raw * scale - offset) * gain          per plane (range)
//   e  = pre * v'                               Y'CbCr -> transfer inputs
//   l  = EOTF(e)                                per channel (HLG: per triple)
//   rgb = post * l                              linear-light matrix + primaries
// (see make_image_spec in engine.cpp for the derivation per matrix).
struct FrontColor {
    int32_t sample;        // SampleKind
    uint32_t mask;         // bit mask of integer samples
    int32_t subw, subh;    // log2 chroma decimation
    int32_t h_cosited;     // horizontal chroma siting: 1 left, 0 centre
    int32_t v_cosited;     // vertical chroma siting: 1 top, 0 centre
    int32_t transfer;      // Vship_TransferFunction_t
    int32_t use_pre;       // pre is not the identity
    int32_t use_post;      // post is not the identity
    float range0[2];       // v * mul + add of plane 0
    float range12[2];      // v * mul + add of planes 1 and 2
    float pre[9];          // row-major
    float post[9];         // row-major
    // Catmull-Rom weights of the two output phases (X = 2m + phase) of a
    // 2x-subsampled axis over the 5-sample window [m-2, m+2] of the chroma
    // plane (the cosited copy phase is the weight 1): hw horizontal, vw
    // vertical. Used when subw <= 1 and subh <= 1.
    float hw[2][5];
    float vw[2][5];
};

// Pyramid geometry of one pair (both images share it).
struct PairGeom {
    int32_t w[kUsedScales];
    int32_t h[kUsedScales];
    int32_t pitch[kUsedScales];          // elements
    int64_t off[kUsedScales][3];         // element offset in the image buffer, -1 if unused
    int64_t image_elems;                 // elements per image buffer
};

// One pair for the front-end kernel: both images of a tile are processed by
// the same workgroup (the FP16 layout stores the distorted image as its
// difference from the reference).
struct FrontJob {
    const uint8_t* plane[2][3];  // device raw planes of image 0 (reference) and 1
    int64_t stride[2][3];        // bytes per row
    void* out[2];                // XYB planes of each image (StoreT elements)
    int32_t color[2];            // index into the FrontColor table
    int32_t width, height;       // level-0 size
    int32_t tiles_x;             // 64-wide tiles along x (set with set_front_tiles)
    uint32_t tiles_x_magic;      // local / tiles_x = (local * magic) >> shift (front_tile_xy)
    int32_t tiles_x_shift;
    int32_t tile_start;          // first global tile index of this job
    int32_t pair;                // index into the PairGeom table
    int32_t shared;              // 1: image 0 was computed by an earlier job
                                 // (out[0] points at its stored planes)
    // FP16 layout, image 0 shared with later jobs: the FP32 values of image 0
    // (the stored plane layout, float elements), written by the job that
    // computes image 0 and read by the jobs that share it, so that their
    // distorted planes fp16(b - a) are the same as without sharing; null
    // otherwise
    float* exact0;
};

// Sets a job's tiles_x and the round-up reciprocal that divides by it
// (tile_div.hpp).
inline void set_front_tiles(FrontJob& j, int tiles_x)
{
    j.tiles_x = tiles_x;
    tile_div_magic(tiles_x, &j.tiles_x_magic, &j.tiles_x_shift);
}

// Tile (x, y) of tile `local` of a job, from the job's reciprocal: with the
// job and local uniform across the workgroup the division is three scalar
// instructions instead of a float-reciprocal sequence on the vector unit.
__device__ inline void front_tile_xy(const FrontJob& j, int local, int& tile_x, int& tile_y)
{
    const int q = (int)(((uint64_t)(uint32_t)local * j.tiles_x_magic) >> j.tiles_x_shift);
    tile_y = q;
    tile_x = local - q * j.tiles_x;
}

// The stored XYB planes per scale (plane X, Y, B), fixed by the weight table:
// a (plane, scale) is stored iff one of its three maps has an L1 or L4 weight
// above 0.01 (reference src/ssimu2/score.hpp:4,122-131). The host derives the
// same set from the weights and checks it against this table.
constexpr bool kStored[5][3] = {
    {false, true, false}, {true, true, false}, {true, true, false}, {true, true, true}, {true, true, true}};

// FP16 layout: the reference image holds fp16(v - offset[plane]) and the
// distorted image fp16(b - a) of the FP32 values. The difference keeps the
// distortion exact to 11 bits wherever the images are close, and identical
// inputs give a zero plane; the offsets centre X', Y', B' for finer steps.
constexpr float kHalfOffset[3] = {0.42f, 0.5f, 0.55f};

// One active (pair, plane, scale) for the blur/map kernel.
struct MapJob {
    const void* a;             // XYB plane of the reference image
    const void* b;             // XYB plane of the distorted image
    int32_t w, h, pitch;
    int32_t flags;             // MapFlags
    int32_t tiles_x;
    int32_t tile_start;        // first work-item index of the job in its launch
    int32_t ntiles;
    int32_t pair;
    int32_t feature_base;      // plane*36 + scale*6
    int32_t part_start;        // partial-sum slot of the job's first item (== tile_start
                               // unless a launch covers a subset of the jobs)
};

// Blur taps (float, like the reference's GaussianHandle, gaussianblur.hpp:7-14)
// and their prefix integral for the per-axis renormalisation.
struct BlurTaps {
    float g[17];
    float integral[18];
    float inv_full;  // 1 / (integral[17] - integral[0]): all 17 taps in bounds
    // The normalised taps g[k] * inv_full (they sum to 1, so the blur of an
    // output whose 17 taps are all in bounds needs no renormalisation) as
    // FP16 hi + lo pairs for the matrix-core blur: entry k + 16 (k = -16..31)
    // holds hi | lo << 16 (hi rounded to nearest, lo = fp16(g - hi)), 0 for k
    // outside 0..16, so a Toeplitz band element is one lookup.
    uint32_t band[48];
};

// Per-pair score job for the reduction kernel.
struct ScoreJob {
    int32_t job_begin;         // first MapJob of the pair
    int32_t job_count;
    int32_t pair;              // slot of its score and features
    int32_t pad;
};

}  // namespace vship
