// kernels.hpp - host-side launchers of the gfx950 kernels.
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

// LDS bytes of the chroma staging for images of this colour description (the
// linear-RGB kernel's dynamic LDS; the front-end kernels size theirs), and the
// largest over every supported chroma subsampling.
size_t frontend_lds_bytes(const FrontColor& fc);
size_t frontend_max_lds_bytes();

// Colour layouts with a front-end instantiation of their own (8- and 16-bit
// YUV 4:2:0 with left-sited chroma, BT.709 transfer, no linear-light
// matrix); kFrontGeneric handles every layout.
enum FrontSpec : int {
    kFrontGeneric = 0,
    kFrontU8Yuv420Bt709 = 1,
    kFrontU16Yuv420Bt709 = 2,
};
// The instantiation serving a pair of images with these descriptors.
int front_spec(const FrontColor& a, const FrontColor& b);
// Whether every sample row of the job's six planes starts within 2^32 bytes
// of its plane's base with a stride below 2^24 (the layout instantiations
// address samples with 32-bit offsets; callers use kFrontGeneric otherwise).
bool front_offsets_fit(const FrontJob& job, const FrontColor& c0, const FrontColor& c1);

// counts: device array indexed by CountSlot. With fewer workgroups than items
// the kernels loop over items (persistent mode).
// spec, phase: the job group this launch runs, tiles
// [counts[kCountFrontGroup + g], counts[kCountFrontGroup + g + 1]) with
// g = phase * 3 + spec: jobs whose images both have layout spec (generic: any
// layout), and, in phase 1, whose image 0 (FrontJob::shared) a phase-0 job
// stores, which must have completed (stream order).
template <typename StoreT>
hipError_t launch_frontend(const FrontJob* jobs, const int* counts, int grid, const FrontColor* colors,
                           const PairGeom* geoms, size_t lds_bytes, int spec, int phase, hipStream_t stream);


// One pair's tiles (tiles = the job's 64x32 tile count) with the job passed
// by value as a kernel argument, one workgroup per tile.
template <typename StoreT>
hipError_t launch_frontend_single(const FrontJob& job, int tiles, const FrontColor* colors, const PairGeom* geoms,
                                  int spec, hipStream_t stream);

// Linear RGB of image 0 of `job` (tiles = job's 64x32 tile count) into three
// planes of job.width x job.height floats, row pitch `pitch` elements.
hipError_t launch_linear_rgb(const FrontJob* job, int tiles, const FrontColor* colors, float* out, int64_t pitch,
                             size_t lds_bytes, hipStream_t stream);

// Separable cubic-Hermite (Catmull-Rom) resize of three float planes,
// reference src/gpuColorToLinear/resize.hpp:26-65: horizontal sw -> rw on h
// rows, then vertical sh -> rh on w columns (dense planes).
hipError_t launch_resize_h(float* dst, const float* src, int sw, int h, int rw, hipStream_t stream);
hipError_t launch_resize_v(float* dst, const float* src, int w, int sh, int rh, hipStream_t stream);

// Blur/map tile shapes: config 0 = 64x32, 1 = 64x16, 2 = 32x64,
// 3 = 32x32, 4 = 64x64 (FP32 VALU kernel); 5 = 64x48 (matrix-core kernel,
// blur_mfma.hip); 6 = 64x64, 7 = 64x128, 8 = 64x256 (streamed matrix-core
// kernel, blur_strip.hip).
struct BlurTile {
    int tw, th;
};
BlurTile blur_tile_shape(int config);
int blur_tile_configs();
// Non-persistent grid covering total_tiles (a multiple of the XCD count).
int blur_grid_for(int total_tiles);

// grid must be a multiple of 8; with fewer workgroups than tiles the kernel
// loops over tiles (persistent mode).
template <typename StoreT>
hipError_t launch_blur(int config, const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps,
                       float* partials, hipStream_t stream);

// Matrix-core variant of launch_blur (tile 64x48, blur config 5).
template <typename StoreT>
hipError_t launch_blur_mfma(const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps, float* partials,
                            hipStream_t stream);

// Streamed matrix-core variant (items of 64 x 16*segb outputs; configs 6-8).
template <typename StoreT>
hipError_t launch_blur_strip(int segb, const MapJob* jobs, const int* counts, int grid, const BlurTaps& taps,
                             float* partials, hipStream_t stream);

hipError_t launch_score(const ScoreJob* pairs, const int* counts, int grid, const MapJob* jobs,
                        const float* partials, double* features, double* scores, hipStream_t stream);

// Plain 16-byte-vector copy (bandwidth probe and counter calibration).
hipError_t launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t stream);

// Planes in pinned host memory (their device addresses) copied into device
// buffers by one kernel; every address 16-byte aligned.
constexpr int kGatherPlanes = 6;
struct GatherPlanes {
    const uint8_t* src[kGatherPlanes];
    uint8_t* dst[kGatherPlanes];
    int64_t bytes[kGatherPlanes];
    int n;
};
hipError_t launch_gather(const GatherPlanes& g, hipStream_t stream);

}  // namespace vship
