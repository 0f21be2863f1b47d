// engine.hpp - host side of the SSIMULACRA2 path on one MI355X.
//
// Ssimu2Plan validates two Vship_Colorspace_t descriptors and derives the
// pyramid geometry, the active (plane, scale) maps and the launch tables.
// Ssimu2Handler owns one HIP stream and every device buffer one pair needs,
// allocated once in the constructor (the reference allocates and frees
// ~100 MB per call, src/ssimu2/main.hpp:126-152); compute() issues two
// kernels (front-end, blur/maps) whose item partial sums land in pinned host
// memory, and forms the score from them in FP64 on the calling thread, as the
// reference completes its score on the host (src/ssimu2/score.hpp:380-460).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/VshipAPI.h"
#include "ssimu2_common.hpp"

namespace vship {

// Validated per-image colour description.
struct ImageSpec {
    FrontColor fc;
    int bytes;                    // bytes per sample
    int64_t width, height;        // input size
    int64_t plane_w[3], plane_h[3];
    bool resample;                // resize and/or crop before the pyramid
    int64_t target_w, target_h;   // size after the resize
    int64_t crop_left, crop_top;
    int64_t final_w, final_h;     // size the metric sees (level 0)
};

// Colour description of linear BT.709 RGB float planes (the resampled path's
// input to the pyramid kernel).
FrontColor linear_rgb_color();

ImageSpec make_image_spec(const Vship_Colorspace_t& cs);

// Blur taps computed like the reference's GaussianHandle::init
// (src/ssimu2/gaussianblur.hpp:7-14): float values and float prefix sums.
BlurTaps make_blur_taps();

// The map jobs of one pair (at most kMaxJobsPerPair), held inline so that
// building a layout allocates nothing.
struct MapJobList {
    MapJob items[kMaxJobsPerPair];
    int count = 0;
    void push_back(const MapJob& m);
    size_t size() const { return (size_t)count; }
    const MapJob* data() const { return items; }
    MapJob* data() { return items; }
    const MapJob* begin() const { return items; }
    const MapJob* end() const { return items + count; }
    MapJob* begin() { return items; }
    MapJob* end() { return items + count; }
};

// Geometry and map jobs of one pair of size w x h; xyb offsets relative to the
// start of each image's XYB buffer.
struct PairLayout {
    PairGeom geom;
    int front_tiles_per_image;
    int front_tiles_x;
    MapJobList maps;              // a/b filled with element offsets, not pointers
    int blur_tiles;
    int blur_config;
};

PairLayout make_pair_layout(int w, int h, int blur_config = kDefaultBlurConfig);

// FP64 completion of one pair's score from its blur items' partial sums
// (kPartialStride floats per item, items numbered by MapJob::tile_start):
// the 108 features and the score, with the operations and order of the GPU's
// score kernel (so both paths give the same features).
double score_from_partials(const MapJob* maps, int nmaps, const float* partials, double features[108]);

class Ssimu2Handler {
public:
    Ssimu2Handler(const Vship_Colorspace_t& src, const Vship_Colorspace_t& dis, int device);
    ~Ssimu2Handler();
    Ssimu2Handler(const Ssimu2Handler&) = delete;
    Ssimu2Handler& operator=(const Ssimu2Handler&) = delete;

    int device() const { return device_; }

    // Host planes (reference semantics): upload, run, wait, return the score.
    double compute_host(const uint8_t* const src1[3], const uint8_t* const src2[3], const int64_t ls1[3],
                        const int64_t ls2[3]);
    // Planes already in device memory of this handler's GPU.
    double compute_device(const uint8_t* const src1[3], const uint8_t* const src2[3], const int64_t ls1[3],
                          const int64_t ls2[3]);

    // Features (108 doubles) of the last compute.
    void last_features(double out[108]);
    // Dense copy of one XYB plane of the last compute (image 0 or 1).
    void debug_plane(int image, int plane, int scale, float* out, int64_t* w, int64_t* h);
    // Average time (ms) of each step over `iters` replays of the last compute:
    // the two kernels with HIP events on the handler's stream, the host's FP64
    // completion of the score with a steady clock.
    void profile(int iters, double* ms_front, double* ms_blur, double* ms_score);
    // Algorithmic work of one pair (bytes and blur-stage FLOPs).
    void work(double* bytes, double* flops) const;

private:
    // run = enqueue (the call's plane pointers into the job, every kernel on
    // the stream) + complete (wait, then the host's FP64 score)
    void run(const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3], const int64_t ls2[3]);
    void enqueue(const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3],
                 const int64_t ls2[3]);
    void complete();
    void launch_all();
    void ensure_staging(int image, const int64_t ls[3]);

    int device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t gather_done_ = nullptr;        // end of a call's host-plane read
    ImageSpec spec_[2];
    PairLayout layout_;
    BlurTaps taps_;
    int width_, height_;
    size_t front_lds_ = 0;
    int front_spec_ = 0;                      // FrontSpec of the single-pair front-end
    // front_spec_, or kFrontGeneric when this call's line sizes do not fit
    // the layout instantiation's 32-bit sample offsets
    int call_spec() const;

    void* xyb_[2] = {nullptr, nullptr};       // float planes per image
    FrontColor* d_colors_ = nullptr;
    PairGeom* d_geom_ = nullptr;
    MapJob* d_maps_ = nullptr;
    int* d_counts_ = nullptr;                  // {front tiles, blur tiles, pairs}
    FrontJob* h_front_jobs_ = nullptr;         // pinned, one pair
    double* h_score_ = nullptr;                // pinned
    // the blur items' partial sums of each call (pinned host memory the blur
    // kernel writes) and the 108 features formed from them on the host
    float* h_partials_ = nullptr;
    double h_features_[108] = {};
    // crop/resize path, per image: linear RGB at the input size, after the
    // horizontal resize, after the vertical resize (3 planes each)
    float* lin_[2] = {nullptr, nullptr};
    float* rsz_h_[2] = {nullptr, nullptr};
    float* rsz_[2] = {nullptr, nullptr};
    FrontJob* d_lin_jobs_ = nullptr;           // one per image
    FrontJob* h_lin_jobs_ = nullptr;           // pinned
    uint8_t* staging_[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    size_t staging_cap_[2][3] = {{0, 0, 0}, {0, 0, 0}};
    bool have_run_ = false;
};

}  // namespace vship
