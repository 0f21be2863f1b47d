// batch.hpp - batched SSIMULACRA2 over many frame pairs of uneven sizes.
//
// One Ssimu2Batch owns device arenas (XYB planes of every image, tile
// partials, features, scores) and a HIP graph of the three persistent kernels
// plus the score read-back. A compute call builds the work tables of all pairs
// on the host (front-end tiles, active maps, pair geometry), uploads them in
// one copy, and replays the graph: the kernels read their item counts from
// device memory, so one captured graph serves any batch composition that fits
// the arenas. Planes are stored in FP16 (or FP32), all arithmetic is FP32, and
// the norms, the 108-weight sum and the polynomial are FP64 on the GPU.
//
// The autotuner times every blur tile shape and persistent grid size on a
// caller-supplied batch and keeps the fastest; the graph is re-captured with
// it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

#include <array>
#include <map>
#include <string>
#include <tuple>

#include "../../include/VshipAPI.h"
#include "engine.hpp"
#include "ssimu2_common.hpp"

namespace vship {

struct BatchTuning {
    int blur_config = kDefaultBatchBlurConfig;  // blur tile shape (blur_tile_shape)
    int blur_waves = kDefaultBatchBlurWaves;    // blur workgroups per CU in the grid (1..64)
    int front_waves = kDefaultBatchFrontWaves;  // front-end workgroups per CU in the grid (1..64)
    int generic_front = 0;                      // 1: no layout instantiation of the front-end
};

struct BatchInput {
    int npairs;
    const Vship_Colorspace_t* src;    // npairs descriptors
    const Vship_Colorspace_t* dis;
    const uint8_t* const* planes1;    // npairs * 3 device pointers
    const uint8_t* const* planes2;
    const int64_t* ls1;               // npairs * 3 line sizes (bytes)
    const int64_t* ls2;
};

// Bit s selects scale s; the score uses scales 0-4.
constexpr int kAllScales = (1 << kUsedScales) - 1;

class Ssimu2Batch {
public:
    Ssimu2Batch(int max_pairs, bool fp16, int device);
    ~Ssimu2Batch();
    Ssimu2Batch(const Ssimu2Batch&) = delete;
    Ssimu2Batch& operator=(const Ssimu2Batch&) = delete;

    int device() const { return device_; }
    // scale_mask < kAllScales blurs only the map jobs of those scales (the
    // planes of every scale are still stored): features of the other scales
    // are 0 and the scores are those of the computed scales only.
    void compute(const BatchInput& in, double* scores, int scale_mask = kAllScales);
    // Blurs the map jobs of the scales in scale_mask for `count` pairs
    // (indices into the last compute) on the planes that compute stored, and
    // writes their full scores and features; the last compute's scales and
    // scale_mask together must cover scales 0-4. Scores and features equal a
    // full compute's bit for bit.
    void complete(int count, const int* pairs, int scale_mask, double* scores);
    void features(double* out);            // npairs * 108 of the last compute
    void autotune(const BatchInput& in, int reps, BatchTuning* chosen, double* best_ms);
    void set_tuning(const BatchTuning& t);
    BatchTuning tuning() const { return tuning_; }
    // mean device ms of each kernel and of the whole graph over `iters` replays
    void profile(int iters, double* ms_front, double* ms_blur, double* ms_score, double* ms_graph);
    void work(double* bytes, double* flops) const;

private:
    void build_tables(const BatchInput& in);
    void ensure_capacity();
    void capture();
    void launch_front_groups(hipStream_t s);  // one front-end launch per non-empty job group
    void launch_kernels(hipStream_t s);

    int device_;
    bool fp16_;
    int max_pairs_;
    int cus_ = 256;
    hipStream_t stream_ = nullptr;
    BatchTuning tuning_;
    BlurTaps taps_;

    // host tables of the current batch
    std::vector<FrontJob> fjobs_;
    std::vector<FrontColor> colors_;
    std::vector<PairGeom> geoms_;
    std::vector<MapJob> maps_;       // every map job of the batch, a/b as arena offsets
    std::vector<MapJob> dev_maps_;   // the same with device plane pointers
    std::vector<ScoreJob> sjobs_;
    std::vector<int64_t> xyb_off_;   // per image (2 per pair): element offset in the arena
    // per pair, FP16 layout: arena offset (FP16 elements) of the FP32 values of
    // the shared image 0 its front-end job writes or reads, -1 if none
    std::vector<int64_t> exact_off_;
    // descriptors and layouts of earlier calls (a batch usually repeats a few
    // colourspaces and sizes): keyed by the descriptor's bytes and by
    // (width, height, blur config)
    std::map<std::string, ImageSpec> spec_cache_;
    std::map<std::tuple<int, int, int>, PairLayout> layout_cache_;
    const ImageSpec& cached_spec(const Vship_Colorspace_t& cs);
    const PairLayout& cached_layout(int w, int h, int blur_config);
    int64_t xyb_elems_ = 0;
    int front_tiles_ = 0, blur_tiles_ = 0;
    int table_pairs_ = 0;  // pairs of the tables built last
    int npairs_ = 0;       // pairs of the last compute that completed, 0 if none
    int last_mask_ = 0;    // scales blurred by that compute
    int call_mask_ = kAllScales;  // scales the tables being built run
    std::vector<int> pair_jobs_;  // per pair: first index into maps_ (and count in maps_ order)
    double work_bytes_ = 0.0, work_flops_ = 0.0;
    // crop/resize: per resampled image its linear-RGB job and its buffers in
    // the resample arena (float offsets; the resized planes are the pyramid
    // job's input, read from the cropped window)
    struct Resample {
        int64_t width, height, target_w, target_h;
        int64_t lin, horiz, vert;    // float offsets of the linear, h-resized and resized planes (-1: none)
    };
    std::vector<Resample> resamples_;
    std::vector<FrontJob> rjobs_;                // linear-RGB jobs, one per entry of resamples_
    std::vector<std::array<int64_t, 3>> window_; // per image (2 per pair): float offset of each plane's window, -1
    int64_t rsz_elems_ = 0;
    void run_resamples();

    // device
    void* xyb_ = nullptr;
    int64_t xyb_cap_ = 0;            // elements
    float* rsz_ = nullptr;           // resample arena
    int64_t rsz_cap_ = 0;            // floats
    FrontJob* d_rjobs_ = nullptr;    // linear-RGB jobs (2 per pair at most)
    FrontJob* h_rjobs_ = nullptr;    // pinned staging
    float* partials_ = nullptr;
    int64_t partials_cap_ = 0;       // tiles
    uint8_t* tables_ = nullptr;      // packed tables
    size_t tables_cap_ = 0;
    uint8_t* h_tables_ = nullptr;    // pinned staging of the packed tables
    double* d_features_ = nullptr;
    double* d_scores_ = nullptr;
    double* h_scores_ = nullptr;     // pinned
    // table views inside tables_: the blur launch reads the jobs of the
    // scales it runs (d_bmaps_), the score kernel every job (d_maps_, flags
    // cleared for jobs that did not run); the completion's own tables follow
    int* d_counts_ = nullptr;
    FrontJob* d_fjobs_ = nullptr;
    FrontColor* d_colors_ = nullptr;
    PairGeom* d_geoms_ = nullptr;
    MapJob* d_maps_ = nullptr;
    MapJob* d_bmaps_ = nullptr;
    ScoreJob* d_sjobs_ = nullptr;
    int* d_ccounts_ = nullptr;
    MapJob* d_cmaps_ = nullptr;
    MapJob* d_cbmaps_ = nullptr;
    ScoreJob* d_csjobs_ = nullptr;
    int cap_fjobs_ = 0, cap_maps_ = 0;

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    bool graph_valid_ = false;
    int front_groups_[kFrontGroups + 1] = {};  // first front-end tile of each job group, then the total
    unsigned front_mask_ = 0;                  // bit g: group g has jobs (a captured launch)
};

}  // namespace vship
