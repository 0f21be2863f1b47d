// batch.cpp - batched, graph-replayed SSIMULACRA2 over pairs of uneven sizes.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "colorspace_key.hpp"
#include "errors.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

constexpr int kScoreGrid = 256;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// scale of a map job, from its feature index plane*36 + scale*6
int job_scale(const MapJob& m) { return (m.feature_base % 36) / 6; }

}  // namespace

Ssimu2Batch::Ssimu2Batch(int max_pairs, bool fp16, int device)
    : device_(device), fp16_(fp16), max_pairs_(max_pairs)
{
    if (max_pairs <= 0 || max_pairs > (1 << 20)) VSHIP_THROW(Vship_BadPointer, "max_pairs must be in [1, 2^20]");
    taps_ = make_blur_taps();
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    hipDeviceProp_t prop;
    VSHIP_HIP_CHECK(hipGetDeviceProperties(&prop, device_));
    cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    VSHIP_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));

    // fixed-capacity packed tables: counts | front jobs | colours | geoms | maps | score jobs
    cap_fjobs_ = max_pairs;
    cap_maps_ = kMaxJobsPerPair * max_pairs;
    size_t off = 0;
    const size_t o_counts = off; off = align_up(off + kCountSlots * sizeof(int), 256);
    const size_t o_fjobs = off; off = align_up(off + cap_fjobs_ * sizeof(FrontJob), 256);
    // colours: each image's own, and the linear-RGB input of a resampled image's pyramid job
    const size_t o_colors = off; off = align_up(off + 4 * cap_fjobs_ * sizeof(FrontColor), 256);
    const size_t o_geoms = off; off = align_up(off + max_pairs * sizeof(PairGeom), 256);
    const size_t o_maps = off; off = align_up(off + cap_maps_ * sizeof(MapJob), 256);
    const size_t o_bmaps = off; off = align_up(off + cap_maps_ * sizeof(MapJob), 256);
    const size_t o_sjobs = off; off = align_up(off + max_pairs * sizeof(ScoreJob), 256);
    // the completion's tables
    const size_t o_ccounts = off; off = align_up(off + kCountSlots * sizeof(int), 256);
    const size_t o_cmaps = off; off = align_up(off + cap_maps_ * sizeof(MapJob), 256);
    const size_t o_cbmaps = off; off = align_up(off + cap_maps_ * sizeof(MapJob), 256);
    const size_t o_csjobs = off; off = align_up(off + max_pairs * sizeof(ScoreJob), 256);
    tables_cap_ = off;
    if (hipMalloc((void**)&tables_, tables_cap_) != hipSuccess ||
        hipMalloc((void**)&d_features_, (size_t)max_pairs * 108 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_scores_, (size_t)max_pairs * sizeof(double)) != hipSuccess)
        VSHIP_THROW(Vship_OutOfVRAM, "batch table allocation failed");
    if (hipHostMalloc((void**)&h_tables_, tables_cap_, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h_scores_, (size_t)max_pairs * sizeof(double), hipHostMallocDefault) != hipSuccess)
        VSHIP_THROW(Vship_OutOfRAM, "batch pinned allocation failed");
    d_counts_ = reinterpret_cast<int*>(tables_ + o_counts);
    d_fjobs_ = reinterpret_cast<FrontJob*>(tables_ + o_fjobs);
    d_colors_ = reinterpret_cast<FrontColor*>(tables_ + o_colors);
    d_geoms_ = reinterpret_cast<PairGeom*>(tables_ + o_geoms);
    d_maps_ = reinterpret_cast<MapJob*>(tables_ + o_maps);
    d_bmaps_ = reinterpret_cast<MapJob*>(tables_ + o_bmaps);
    d_sjobs_ = reinterpret_cast<ScoreJob*>(tables_ + o_sjobs);
    d_ccounts_ = reinterpret_cast<int*>(tables_ + o_ccounts);
    d_cmaps_ = reinterpret_cast<MapJob*>(tables_ + o_cmaps);
    d_cbmaps_ = reinterpret_cast<MapJob*>(tables_ + o_cbmaps);
    d_csjobs_ = reinterpret_cast<ScoreJob*>(tables_ + o_csjobs);
}

Ssimu2Batch::~Ssimu2Batch()
{
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    (void)hipFree(xyb_);
    (void)hipFree(rsz_);
    (void)hipFree(d_rjobs_);
    (void)hipHostFree(h_rjobs_);
    (void)hipFree(partials_);
    (void)hipFree(tables_);
    (void)hipFree(d_features_);
    (void)hipFree(d_scores_);
    (void)hipHostFree(h_tables_);
    (void)hipHostFree(h_scores_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

const ImageSpec& Ssimu2Batch::cached_spec(const Vship_Colorspace_t& cs)
{
    std::string key = colorspace_key(cs);
    auto it = spec_cache_.find(key);
    if (it != spec_cache_.end()) return it->second;
    if (spec_cache_.size() > 256) spec_cache_.clear();
    return spec_cache_.emplace(std::move(key), make_image_spec(cs)).first->second;
}

const PairLayout& Ssimu2Batch::cached_layout(int w, int h, int blur_config)
{
    const auto key = std::make_tuple(w, h, blur_config);
    auto it = layout_cache_.find(key);
    if (it != layout_cache_.end()) return it->second;
    if (layout_cache_.size() > 256) layout_cache_.clear();
    return layout_cache_.emplace(key, make_pair_layout(w, h, blur_config)).first->second;
}

void Ssimu2Batch::build_tables(const BatchInput& in)
{
    if (in.npairs <= 0 || in.npairs > max_pairs_) VSHIP_THROW(Vship_BadPointer, "npairs outside [1, max_pairs]");
    fjobs_.clear();
    colors_.clear();
    geoms_.clear();
    maps_.clear();
    sjobs_.clear();
    xyb_off_.clear();
    pair_jobs_.clear();
    resamples_.clear();
    rjobs_.clear();
    window_.assign(2 * (size_t)in.npairs, std::array<int64_t, 3>{-1, -1, -1});
    rsz_elems_ = 0;
    exact_off_.assign(in.npairs, -1);
    const size_t elem = fp16_ ? 2 : 4;
    int64_t elems = 0;
    int btiles = 0;
    double in_bytes = 0.0, active = 0.0, written = 0.0, flops = 0.0, blocks = 0.0;
    // pairs whose reference image (descriptor, planes, line sizes) repeats an
    // earlier pair's share its stored pyramid: their front-end job computes
    // image 1 only, after the first pair's job (phase 1). In the FP16 layout
    // the first pair's job also keeps image 0's FP32 values for them.
    std::map<std::string, int> first_of;
    std::vector<FrontJob> own_jobs, shared_jobs;
    for (int i = 0; i < in.npairs; i++) {
        const ImageSpec s1 = cached_spec(in.src[i]);
        const ImageSpec s2 = cached_spec(in.dis[i]);
        // the size the metric sees: after the resize and crop of a resampled image
        if (s1.final_w != s2.final_w || s1.final_h != s2.final_h)
            VSHIP_THROW(Vship_DifferingInputType, "pair " + std::to_string(i) + ": sizes differ");
        const ImageSpec* specs[2] = {&s1, &s2};
        const uint8_t* const* pl[2] = {in.planes1 + 3 * i, in.planes2 + 3 * i};
        const int64_t* ls[2] = {in.ls1 + 3 * i, in.ls2 + 3 * i};
        for (int k = 0; k < 2; k++)
            for (int p = 0; p < 3; p++) {
                if (pl[k][p] == nullptr) VSHIP_THROW(Vship_BadPointer, "null plane pointer");
                if (ls[k][p] < specs[k]->plane_w[p] * specs[k]->bytes)
                    VSHIP_THROW(Vship_BadPointer, "lineSize smaller than one row of samples");
                in_bytes += (double)specs[k]->plane_w[p] * specs[k]->plane_h[p] * specs[k]->bytes;
            }
        std::string key = colorspace_key(in.src[i]);
        key.append(reinterpret_cast<const char*>(pl[0]), 3 * sizeof(pl[0][0]));
        key.append(reinterpret_cast<const char*>(ls[0]), 3 * sizeof(ls[0][0]));
        const auto found = first_of.find(key);
        const int ref = found == first_of.end() ? -1 : found->second;
        if (ref < 0) first_of.emplace(std::move(key), i);
        const PairLayout& L = cached_layout((int)s1.final_w, (int)s1.final_h, tuning_.blur_config);
        geoms_.push_back(L.geom);
        FrontJob j;
        std::memset(&j, 0, sizeof(j));
        j.shared = ref >= 0 ? 1 : 0;
        if (j.shared) {  // the raw reference planes are not read again
            for (int p = 0; p < 3; p++) in_bytes -= (double)s1.plane_w[p] * s1.plane_h[p] * s1.bytes;
            if (fp16_ && exact_off_[ref] < 0) {
                // FP32 elements take two FP16 slots of the arena
                elems = (int64_t)align_up((size_t)elems, 256 / elem);
                exact_off_[ref] = elems;
                elems += 2 * L.geom.image_elems;
            }
            exact_off_[i] = exact_off_[ref];
        }
        for (int k = 0; k < 2; k++) {
            if (k == 0 && j.shared) {
                xyb_off_.push_back(xyb_off_[2 * ref]);
            } else {
                // keep every image buffer 256-byte aligned
                elems = (int64_t)align_up((size_t)elems, 256 / elem);
                xyb_off_.push_back(elems);
                elems += L.geom.image_elems;
            }
            for (int p = 0; p < 3; p++) {
                j.plane[k][p] = pl[k][p];
                j.stride[k][p] = ls[k][p];
            }
            j.color[k] = (int32_t)colors_.size();
            colors_.push_back(specs[k]->fc);
            if (specs[k]->resample && !(k == 0 && j.shared)) {
                // crop/resize (reference vshipColor.hpp:87-103): linear RGB at
                // the input size, resized; the pyramid job reads the cropped
                // window of the resized planes as float linear RGB
                const ImageSpec& sp = *specs[k];
                Resample r;
                r.width = sp.width;
                r.height = sp.height;
                r.target_w = sp.target_w;
                r.target_h = sp.target_h;
                r.lin = rsz_elems_;
                rsz_elems_ += (int64_t)align_up((size_t)(3 * sp.width * sp.height), 64);
                int64_t cur = r.lin;
                r.horiz = r.vert = -1;
                if (sp.target_w != sp.width) {
                    r.horiz = cur = rsz_elems_;
                    rsz_elems_ += (int64_t)align_up((size_t)(3 * sp.target_w * sp.height), 64);
                }
                if (sp.target_h != sp.height) {
                    r.vert = cur = rsz_elems_;
                    rsz_elems_ += (int64_t)align_up((size_t)(3 * sp.target_w * sp.target_h), 64);
                }
                FrontJob lj;
                std::memset(&lj, 0, sizeof(lj));
                for (int p = 0; p < 3; p++) {
                    lj.plane[0][p] = pl[k][p];
                    lj.stride[0][p] = ls[k][p];
                }
                lj.color[0] = j.color[k];
                lj.width = (int32_t)sp.width;
                lj.height = (int32_t)sp.height;
                set_front_tiles(lj, (int)((sp.width + kFrontTileW - 1) / kFrontTileW));
                resamples_.push_back(r);
                rjobs_.push_back(lj);
                for (int p = 0; p < 3; p++) {
                    window_[2 * (size_t)i + k][p] = cur + p * sp.target_w * sp.target_h + sp.crop_top * sp.target_w +
                                                    sp.crop_left;
                    j.stride[k][p] = sp.target_w * (int64_t)sizeof(float);
                }
                j.color[k] = (int32_t)colors_.size();
                colors_.push_back(linear_rgb_color());
            }
        }
        j.width = (int32_t)s1.final_w;
        j.height = (int32_t)s1.final_h;
        set_front_tiles(j, L.front_tiles_x);
        j.tile_start = L.front_tiles_per_image;  // tile count until the jobs are ordered
        j.pair = i;
        (j.shared ? shared_jobs : own_jobs).push_back(j);
        ScoreJob sj = {(int32_t)maps_.size(), (int32_t)L.maps.size(), (int32_t)i, 0};
        sjobs_.push_back(sj);
        pair_jobs_.push_back((int)maps_.size());
        for (const MapJob& m0 : L.maps) {
            MapJob m = m0;
            m.pair = i;
            m.tile_start = btiles + m0.tile_start;
            m.part_start = m.tile_start;
            // a/b hold element offsets until the arena is known
            m.a = reinterpret_cast<const void*>(xyb_off_[2 * i] + reinterpret_cast<int64_t>(m0.a));
            m.b = reinterpret_cast<const void*>(xyb_off_[2 * i + 1] + reinterpret_cast<int64_t>(m0.b));
            maps_.push_back(m);
            const double px = (double)m.w * m.h;
            active += px;
            written += j.shared ? px : 2.0 * px;  // a shared image 0 is stored once
            const bool ssim = m.flags & kMapSsim;
            const bool edge = m.flags & (kMapArtifact | kMapDetail);
            const int n_edge = ((m.flags & kMapArtifact) ? 1 : 0) + ((m.flags & kMapDetail) ? 1 : 0);
            flops += px * (70.0 * (ssim ? 4 : 2) + 23.0 * ssim + 10.0 * edge + 4.0 * n_edge);
            blocks += std::ceil(m.w / 16.0) * std::ceil(m.h / 16.0);
        }
        btiles += L.blur_tiles;
    }
    // front-end jobs in groups g = phase * 3 + layout instantiation: the pairs
    // that compute both images (phase 0) before those whose image 0 a phase-0
    // job stores (phase 1), each phase by layout, so that every pair of a
    // mixed batch runs its layout's instantiation (one launch per group)
    std::vector<FrontJob> groups[kFrontGroups];
    for (auto* list : {&own_jobs, &shared_jobs}) {
        for (const FrontJob& j : *list) {
            const FrontColor& c0 = colors_[j.color[0]];
            const FrontColor& c1 = colors_[j.color[1]];
            int spec = front_spec(c0, c1);
            if (spec != kFrontGeneric && (tuning_.generic_front || !front_offsets_fit(j, c0, c1)))
                spec = kFrontGeneric;
            groups[(list == &shared_jobs ? 3 : 0) + spec].push_back(j);
        }
    }
    int ftiles = 0;
    unsigned mask = 0;
    for (int g = 0; g < kFrontGroups; g++) {
        front_groups_[g] = ftiles;
        if (!groups[g].empty()) mask |= 1u << g;
        for (FrontJob j : groups[g]) {
            const int n = j.tile_start;
            j.tile_start = ftiles;
            ftiles += n;
            fjobs_.push_back(j);
        }
    }
    front_groups_[kFrontGroups] = ftiles;
    // the graph holds one launch per non-empty group
    if (mask != front_mask_) {
        front_mask_ = mask;
        graph_valid_ = false;
    }
    xyb_elems_ = elems;
    front_tiles_ = ftiles;
    blur_tiles_ = btiles;
    table_pairs_ = in.npairs;
    // raw inputs read, stored planes written by the front-end and read twice
    // (both images) by the blur, 48 B of partials per 16x16 block
    work_bytes_ = in_bytes + (double)elem * (written + 2.0 * active) + 48.0 * blocks;
    work_flops_ = flops;
}

void Ssimu2Batch::ensure_capacity()
{
    const size_t elem = fp16_ ? 2 : 4;
    if (xyb_elems_ > xyb_cap_) {
        (void)hipFree(xyb_);
        xyb_ = nullptr;
        xyb_cap_ = 0;
        graph_valid_ = false;
        const int64_t want = xyb_elems_ + xyb_elems_ / 8;  // headroom for the next batches
        if (hipMalloc(&xyb_, (size_t)want * elem) != hipSuccess)
            VSHIP_THROW(Vship_OutOfVRAM, "batch XYB arena allocation failed");
        xyb_cap_ = want;
    }
    if (blur_tiles_ > partials_cap_) {
        (void)hipFree(partials_);
        partials_ = nullptr;
        partials_cap_ = 0;
        const int64_t want = blur_tiles_ + blur_tiles_ / 8 + 64;
        if (hipMalloc((void**)&partials_, (size_t)want * kPartialStride * sizeof(float)) != hipSuccess)
            VSHIP_THROW(Vship_OutOfVRAM, "batch partial-sum allocation failed");
        partials_cap_ = want;
        graph_valid_ = false;  // the kernels take the partials pointer as an argument
    }
    if (rsz_elems_ > rsz_cap_) {
        (void)hipFree(rsz_);
        rsz_ = nullptr;
        rsz_cap_ = 0;
        const int64_t want = rsz_elems_ + rsz_elems_ / 8;
        if (hipMalloc((void**)&rsz_, (size_t)want * sizeof(float)) != hipSuccess)
            VSHIP_THROW(Vship_OutOfVRAM, "batch resample arena allocation failed");
        rsz_cap_ = want;
    }
    if (!rjobs_.empty() && d_rjobs_ == nullptr) {
        if (hipMalloc((void**)&d_rjobs_, 2 * (size_t)max_pairs_ * sizeof(FrontJob)) != hipSuccess)
            VSHIP_THROW(Vship_OutOfVRAM, "batch resample job allocation failed");
        if (hipHostMalloc((void**)&h_rjobs_, 2 * (size_t)max_pairs_ * sizeof(FrontJob), hipHostMallocDefault) !=
            hipSuccess)
            VSHIP_THROW(Vship_OutOfRAM, "batch pinned allocation failed");
    }
    // final pointers into the arena, then one packed upload
    for (FrontJob& j : fjobs_) {
        for (int k = 0; k < 2; k++) j.out[k] = static_cast<uint8_t*>(xyb_) + xyb_off_[2 * j.pair + k] * elem;
        for (int k = 0; k < 2; k++) {
            const auto& w = window_[2 * (size_t)j.pair + k];
            if (w[0] < 0 || (k == 0 && j.shared)) continue;
            for (int p = 0; p < 3; p++) j.plane[k][p] = reinterpret_cast<const uint8_t*>(rsz_ + w[p]);
        }
        const int64_t e = exact_off_[j.pair];
        j.exact0 = e < 0 ? nullptr : reinterpret_cast<float*>(static_cast<uint8_t*>(xyb_) + e * elem);
    }
    dev_maps_ = maps_;
    for (auto& m : dev_maps_) {
        m.a = static_cast<const uint8_t*>(xyb_) + reinterpret_cast<int64_t>(m.a) * elem;
        m.b = static_cast<const uint8_t*>(xyb_) + reinterpret_cast<int64_t>(m.b) * elem;
    }
    // the blur launch takes the jobs of the scales this call runs, numbered
    // afresh (their partial-sum slots keep the numbering over every job); the
    // score kernel takes every job, with no map flag on those that do not run
    std::vector<MapJob> bmaps, smaps = dev_maps_;
    int btotal = 0;
    for (size_t k = 0; k < dev_maps_.size(); k++) {
        if (call_mask_ & (1 << job_scale(dev_maps_[k]))) {
            MapJob m = dev_maps_[k];
            m.tile_start = btotal;
            btotal += m.ntiles;
            bmaps.push_back(m);
        } else {
            smaps[k].flags = 0;
        }
    }
    int counts[kCountSlots] = {};
    counts[kCountFrontTiles] = front_tiles_;
    counts[kCountFrontJobs] = (int)fjobs_.size();
    counts[kCountBlurTiles] = btotal;
    counts[kCountBlurJobs] = (int)bmaps.size();
    counts[kCountPairs] = table_pairs_;
    for (int g = 0; g <= kFrontGroups; g++) counts[kCountFrontGroup + g] = front_groups_[g];
    auto put = [&](const void* dev, const void* src, size_t bytes) {
        std::memcpy(h_tables_ + (static_cast<const uint8_t*>(dev) - tables_), src, bytes);
    };
    put(d_counts_, counts, sizeof(counts));
    put(d_fjobs_, fjobs_.data(), fjobs_.size() * sizeof(FrontJob));
    put(d_colors_, colors_.data(), colors_.size() * sizeof(FrontColor));
    put(d_geoms_, geoms_.data(), geoms_.size() * sizeof(PairGeom));
    put(d_maps_, smaps.data(), smaps.size() * sizeof(MapJob));
    put(d_bmaps_, bmaps.data(), bmaps.size() * sizeof(MapJob));
    put(d_sjobs_, sjobs_.data(), sjobs_.size() * sizeof(ScoreJob));
    const size_t used = (reinterpret_cast<uint8_t*>(d_sjobs_) - tables_) + sjobs_.size() * sizeof(ScoreJob);
    VSHIP_HIP_CHECK(hipMemcpyAsync(tables_, h_tables_, used, hipMemcpyHostToDevice, stream_));
}

void Ssimu2Batch::launch_front_groups(hipStream_t s)
{
    const int fgrid = std::max(1, tuning_.front_waves) * cus_;
    const size_t lds = frontend_max_lds_bytes();
    for (int g = 0; g < kFrontGroups; g++) {
        if (!(front_mask_ & (1u << g))) continue;
        if (fp16_)
            VSHIP_HIP_CHECK(
                launch_frontend<__half>(d_fjobs_, d_counts_, fgrid, d_colors_, d_geoms_, lds, g % 3, g / 3, s));
        else
            VSHIP_HIP_CHECK(
                launch_frontend<float>(d_fjobs_, d_counts_, fgrid, d_colors_, d_geoms_, lds, g % 3, g / 3, s));
    }
}

void Ssimu2Batch::launch_kernels(hipStream_t s)
{
    const int bgrid = (std::max(1, tuning_.blur_waves) * cus_ + 7) / 8 * 8;
    launch_front_groups(s);
    if (fp16_)
        VSHIP_HIP_CHECK(launch_blur<__half>(tuning_.blur_config, d_bmaps_, d_counts_, bgrid, taps_, partials_, s));
    else
        VSHIP_HIP_CHECK(launch_blur<float>(tuning_.blur_config, d_bmaps_, d_counts_, bgrid, taps_, partials_, s));
    VSHIP_HIP_CHECK(launch_score(d_sjobs_, d_counts_, kScoreGrid, d_maps_, partials_, d_features_, d_scores_, s));
}

void Ssimu2Batch::capture()
{
    if (exec_) { (void)hipGraphExecDestroy(exec_); exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    VSHIP_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        launch_kernels(stream_);
        VSHIP_HIP_CHECK(hipMemcpyAsync(h_scores_, d_scores_, (size_t)max_pairs_ * sizeof(double),
                                       hipMemcpyDeviceToHost, stream_));
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    VSHIP_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
    VSHIP_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    graph_valid_ = true;
}

// Crop/resize of the resampled images (linear RGB at the input size, then
// the horizontal and vertical resizes, as the single-pair handler does),
// launched on the batch stream after the table upload and before the graph,
// whose front-end reads the resized planes: one set of launches per
// resampled image, outside the captured graph because their sizes vary
// from batch to batch.
void Ssimu2Batch::run_resamples()
{
    if (resamples_.empty()) return;
    std::memcpy(h_rjobs_, rjobs_.data(), rjobs_.size() * sizeof(FrontJob));
    VSHIP_HIP_CHECK(
        hipMemcpyAsync(d_rjobs_, h_rjobs_, rjobs_.size() * sizeof(FrontJob), hipMemcpyHostToDevice, stream_));
    const size_t lds = frontend_max_lds_bytes();
    for (size_t r = 0; r < resamples_.size(); r++) {
        const Resample& x = resamples_[r];
        const int W = (int)x.width, H = (int)x.height, tw = (int)x.target_w, th = (int)x.target_h;
        const int tiles = rjobs_[r].tiles_x * ((H + kFrontTileH - 1) / kFrontTileH);
        VSHIP_HIP_CHECK(launch_linear_rgb(d_rjobs_ + r, tiles, d_colors_, rsz_ + x.lin, W, lds, stream_));
        const float* cur = rsz_ + x.lin;
        if (x.horiz >= 0) {
            VSHIP_HIP_CHECK(launch_resize_h(rsz_ + x.horiz, cur, W, H, tw, stream_));
            cur = rsz_ + x.horiz;
        }
        if (x.vert >= 0) VSHIP_HIP_CHECK(launch_resize_v(rsz_ + x.vert, cur, tw, H, th, stream_));
    }
}

void Ssimu2Batch::compute(const BatchInput& in, double* scores, int scale_mask)
{
    // features(), profile() and complete() refer to the last call that
    // completed: none until this one has
    npairs_ = 0;
    if (scale_mask <= 0 || scale_mask > kAllScales)
        VSHIP_THROW(Vship_BadPointer, "scale_mask must select scales among 0..4 (bits 0x1f)");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    call_mask_ = scale_mask;
    build_tables(in);
    ensure_capacity();
    run_resamples();
    if (!graph_valid_) capture();
    VSHIP_HIP_CHECK(hipGraphLaunch(exec_, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    npairs_ = table_pairs_;
    last_mask_ = scale_mask;
    std::memcpy(scores, h_scores_, (size_t)npairs_ * sizeof(double));
}

void Ssimu2Batch::complete(int count, const int* pairs, int scale_mask, double* scores)
{
    if (npairs_ == 0) VSHIP_THROW(Vship_BadPointer, "completion needs one successful compute call first");
    if (scale_mask <= 0 || scale_mask > kAllScales)
        VSHIP_THROW(Vship_BadPointer, "scale_mask must select scales among 0..4 (bits 0x1f)");
    if ((scale_mask | last_mask_) != kAllScales)
        VSHIP_THROW(Vship_BadPointer, "the last compute's scales and scale_mask must together cover scales 0..4");
    if (count <= 0 || count > npairs_ || !pairs || !scores)
        VSHIP_THROW(Vship_BadPointer, "count outside [1, pairs of the last compute] or null array");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    std::vector<char> seen((size_t)npairs_, 0);
    std::vector<MapJob> bmaps, smaps;
    std::vector<ScoreJob> sj;
    int btotal = 0;
    for (int c = 0; c < count; c++) {
        const int i = pairs[c];
        if (i < 0 || i >= npairs_) VSHIP_THROW(Vship_BadPointer, "pair index outside the last compute");
        if (seen[(size_t)i]) VSHIP_THROW(Vship_BadPointer, "pair index listed twice");
        seen[(size_t)i] = 1;
        const int begin = pair_jobs_[(size_t)i], n = sjobs_[(size_t)i].job_count;
        sj.push_back(ScoreJob{(int32_t)smaps.size(), n, i, 0});
        for (int k = 0; k < n; k++) {
            const MapJob& m = dev_maps_[(size_t)(begin + k)];
            smaps.push_back(m);
            if (scale_mask & (1 << job_scale(m))) {
                MapJob b = m;
                b.tile_start = btotal;
                btotal += b.ntiles;
                bmaps.push_back(b);
            }
        }
    }
    int counts[kCountSlots] = {};
    counts[kCountBlurTiles] = btotal;
    counts[kCountBlurJobs] = (int)bmaps.size();
    counts[kCountPairs] = count;
    auto put = [&](const void* dev, const void* src, size_t bytes) {
        std::memcpy(h_tables_ + (static_cast<const uint8_t*>(dev) - tables_), src, bytes);
    };
    put(d_ccounts_, counts, sizeof(counts));
    put(d_cmaps_, smaps.data(), smaps.size() * sizeof(MapJob));
    put(d_cbmaps_, bmaps.data(), bmaps.size() * sizeof(MapJob));
    put(d_csjobs_, sj.data(), sj.size() * sizeof(ScoreJob));
    const size_t first = reinterpret_cast<uint8_t*>(d_ccounts_) - tables_;
    const size_t last = (reinterpret_cast<uint8_t*>(d_csjobs_) - tables_) + sj.size() * sizeof(ScoreJob);
    VSHIP_HIP_CHECK(hipMemcpyAsync(tables_ + first, h_tables_ + first, last - first, hipMemcpyHostToDevice, stream_));
    const int bgrid = std::min((std::max(1, tuning_.blur_waves) * cus_ + 7) / 8 * 8, blur_grid_for(btotal));
    if (btotal > 0) {
        if (fp16_)
            VSHIP_HIP_CHECK(launch_blur<__half>(tuning_.blur_config, d_cbmaps_, d_ccounts_, bgrid, taps_, partials_,
                                                stream_));
        else
            VSHIP_HIP_CHECK(launch_blur<float>(tuning_.blur_config, d_cbmaps_, d_ccounts_, bgrid, taps_, partials_,
                                               stream_));
    }
    VSHIP_HIP_CHECK(launch_score(d_csjobs_, d_ccounts_, std::min(count, kScoreGrid), d_cmaps_, partials_, d_features_,
                                 d_scores_, stream_));
    VSHIP_HIP_CHECK(hipMemcpyAsync(h_scores_, d_scores_, (size_t)npairs_ * sizeof(double), hipMemcpyDeviceToHost,
                                   stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    for (int c = 0; c < count; c++) scores[c] = h_scores_[pairs[c]];
}

void Ssimu2Batch::features(double* out)
{
    if (npairs_ == 0) VSHIP_THROW(Vship_BadPointer, "features need one successful compute call first");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipMemcpy(out, d_features_, (size_t)npairs_ * 108 * sizeof(double), hipMemcpyDeviceToHost));
}

void Ssimu2Batch::set_tuning(const BatchTuning& t)
{
    if (t.blur_config < 0 || t.blur_config >= blur_tile_configs() || t.blur_waves < 1 || t.blur_waves > 64 ||
        t.front_waves < 1 || t.front_waves > 64 || (t.generic_front != 0 && t.generic_front != 1))
        VSHIP_THROW(Vship_BadPointer, "tuning parameters out of range");
    tuning_ = t;
    graph_valid_ = false;
}

void Ssimu2Batch::autotune(const BatchInput& in, int reps, BatchTuning* chosen, double* best_ms)
{
    npairs_ = 0;  // the replays leave features of no completed compute
    call_mask_ = kAllScales;
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    reps = std::max(1, reps);
    hipEvent_t e0, e1;
    VSHIP_HIP_CHECK(hipEventCreate(&e0));
    VSHIP_HIP_CHECK(hipEventCreate(&e1));
    auto time_config = [&](const BatchTuning& t) {
        tuning_ = t;
        build_tables(in);
        ensure_capacity();
        run_resamples();          // the front-end's input (not timed: the same for every tuning)
        launch_kernels(stream_);  // warm-up
        VSHIP_HIP_CHECK(hipEventRecord(e0, stream_));
        for (int r = 0; r < reps; r++) launch_kernels(stream_);
        VSHIP_HIP_CHECK(hipEventRecord(e1, stream_));
        VSHIP_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        VSHIP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
    };
    BatchTuning best = tuning_;
    double best_t = 1e30;
    // blur tile shape and persistent blur grid, front-end grid fixed
    for (int cfg = 0; cfg < blur_tile_configs(); cfg++)
        for (int bw : {2, 4, 16, 32, 64}) {
            BatchTuning t = best_t < 1e30 ? best : tuning_;
            t.blur_config = cfg;
            t.blur_waves = bw;
            const double ms = time_config(t);
            if (ms < best_t) { best_t = ms; best = t; }
        }
    for (int fw : {2, 4, 8, 16, 32, 64}) {
        BatchTuning t = best;
        t.front_waves = fw;
        const double ms = time_config(t);
        if (ms < best_t) { best_t = ms; best = t; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tuning_ = best;
    graph_valid_ = false;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
}

void Ssimu2Batch::profile(int iters, double* ms_front, double* ms_blur, double* ms_score, double* ms_graph)
{
    if (npairs_ == 0) VSHIP_THROW(Vship_BadPointer, "profile needs one successful compute call first");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    iters = std::max(1, iters);
    hipEvent_t ev[4];
    for (auto& e : ev) VSHIP_HIP_CHECK(hipEventCreate(&e));
    const int bgrid = (std::max(1, tuning_.blur_waves) * cus_ + 7) / 8 * 8;
    double acc[3] = {0, 0, 0};
    for (int it = 0; it < iters; it++) {
        VSHIP_HIP_CHECK(hipEventRecord(ev[0], stream_));
        launch_front_groups(stream_);
        VSHIP_HIP_CHECK(hipEventRecord(ev[1], stream_));
        if (fp16_) VSHIP_HIP_CHECK(launch_blur<__half>(tuning_.blur_config, d_bmaps_, d_counts_, bgrid, taps_, partials_, stream_));
        else VSHIP_HIP_CHECK(launch_blur<float>(tuning_.blur_config, d_bmaps_, d_counts_, bgrid, taps_, partials_, stream_));
        VSHIP_HIP_CHECK(hipEventRecord(ev[2], stream_));
        VSHIP_HIP_CHECK(launch_score(d_sjobs_, d_counts_, kScoreGrid, d_maps_, partials_, d_features_, d_scores_, stream_));
        VSHIP_HIP_CHECK(hipEventRecord(ev[3], stream_));
        VSHIP_HIP_CHECK(hipEventSynchronize(ev[3]));
        for (int k = 0; k < 3; k++) {
            float ms = 0.0f;
            VSHIP_HIP_CHECK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    // whole graph replays
    if (!graph_valid_) capture();
    VSHIP_HIP_CHECK(hipEventRecord(ev[0], stream_));
    for (int it = 0; it < iters; it++) VSHIP_HIP_CHECK(hipGraphLaunch(exec_, stream_));
    VSHIP_HIP_CHECK(hipEventRecord(ev[1], stream_));
    VSHIP_HIP_CHECK(hipEventSynchronize(ev[1]));
    float gms = 0.0f;
    VSHIP_HIP_CHECK(hipEventElapsedTime(&gms, ev[0], ev[1]));
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_front = acc[0] / iters;
    *ms_blur = acc[1] / iters;
    *ms_score = acc[2] / iters;
    *ms_graph = gms / iters;
}

void Ssimu2Batch::work(double* bytes, double* flops) const
{
    *bytes = work_bytes_;
    *flops = work_flops_;
}

}  // namespace vship
