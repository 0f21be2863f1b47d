// noalloc.hip - one SSIMULACRA2 pair with no device allocation: the tables,
// both images' XYB planes and the blur items' partial sums live in a
// file-scope device array (`scratch`) that every GPU the library runs on holds
// for the life of the process. The kernels are the plain FP32 path's
// (front-end tile kernel, 64x128-strip matrix-core blur) and the score is
// completed on the calling thread from the same partial sums in the same
// order, so the result equals VshipAMD_ComputeSSIMU2Device's bit for bit.
//
// All calls on one GPU share that GPU's array: calls that overlap on one GPU
// race (the caller serialises them). Nothing in the array has to start at
// zero: every table, plane and partial sum a call reads is written by that
// call first (row padding and out-of-image columns of the planes are read
// but masked by the blur, as in the handler's own hipMalloc'ed buffers).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "engine.hpp"
#include "errors.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"

namespace vship {

namespace {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU

// The device address of this GPU's `scratch` (the symbol's host-side name is
// not a device address). Looked up once per device, with that device current.
uint8_t* scratch_on(int device)
{
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            VSHIP_THROW(Vship_HIPError, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                                            hipGetErrorString(e));
        g_scratch[device] = p;
    }
    return static_cast<uint8_t*>(g_scratch[device]);
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Device-side tables of one pair, uploaded with one copy.
struct Tables {
    FrontColor colors[2];
    PairGeom geom;
    MapJob maps[kMaxJobsPerPair];
    int counts[kCountSlots];
};

}  // namespace

double compute_ssimu2_noalloc(const Vship_Colorspace_t& src, const Vship_Colorspace_t& dis,
                              const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3],
                              const int64_t ls2[3], int device)
{
    const ImageSpec s1 = make_image_spec(src);
    const ImageSpec s2 = make_image_spec(dis);
    if (s1.resample || s2.resample)
        VSHIP_THROW(Vship_NonRGBSInput, "crop and resize need the handler's buffers: use Vship_SSIMU2Init2");
    if (s1.width != s2.width || s1.height != s2.height)
        VSHIP_THROW(Vship_DifferingInputType, "reference and distorted sizes differ");
    const ImageSpec* specs[2] = {&s1, &s2};
    const uint8_t* const* planes[2] = {p1, p2};
    const int64_t* strides[2] = {ls1, ls2};
    for (int k = 0; k < 2; k++)
        for (int p = 0; p < 3; p++) {
            if (planes[k][p] == nullptr) VSHIP_THROW(Vship_BadPointer, "null plane pointer");
            if (strides[k][p] < specs[k]->plane_w[p] * specs[k]->bytes)
                VSHIP_THROW(Vship_BadPointer, "lineSize smaller than one row of samples");
        }
    const PairLayout L = make_pair_layout((int)s1.width, (int)s1.height, kDefaultBlurConfig);

    // scratch layout: tables | image 0 planes | image 1 planes | partial sums
    const size_t o_tables = 0;
    const size_t o_xyb0 = align256(sizeof(Tables));
    const size_t image_bytes = align256((size_t)L.geom.image_elems * sizeof(float));
    const size_t o_xyb1 = o_xyb0 + image_bytes;
    const size_t o_part = o_xyb1 + image_bytes;
    const size_t need = o_part + (size_t)L.blur_tiles * kPartialStride * sizeof(float);
    if (need > kNoAllocScratchBytes || L.blur_tiles > kNoAllocMaxItems)
        VSHIP_THROW(Vship_OutOfVRAM, "frame of " + std::to_string(s1.width) + "x" + std::to_string(s1.height) +
                                         " needs " + std::to_string(need) + " bytes and " +
                                         std::to_string(L.blur_tiles) +
                                         " blur items; the no-allocation entry point holds at most " +
                                         std::to_string(kNoAllocScratchBytes) + " bytes and " +
                                         std::to_string(kNoAllocMaxItems) +
                                         " items (1920x1080 and 2048x1080 fit): use Vship_SSIMU2Init2");

    int ndev = 0;
    VSHIP_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= kMaxDevices)
        VSHIP_THROW(Vship_BadDeviceArgument, "In VshipAMD_ComputeSSIMU2DeviceNoAlloc");
    VSHIP_HIP_CHECK(hipSetDevice(device));
    uint8_t* base = scratch_on(device);

    Tables t;
    std::memset(&t, 0, sizeof(t));
    t.colors[0] = s1.fc;
    t.colors[1] = s2.fc;
    t.geom = L.geom;
    float* xyb[2] = {reinterpret_cast<float*>(base + o_xyb0), reinterpret_cast<float*>(base + o_xyb1)};
    for (int m = 0; m < L.maps.count; m++) {
        MapJob job = L.maps.items[m];
        const int64_t off = reinterpret_cast<int64_t>(job.a);
        job.a = xyb[0] + off;
        job.b = xyb[1] + off;
        t.maps[m] = job;
    }
    t.counts[kCountFrontTiles] = L.front_tiles_per_image;
    t.counts[kCountFrontJobs] = 1;
    t.counts[kCountBlurTiles] = L.blur_tiles;
    t.counts[kCountBlurJobs] = L.maps.count;
    t.counts[kCountPairs] = 1;

    FrontJob job;
    std::memset(&job, 0, sizeof(job));
    for (int k = 0; k < 2; k++) {
        for (int p = 0; p < 3; p++) {
            job.plane[k][p] = planes[k][p];
            job.stride[k][p] = strides[k][p];
        }
        job.out[k] = xyb[k];
        job.color[k] = k;
    }
    job.width = (int32_t)s1.width;
    job.height = (int32_t)s1.height;
    set_front_tiles(job, L.front_tiles_x);
    int spec = front_spec(s1.fc, s2.fc);
    if (spec != kFrontGeneric && !front_offsets_fit(job, s1.fc, s2.fc)) spec = kFrontGeneric;

    Tables* dt = reinterpret_cast<Tables*>(base + o_tables);
    float* dpart = reinterpret_cast<float*>(base + o_part);
    float partials[kNoAllocMaxItems * kPartialStride];
    const hipStream_t stream = nullptr;  // the device's null stream
    VSHIP_HIP_CHECK(hipMemcpyAsync(dt, &t, sizeof(t), hipMemcpyHostToDevice, stream));
    VSHIP_HIP_CHECK(
        launch_frontend_single<float>(job, L.front_tiles_per_image, dt->colors, &dt->geom, spec, stream));
    VSHIP_HIP_CHECK(launch_blur<float>(kDefaultBlurConfig, dt->maps, dt->counts, blur_grid_for(L.blur_tiles),
                                       make_blur_taps(), dpart, stream));
    VSHIP_HIP_CHECK(hipMemcpyAsync(partials, dpart, (size_t)L.blur_tiles * kPartialStride * sizeof(float),
                                   hipMemcpyDeviceToHost, stream));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream));
    double features[108];
    return score_from_partials(L.maps.data(), L.maps.count, partials, features);
}

}  // namespace vship
