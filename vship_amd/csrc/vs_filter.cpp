// vs_filter.cpp - VapourSynth plugin entry: vship.SSIMULACRA2 and
// vship.GpuInfo on the gfx950 engine.
//
// Contract followed (reference src/VshipLib.cpp:9-68,
// src/ssimu2/vapoursynth.hpp:8-195, src/util/torgbs.hpp:8-36):
// * plugin id "com.lumen.vship", namespace "vship", version 4.1;
//   SSIMULACRA2(reference:vnode, distorted:vnode, numStream:int:opt,
//   gpu_id:int:opt) -> clip:vnode; GpuInfo(gpu_id:int:opt) ->
//   gpu_human_data:data.
// * Inputs that are not 32-bit float RGB are converted by the host's
//   resize.Bicubic to RGBS (matrix_in BT.709 above 650 rows, else BT.601;
//   BT.709 transfer and primaries; limited -> full range).
// * Both clips must have the same VSVideoInfo (else DifferingInputType) and be
//   RGBS afterwards (else NonRGBSInput). Each clip is computed as float RGB,
//   full range, BT.709 transfer and primaries.
// * The output is a copy of the (converted) distorted frame with the float
//   property _SSIMULACRA2; fmParallel, strict-spatial requests of both clips.
// * numStream handlers (default and upper bound: the core's thread count, at
//   least 1) are shared by the frame requests through a blocking pool.
#include <algorithm>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/VshipAPI.h"
#include "engine.hpp"
#include "errors.hpp"
#include "vs4_abi.hpp"

namespace {

using namespace vs4;

const char kResizePluginId[] = "com.vapoursynth.resize";

std::string error_message(Vship_Exception code, const char* file, int line, const std::string& detail = "")
{
    return vship::Error(code, file, line, detail).message();
}

// Blocking pool of handler indices.
class HandlerPool {
public:
    int acquire()
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return !idle_.empty(); });
        const int k = idle_.back();
        idle_.pop_back();
        return k;
    }
    void release(int k)
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            idle_.push_back(k);
        }
        cv_.notify_one();
    }
    void add(int k) { idle_.push_back(k); }

private:
    std::mutex m_;
    std::condition_variable cv_;
    std::vector<int> idle_;
};

struct FilterState {
    VSNode* reference = nullptr;
    VSNode* distorted = nullptr;
    std::vector<std::unique_ptr<vship::Ssimu2Handler>> handlers;
    HandlerPool pool;
};

bool is_rgbs(const VSVideoInfo* vi)
{
    return vi->format.colorFamily == cfRGB && vi->format.sampleType == stFloat && vi->format.bitsPerSample == 32;
}

bool same_video_info(const VSVideoInfo* a, const VSVideoInfo* b)
{
    return a->width == b->width && a->height == b->height && a->format.colorFamily == b->format.colorFamily &&
           a->format.sampleType == b->format.sampleType && a->format.bitsPerSample == b->format.bitsPerSample &&
           a->format.subSamplingW == b->format.subSamplingW && a->format.subSamplingH == b->format.subSamplingH;
}

// Consumes `src`; returns an RGBS node (src itself when already RGBS), or
// nullptr when the host's resize plugin is missing or fails.
VSNode* to_rgbs(VSNode* src, VSCore* core, const VSAPI* api)
{
    const VSVideoInfo* vi = api->getVideoInfo(src);
    if (is_rgbs(vi)) return src;
    VSMap* args = api->createMap();
    api->mapSetInt(args, "matrix_in", vi->height > 650 ? 1 : 6, maReplace);  // BT.709 / BT.601
    api->mapConsumeNode(args, "clip", src, maReplace);
    api->mapSetInt(args, "transfer_in", 1, maReplace);
    api->mapSetInt(args, "primaries_in", 1, maReplace);
    api->mapSetInt(args, "range_in", 0, maReplace);
    api->mapSetInt(args, "matrix", 0, maReplace);
    api->mapSetInt(args, "transfer", 1, maReplace);
    api->mapSetInt(args, "primaries", 1, maReplace);
    api->mapSetInt(args, "range", 1, maReplace);
    api->mapSetInt(args, "format", pfRGBS, maReplace);
    VSPlugin* resize = api->getPluginByID(kResizePluginId, core);
    VSNode* out = nullptr;
    if (resize) {
        VSMap* ret = api->invoke(resize, "Bicubic", args);
        if (ret) {
            int err = 0;
            out = api->mapGetError(ret) ? nullptr : api->mapGetNode(ret, "clip", 0, &err);
            api->freeMap(ret);
        }
    }
    api->freeMap(args);
    return out;
}

const VSFrame* ssimu2_get_frame(int n, int reason, void* instance, void** frame_data, VSFrameContext* ctx,
                                VSCore* core, const VSAPI* api)
{
    (void)frame_data;
    FilterState* d = static_cast<FilterState*>(instance);
    if (reason == arInitial) {
        api->requestFrameFilter(n, d->reference, ctx);
        api->requestFrameFilter(n, d->distorted, ctx);
        return nullptr;
    }
    if (reason != arAllFramesReady) return nullptr;
    const VSFrame* f1 = api->getFrameFilter(n, d->reference, ctx);
    const VSFrame* f2 = api->getFrameFilter(n, d->distorted, ctx);
    const uint8_t* p1[3];
    const uint8_t* p2[3];
    int64_t l1[3], l2[3];
    for (int p = 0; p < 3; p++) {
        p1[p] = api->getReadPtr(f1, p);
        p2[p] = api->getReadPtr(f2, p);
        l1[p] = (int64_t)api->getStride(f1, p);
        l2[p] = (int64_t)api->getStride(f2, p);
    }
    const int k = d->pool.acquire();
    double score = 0.0;
    std::string err;
    try {
        score = d->handlers[k]->compute_host(p1, p2, l1, l2);
    } catch (const vship::Error& e) {
        err = e.message();
    } catch (...) {
        err = error_message(Vship_BadErrorType, __FILE__, __LINE__);
    }
    d->pool.release(k);
    if (!err.empty()) {
        api->setFilterError(err.c_str(), ctx);
        api->freeFrame(f1);
        api->freeFrame(f2);
        return nullptr;
    }
    VSFrame* out = api->copyFrame(f2, core);
    api->mapSetFloat(api->getFramePropertiesRW(out), "_SSIMULACRA2", score, maReplace);
    api->freeFrame(f1);
    api->freeFrame(f2);
    return out;
}

void ssimu2_free(void* instance, VSCore* core, const VSAPI* api)
{
    (void)core;
    FilterState* d = static_cast<FilterState*>(instance);
    api->freeNode(d->reference);
    api->freeNode(d->distorted);
    delete d;
}

void ssimu2_create(const VSMap* in, VSMap* out, void* user, VSCore* core, const VSAPI* api)
{
    (void)user;
    std::unique_ptr<FilterState> d(new FilterState);
    int err = 0;
    VSNode* ref_in = api->mapGetNode(in, "reference", 0, &err);
    VSNode* dis_in = api->mapGetNode(in, "distorted", 0, &err);
    d->reference = ref_in ? to_rgbs(ref_in, core, api) : nullptr;
    d->distorted = dis_in ? to_rgbs(dis_in, core, api) : nullptr;
    auto fail = [&](const std::string& msg) {
        api->mapSetError(out, msg.c_str());
        if (d->reference) api->freeNode(d->reference);
        if (d->distorted) api->freeNode(d->distorted);
        d->reference = d->distorted = nullptr;
    };
    if (!d->reference || !d->distorted) {
        fail(error_message(Vship_NonRGBSInput, __FILE__, __LINE__, "conversion to RGBS failed"));
        return;
    }
    const VSVideoInfo* vr = api->getVideoInfo(d->reference);
    const VSVideoInfo* vd = api->getVideoInfo(d->distorted);
    if (!same_video_info(vr, vd)) {
        fail(error_message(Vship_DifferingInputType, __FILE__, __LINE__));
        return;
    }
    if (!is_rgbs(vr)) {
        fail(error_message(Vship_NonRGBSInput, __FILE__, __LINE__));
        return;
    }
    int gpu = (int)api->mapGetInt(in, "gpu_id", 0, &err);
    if (err != peSuccess) gpu = 0;
    const Vship_Exception chk = Vship_GPUFullCheck(gpu);
    if (chk != Vship_NoError) {
        fail(error_message(chk, __FILE__, __LINE__));
        return;
    }
    VSCoreInfo info{};
    api->getCoreInfo(core, &info);
    int streams = (int)api->mapGetInt(in, "numStream", 0, &err);
    if (err != peSuccess) streams = info.numThreads;
    streams = std::max(1, std::min(streams, info.numThreads));

    Vship_Colorspace_t cs{};
    cs.width = vr->width;
    cs.height = vr->height;
    cs.target_width = -1;
    cs.target_height = -1;
    cs.sample = Vship_SampleFLOAT;
    cs.range = Vship_RangeFull;
    cs.subsampling.subw = 0;
    cs.subsampling.subh = 0;
    cs.chromaLocation = Vship_ChromaLoc_Left;
    cs.colorFamily = Vship_ColorRGB;
    cs.YUVMatrix = Vship_MATRIX_RGB;
    cs.transferFunction = Vship_TRC_BT709;
    cs.primaries = Vship_PRIMARIES_BT709;
    try {
        for (int i = 0; i < streams; i++) {
            d->handlers.emplace_back(new vship::Ssimu2Handler(cs, cs, gpu));
            d->pool.add(i);
        }
    } catch (const vship::Error& e) {
        fail(e.message());
        return;
    } catch (...) {
        fail(error_message(Vship_BadErrorType, __FILE__, __LINE__));
        return;
    }
    VSFilterDependency deps[2] = {{d->reference, rpStrictSpatial}, {d->distorted, rpStrictSpatial}};
    VSVideoInfo vi = *vr;
    FilterState* state = d.release();
    api->createVideoFilter(out, "vship", &vi, ssimu2_get_frame, ssimu2_free, fmParallel, deps, 2, state, core);
}

void gpu_info(const VSMap* in, VSMap* out, void* user, VSCore* core, const VSAPI* api)
{
    (void)user;
    (void)core;
    int count = 0;
    if (Vship_GetDeviceCount(&count) != Vship_NoError || count <= 0) {
        api->mapSetError(out, error_message(Vship_NoDeviceDetected, __FILE__, __LINE__).c_str());
        return;
    }
    int err = 0;
    int gpu = (int)api->mapGetInt(in, "gpu_id", 0, &err);
    const bool chosen = err == peSuccess;
    if (!chosen) gpu = 0;
    if (gpu < 0 || gpu >= count) {
        api->mapSetError(out, error_message(Vship_BadDeviceArgument, __FILE__, __LINE__).c_str());
        return;
    }
    std::ostringstream ss;
    hipDeviceProp_t prop;
    if (!chosen) {
        for (int i = 0; i < count; i++) {
            if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
            ss << "GPU " << i << ": " << prop.name << "\n";
        }
    } else if (hipGetDeviceProperties(&prop, gpu) == hipSuccess) {
        ss << "Name: " << prop.name << "\n"
           << "MultiProcessorCount: " << prop.multiProcessorCount << "\n"
           << "MaxSharedMemoryPerBlock: " << prop.sharedMemPerBlock << " bytes\n"
           << "WarpSize: " << prop.warpSize << "\n"
           << "VRAMCapacity: " << (float)prop.totalGlobalMem / 1e9f << " GB\n"
           << "MemoryBusWidth: " << prop.memoryBusWidth << " bits\n"
           << "Integrated: " << prop.integrated << "\n"
           << "PassKernelCheck : " << (Vship_GPUFullCheck(gpu) == Vship_NoError ? 1 : 0) << "\n";
    }
    const std::string s = ss.str();
    api->mapSetData(out, "gpu_human_data", s.data(), (int)s.size(), dtUtf8, maReplace);
}

}  // namespace

extern "C" __attribute__((visibility("default"))) void VapourSynthPluginInit2(vs4::VSPlugin* plugin,
                                                                               const vs4::VSPLUGINAPI* vspapi)
{
    vspapi->configPlugin("com.lumen.vship", "vship", "VapourSynth SSIMULACRA2 on GPU", vs4::make_version(4, 1),
                         vs4::kApiVersion, 0, plugin);
    vspapi->registerFunction("SSIMULACRA2", "reference:vnode;distorted:vnode;numStream:int:opt;gpu_id:int:opt;",
                             "clip:vnode;", ssimu2_create, nullptr, plugin);
    vspapi->registerFunction("GpuInfo", "gpu_id:int:opt;", "gpu_human_data:data;", gpu_info, nullptr, plugin);
}
