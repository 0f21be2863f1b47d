// api.cpp - the C ABI (include/VshipAPI.h, include/vship_amd.h).
//
// Behaviour follows the reference's src/VshipLib.cpp:70-563: every entry point
// returns a Vship_Exception, records the detailed message of a failure
// globally and per handler, and never lets a C++ exception escape. Handler ids
// index tables (handle_table.hpp); unlike the reference (VshipLib.cpp:242-256)
// a freed id is marked empty, so using it again returns Vship_BadHandler
// instead of touching freed memory, a call holds a reference to its handler so
// that a concurrent Free cannot destroy it mid-call, and calls on one handler
// are serialised.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/VshipAPI.h"
#include "../../include/vship_amd.h"
#include "batch.hpp"
#include "engine.hpp"
#include "errors.hpp"
#include "handle_table.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "noalloc.hpp"

namespace vship {

const char* error_text(Vship_Exception code)
{
    switch (code) {
    case Vship_NoError: return "NoError: the call succeeded.";
    case Vship_OutOfVRAM: return "OutOfVRAM: a GPU memory allocation failed. (Advice) use fewer concurrent handlers.";
    case Vship_OutOfRAM: return "OutOfRAM: a host memory allocation failed. (Advice) check the process memory usage.";
    case Vship_HIPError: return "InternalError: a HIP runtime call failed inside the library.";
    case Vship_BadDisplayModel: return "BadDisplayModel: unknown display model.";
    case Vship_DifferingInputType: return "DifferingInputType: the two inputs do not have the same final width and height.";
    case Vship_NonRGBSInput: return "NonRGBSInput: the input colourspace cannot be converted to linear RGB by this library.";
    case Vship_BadPath: return "BadPath: a file could not be opened.";
    case Vship_BadJson: return "BadJson: a JSON document could not be parsed.";
    case Vship_DeviceCountError: return "DeviceCountError: the number of GPUs could not be queried. (Advice) check that the ROCm driver is loaded and the user is in the video and render groups.";
    case Vship_NoDeviceDetected: return "NoDeviceDetected: no GPU was found. (Advice) check that the ROCm driver is loaded and the user is in the video and render groups.";
    case Vship_BadDeviceArgument: return "BadDeviceArgument: gpu_id is negative or not below the number of GPUs.";
    case Vship_BadDeviceCode: return "BadDeviceCode: a probe kernel did not run. (Advice) this library is built for gfx950 (MI355X) only.";
    case Vship_BadHandler: return "BadHandler: the handler was never initialised or was already freed.";
    case Vship_BadPointer: return "BadPointer: a pointer or size argument is not valid.";
    case Vship_BadErrorType: return "BadErrorType: unknown error.";
    }
    return "BadErrorType: unknown error.";
}

std::string Error::message() const
{
    std::ostringstream ss;
    ss << "VshipException" << '\n' << error_text(type_) << '\n' << " - At line " << line_ << " of " << file_ << '\n';
    if (!detail_.empty()) ss << "Detail: " << detail_ << '\n';
    return ss.str();
}

namespace {

// One object of a table plus the lock that serialises calls on it: a handler
// is not reentrant (the reference leaves concurrent use of one handler
// undefined, src/VshipLib.cpp:258); here the second call waits.
template <typename T>
struct Guarded {
    explicit Guarded(std::unique_ptr<T> p) : impl(std::move(p)) {}
    std::unique_ptr<T> impl;  // null when its init failed
    std::mutex mu;
};

std::mutex g_lock;  // g_last_error, g_device
HandleTable<Guarded<Ssimu2Handler>> g_handlers;
HandleTable<Guarded<Ssimu2Batch>> g_batches;
HandleTable<Guarded<Surrogate>> g_mlps;
std::string g_last_error = Error(Vship_NoError, __FILE__, __LINE__).message();
int g_device = 0;

void set_last(const std::string& msg)
{
    std::lock_guard<std::mutex> g(g_lock);
    g_last_error = msg;
}

int copy_message(const std::string& s, char* out, int len)
{
    if (len <= 0 || out == nullptr) return (int)s.size() + 1;
    const int n = std::min(len - 1, (int)s.size());
    std::memcpy(out, s.data(), (size_t)n);
    out[n] = '\0';
    return (int)s.size() + 1;
}

int device_count()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) VSHIP_THROW(Vship_DeviceCountError, "");
    if (n == 0) VSHIP_THROW(Vship_NoDeviceDetected, "");
    return n;
}

__global__ void probe_kernel(int* p) { p[0] = 4320984; }

void record_handler_error(unsigned id, const std::string& msg)
{
    {
        std::lock_guard<std::mutex> g(g_lock);
        g_last_error = msg;
    }
    (void)g_handlers.set_error(id, msg);
}

template <typename F>
Vship_Exception guarded(F&& f)
{
    try {
        f();
    } catch (const Error& e) {
        set_last(e.message());
        return e.type();
    } catch (const std::bad_alloc&) {
        set_last(Error(Vship_OutOfRAM, __FILE__, __LINE__).message());
        return Vship_OutOfRAM;
    } catch (...) {
        set_last(Error(Vship_BadErrorType, __FILE__, __LINE__).message());
        return Vship_BadErrorType;
    }
    return Vship_NoError;
}

template <typename F>
Vship_Exception guarded_handler(unsigned id, F&& f)
{
    try {
        f();
    } catch (const Error& e) {
        record_handler_error(id, e.message());
        return e.type();
    } catch (const std::bad_alloc&) {
        record_handler_error(id, Error(Vship_OutOfRAM, __FILE__, __LINE__).message());
        return Vship_OutOfRAM;
    } catch (...) {
        record_handler_error(id, Error(Vship_BadErrorType, __FILE__, __LINE__).message());
        return Vship_BadErrorType;
    }
    return Vship_NoError;
}

// Runs f(object) holding a reference to a live table entry and its call
// lock, so a concurrent Free cannot destroy the object under the call.
template <typename T, typename F>
void with_object(const HandleTable<Guarded<T>>& table, unsigned id, const char* where, F&& f)
{
    std::shared_ptr<Guarded<T>> h;
    if (!table.lookup(id, &h)) throw Error(Vship_BadHandler, __FILE__, __LINE__, std::string("invalid id in ") + where);
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->impl) throw Error(Vship_BadHandler, __FILE__, __LINE__, std::string("initialisation had failed: ") + where);
    f(*h->impl);
}

// The same for SSIMU2 handlers, recording a failure as the handler's detailed
// error too (reference src/VshipLib.cpp:262-286).
template <typename F>
Vship_Exception handler_call(unsigned id, const char* where, F&& f)
{
    std::shared_ptr<Guarded<Ssimu2Handler>> h;
    if (!g_handlers.lookup(id, &h)) {
        set_last(Error(Vship_BadHandler, __FILE__, __LINE__, std::string("invalid handler id in ") + where).message());
        return Vship_BadHandler;
    }
    return guarded_handler(id, [&] {
        std::lock_guard<std::mutex> g(h->mu);
        if (!h->impl) VSHIP_THROW(Vship_BadHandler, "handler initialisation had failed");
        f(*h->impl);
    });
}

BatchInput make_batch_input(int npairs, const Vship_Colorspace_t* src, const Vship_Colorspace_t* dis,
                            const uint8_t* const* p1, const uint8_t* const* p2, const int64_t* l1, const int64_t* l2)
{
    if (!src || !dis || !p1 || !p2 || !l1 || !l2) VSHIP_THROW(Vship_BadPointer, "null argument");
    BatchInput in;
    in.npairs = npairs;
    in.src = src;
    in.dis = dis;
    in.planes1 = p1;
    in.planes2 = p2;
    in.ls1 = l1;
    in.ls2 = l2;
    return in;
}

Vship_Exception not_in_library(const char* what)
{
    set_last(Error(Vship_BadErrorType, __FILE__, __LINE__,
                   std::string(what) + " is not part of this library (SSIMULACRA2 only)").message());
    return Vship_BadErrorType;
}

}  // namespace
}  // namespace vship

using namespace vship;

extern "C" {

Vship_Version Vship_GetVersion(void)
{
    Vship_Version v;
    v.major = 4;
    v.minor = 1;
    v.minorMinor = 0;
    v.backend = Vship_HIP;
    return v;
}

Vship_Exception Vship_GetDeviceCount(int* number)
{
    if (!number) return Vship_BadPointer;
    return guarded([&] { *number = device_count(); });
}

Vship_Exception Vship_SetDevice(int gpu_id)
{
    return guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "In Vship_SetDevice");
        std::lock_guard<std::mutex> g(g_lock);
        g_device = gpu_id;
    });
}

Vship_Exception Vship_GetDeviceInfo(Vship_DeviceInfo* info, int gpu_id)
{
    if (!info) return Vship_BadPointer;
    return guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "while using Vship_GetDeviceInfo");
        hipDeviceProp_t prop;
        VSHIP_HIP_CHECK(hipGetDeviceProperties(&prop, gpu_id));
        std::memcpy(info->name, prop.name, 256);
        info->name[255] = '\0';
        info->VRAMSize = prop.totalGlobalMem;
        info->integrated = prop.integrated;
        info->MultiProcessorCount = prop.multiProcessorCount;
        info->WarpSize = prop.warpSize;
    });
}

Vship_Exception Vship_GPUFullCheck(int gpu_id)
{
    return guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "");
        VSHIP_HIP_CHECK(hipSetDevice(gpu_id));
        int* d = nullptr;
        int h = 0;
        VSHIP_HIP_CHECK(hipMalloc((void**)&d, sizeof(int)));
        VSHIP_HIP_CHECK(hipMemset(d, 0, sizeof(int)));
        hipLaunchKernelGGL(probe_kernel, dim3(1), dim3(1), 0, 0, d);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        if (le != hipSuccess || ce != hipSuccess || h != 4320984) VSHIP_THROW(Vship_BadDeviceCode, "");
    });
}

int Vship_GetErrorMessage(Vship_Exception exception, char* out_message, int len)
{
    return copy_message(error_text(exception), out_message, len);
}

int Vship_GetDetailedLastError(char* out_message, int len)
{
    std::string s;
    {
        std::lock_guard<std::mutex> g(g_lock);
        s = g_last_error;
    }
    return copy_message(s, out_message, len);
}

Vship_Exception Vship_PinnedMalloc(void** ptr, uint64_t size)
{
    if (!ptr) return Vship_BadPointer;
    if (hipHostMalloc(ptr, size, hipHostMallocDefault) != hipSuccess) {
        set_last(Error(Vship_OutOfRAM, __FILE__, __LINE__, "In Vship_PinnedMalloc").message());
        return Vship_OutOfRAM;
    }
    return Vship_NoError;
}

Vship_Exception Vship_PinnedFree(void* ptr)
{
    if (hipHostFree(ptr) != hipSuccess) {
        set_last(Error(Vship_BadPointer, __FILE__, __LINE__, "hipHostFree failed on the pointer given to Vship_PinnedFree").message());
        return Vship_BadPointer;
    }
    return Vship_NoError;
}

Vship_Exception Vship_SSIMU2Init2(Vship_SSIMU2Handler* handler, Vship_Colorspace_t src_colorspace,
                                  Vship_Colorspace_t dis_colorspace, int gpu_id)
{
    if (!handler) return Vship_BadPointer;
    std::unique_ptr<Ssimu2Handler> impl;
    const Vship_Exception err = guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "In Vship_SSIMU2Init2");
        impl.reset(new Ssimu2Handler(src_colorspace, dis_colorspace, gpu_id));
    });
    // like the reference, an id is handed out even when init fails, so that
    // Vship_SSIMU2GetDetailedLastError can report it; compute on it fails.
    std::string msg;
    if (err == Vship_NoError) {
        msg = Error(Vship_NoError, __FILE__, __LINE__).message();
    } else {
        std::lock_guard<std::mutex> g(g_lock);
        msg = g_last_error;
    }
    handler->id = g_handlers.insert(std::make_shared<Guarded<Ssimu2Handler>>(std::move(impl)), std::move(msg));
    return err;
}

Vship_Exception Vship_SSIMU2Init(Vship_SSIMU2Handler* handler, Vship_Colorspace_t src_colorspace,
                                 Vship_Colorspace_t dis_colorspace)
{
    int dev;
    {
        std::lock_guard<std::mutex> g(g_lock);
        dev = g_device;
    }
    return Vship_SSIMU2Init2(handler, src_colorspace, dis_colorspace, dev);
}

Vship_Exception Vship_SSIMU2Free(Vship_SSIMU2Handler handler)
{
    std::shared_ptr<Guarded<Ssimu2Handler>> h;
    if (!g_handlers.remove(handler.id, &h)) {
        set_last(Error(Vship_BadHandler, __FILE__, __LINE__,
                       "Handler in Vship_SSIMU2Free is not valid: never allocated or already freed").message());
        return Vship_BadHandler;
    }
    // a call still running on another thread holds its own reference: the
    // handler is destroyed when that call returns
    return guarded([&] { h.reset(); });
}

Vship_Exception Vship_ComputeSSIMU2(Vship_SSIMU2Handler handler, double* score, const uint8_t* srcp1[3],
                                    const uint8_t* srcp2[3], const int64_t lineSize[3], const int64_t lineSize2[3])
{
    return handler_call(handler.id, "Vship_ComputeSSIMU2", [&](Ssimu2Handler& h) {
        if (!score || !srcp1 || !srcp2 || !lineSize || !lineSize2) VSHIP_THROW(Vship_BadPointer, "null argument");
        *score = h.compute_host(srcp1, srcp2, lineSize, lineSize2);
    });
}

int Vship_SSIMU2GetDetailedLastError(Vship_SSIMU2Handler handler, char* out_message, int len)
{
    std::string s;
    if (!g_handlers.get_error(handler.id, &s)) {
        set_last(Error(Vship_BadHandler, __FILE__, __LINE__, "Handler in Vship_SSIMU2GetDetailedLastError is not valid")
                     .message());
        return 0;
    }
    return copy_message(s, out_message, len);
}

// ---- Butteraugli and CVVDP: declared for link compatibility only

Vship_Exception Vship_ButteraugliInit(Vship_ButteraugliHandler*, Vship_Colorspace_t, Vship_Colorspace_t, int, float)
{
    return not_in_library("Butteraugli");
}
Vship_Exception Vship_ButteraugliInit2(Vship_ButteraugliHandler*, Vship_Colorspace_t, Vship_Colorspace_t, int, float, int)
{
    return not_in_library("Butteraugli");
}
Vship_Exception Vship_ButteraugliFree(Vship_ButteraugliHandler) { return not_in_library("Butteraugli"); }
Vship_Exception Vship_ComputeButteraugli(Vship_ButteraugliHandler, Vship_ButteraugliScore*, const uint8_t*, int64_t,
                                         const uint8_t*[3], const uint8_t*[3], const int64_t[3], const int64_t[3])
{
    return not_in_library("Butteraugli");
}
int Vship_ButteraugliGetDetailedLastError(Vship_ButteraugliHandler, char*, int) { return 0; }

Vship_Exception Vship_CVVDPInit(Vship_CVVDPHandler*, Vship_Colorspace_t, Vship_Colorspace_t, float, bool, const char*)
{
    return not_in_library("CVVDP");
}
Vship_Exception Vship_CVVDPInit2(Vship_CVVDPHandler*, Vship_Colorspace_t, Vship_Colorspace_t, float, bool, const char*,
                                 const char*)
{
    return not_in_library("CVVDP");
}
Vship_Exception Vship_CVVDPInit3(Vship_CVVDPHandler*, Vship_Colorspace_t, Vship_Colorspace_t, float, bool, const char*,
                                 const char*, int)
{
    return not_in_library("CVVDP");
}
Vship_Exception Vship_CVVDPFree(Vship_CVVDPHandler) { return not_in_library("CVVDP"); }
Vship_Exception Vship_CVVDPGetDistmapResolution(Vship_CVVDPHandler, int64_t*, int64_t*) { return not_in_library("CVVDP"); }
Vship_Exception Vship_ResetCVVDP(Vship_CVVDPHandler) { return not_in_library("CVVDP"); }
Vship_Exception Vship_ResetScoreCVVDP(Vship_CVVDPHandler) { return not_in_library("CVVDP"); }
Vship_Exception Vship_LoadTemporalCVVDP(Vship_CVVDPHandler, const uint8_t*[3], const uint8_t*[3], const int64_t[3],
                                        const int64_t[3])
{
    return not_in_library("CVVDP");
}
Vship_Exception Vship_ComputeCVVDP(Vship_CVVDPHandler, double*, const uint8_t*, int64_t, const uint8_t*[3],
                                   const uint8_t*[3], const int64_t[3], const int64_t[3])
{
    return not_in_library("CVVDP");
}
int Vship_CVVDPGetDetailedLastError(Vship_CVVDPHandler, char*, int) { return 0; }

// ---- extensions (include/vship_amd.h)

Vship_Exception VshipAMD_ComputeSSIMU2Device(Vship_SSIMU2Handler handler, double* score, const uint8_t* dsrcp1[3],
                                             const uint8_t* dsrcp2[3], const int64_t lineSize[3],
                                             const int64_t lineSize2[3])
{
    return handler_call(handler.id, "VshipAMD_ComputeSSIMU2Device", [&](Ssimu2Handler& h) {
        if (!score || !dsrcp1 || !dsrcp2 || !lineSize || !lineSize2) VSHIP_THROW(Vship_BadPointer, "null argument");
        *score = h.compute_device(dsrcp1, dsrcp2, lineSize, lineSize2);
    });
}

Vship_Exception VshipAMD_ComputeSSIMU2DeviceNoAlloc(Vship_Colorspace_t src_colorspace,
                                                    Vship_Colorspace_t dis_colorspace, double* score,
                                                    const uint8_t* dsrcp1[3], const uint8_t* dsrcp2[3],
                                                    const int64_t lineSize[3], const int64_t lineSize2[3], int gpu_id)
{
    return guarded([&] {
        if (!score || !dsrcp1 || !dsrcp2 || !lineSize || !lineSize2) VSHIP_THROW(Vship_BadPointer, "null argument");
        *score = compute_ssimu2_noalloc(src_colorspace, dis_colorspace, dsrcp1, dsrcp2, lineSize, lineSize2, gpu_id);
    });
}

Vship_Exception VshipAMD_SSIMU2Features(Vship_SSIMU2Handler handler, double features[108])
{
    return handler_call(handler.id, "VshipAMD_SSIMU2Features", [&](Ssimu2Handler& h) {
        if (!features) VSHIP_THROW(Vship_BadPointer, "null argument");
        h.last_features(features);
    });
}

Vship_Exception VshipAMD_SSIMU2DebugPlane(Vship_SSIMU2Handler handler, int image, int plane, int scale, float* out,
                                          int64_t* width, int64_t* height)
{
    return handler_call(handler.id, "VshipAMD_SSIMU2DebugPlane", [&](Ssimu2Handler& h) {
        if (!width || !height) VSHIP_THROW(Vship_BadPointer, "null argument");
        h.debug_plane(image, plane, scale, out, width, height);
    });
}

Vship_Exception VshipAMD_SSIMU2Profile(Vship_SSIMU2Handler handler, int iters, double* ms_frontend, double* ms_blur,
                                       double* ms_score)
{
    return handler_call(handler.id, "VshipAMD_SSIMU2Profile", [&](Ssimu2Handler& h) {
        if (iters <= 0 || !ms_frontend || !ms_blur || !ms_score) VSHIP_THROW(Vship_BadPointer, "bad argument");
        h.profile(iters, ms_frontend, ms_blur, ms_score);
    });
}

Vship_Exception VshipAMD_SSIMU2Work(Vship_SSIMU2Handler handler, double* bytes, double* flops)
{
    return handler_call(handler.id, "VshipAMD_SSIMU2Work", [&](Ssimu2Handler& h) {
        if (!bytes || !flops) VSHIP_THROW(Vship_BadPointer, "null argument");
        h.work(bytes, flops);
    });
}

Vship_Exception VshipAMD_SSIMU2BatchInit(VshipAMD_SSIMU2Batch* batch, int max_pairs, int storage, int gpu_id)
{
    if (!batch) return Vship_BadPointer;
    return guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "In VshipAMD_SSIMU2BatchInit");
        if (storage != VshipAMD_StorageFP32 && storage != VshipAMD_StorageFP16)
            VSHIP_THROW(Vship_BadPointer, "storage must be VshipAMD_StorageFP32 or VshipAMD_StorageFP16");
        std::unique_ptr<Ssimu2Batch> b(new Ssimu2Batch(max_pairs, storage == VshipAMD_StorageFP16, gpu_id));
        batch->id = g_batches.insert(std::make_shared<Guarded<Ssimu2Batch>>(std::move(b)));
    });
}

Vship_Exception VshipAMD_SSIMU2BatchFree(VshipAMD_SSIMU2Batch batch)
{
    std::shared_ptr<Guarded<Ssimu2Batch>> b;
    if (!g_batches.remove(batch.id, &b)) {
        set_last(Error(Vship_BadHandler, __FILE__, __LINE__, "invalid batch id in VshipAMD_SSIMU2BatchFree").message());
        return Vship_BadHandler;
    }
    return guarded([&] { b.reset(); });
}

Vship_Exception VshipAMD_SSIMU2BatchCompute(VshipAMD_SSIMU2Batch batch, int npairs, const Vship_Colorspace_t* src_cs,
                                            const Vship_Colorspace_t* dis_cs, const uint8_t* const* dsrcp1,
                                            const uint8_t* const* dsrcp2, const int64_t* lineSize1,
                                            const int64_t* lineSize2, double* scores)
{
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchCompute", [&](Ssimu2Batch& b) {
            if (!scores) VSHIP_THROW(Vship_BadPointer, "null scores");
            b.compute(make_batch_input(npairs, src_cs, dis_cs, dsrcp1, dsrcp2, lineSize1, lineSize2), scores);
        });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchComputeScales(VshipAMD_SSIMU2Batch batch, int npairs,
                                                  const Vship_Colorspace_t* src_cs, const Vship_Colorspace_t* dis_cs,
                                                  const uint8_t* const* dsrcp1, const uint8_t* const* dsrcp2,
                                                  const int64_t* lineSize1, const int64_t* lineSize2, int scale_mask,
                                                  double* scores)
{
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchComputeScales", [&](Ssimu2Batch& b) {
            if (!scores) VSHIP_THROW(Vship_BadPointer, "null scores");
            b.compute(make_batch_input(npairs, src_cs, dis_cs, dsrcp1, dsrcp2, lineSize1, lineSize2), scores,
                      scale_mask);
        });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchCompleteScales(VshipAMD_SSIMU2Batch batch, int count, const int* pairs,
                                                   int scale_mask, double* scores)
{
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchCompleteScales",
                    [&](Ssimu2Batch& b) { b.complete(count, pairs, scale_mask, scores); });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchAutotune(VshipAMD_SSIMU2Batch batch, int npairs, const Vship_Colorspace_t* src_cs,
                                             const Vship_Colorspace_t* dis_cs, const uint8_t* const* dsrcp1,
                                             const uint8_t* const* dsrcp2, const int64_t* lineSize1,
                                             const int64_t* lineSize2, int reps, VshipAMD_BatchTuning* chosen,
                                             double* best_ms)
{
    return guarded([&] {
        BatchTuning t;
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchAutotune", [&](Ssimu2Batch& b) {
            b.autotune(make_batch_input(npairs, src_cs, dis_cs, dsrcp1, dsrcp2, lineSize1, lineSize2), reps, &t,
                       best_ms);
        });
        if (chosen) {
            chosen->blur_config = t.blur_config;
            chosen->blur_waves = t.blur_waves;
            chosen->front_waves = t.front_waves;
            chosen->generic_front = t.generic_front;
        }
    });
}

Vship_Exception VshipAMD_SSIMU2BatchSetTuning(VshipAMD_SSIMU2Batch batch, VshipAMD_BatchTuning tuning)
{
    return guarded([&] {
        BatchTuning t;
        t.blur_config = tuning.blur_config;
        t.blur_waves = tuning.blur_waves;
        t.front_waves = tuning.front_waves;
        t.generic_front = tuning.generic_front;
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchSetTuning", [&](Ssimu2Batch& b) { b.set_tuning(t); });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchGetTuning(VshipAMD_SSIMU2Batch batch, VshipAMD_BatchTuning* tuning)
{
    if (!tuning) return Vship_BadPointer;
    return guarded([&] {
        BatchTuning t;
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchGetTuning", [&](Ssimu2Batch& b) { t = b.tuning(); });
        tuning->blur_config = t.blur_config;
        tuning->blur_waves = t.blur_waves;
        tuning->front_waves = t.front_waves;
        tuning->generic_front = t.generic_front;
    });
}

Vship_Exception VshipAMD_SSIMU2BatchFeatures(VshipAMD_SSIMU2Batch batch, double* features)
{
    if (!features) return Vship_BadPointer;
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchFeatures", [&](Ssimu2Batch& b) { b.features(features); });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchProfile(VshipAMD_SSIMU2Batch batch, int iters, double* ms_frontend,
                                            double* ms_blur, double* ms_score, double* ms_graph)
{
    if (!ms_frontend || !ms_blur || !ms_score || !ms_graph) return Vship_BadPointer;
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchProfile",
                    [&](Ssimu2Batch& b) { b.profile(iters, ms_frontend, ms_blur, ms_score, ms_graph); });
    });
}

Vship_Exception VshipAMD_SSIMU2BatchWork(VshipAMD_SSIMU2Batch batch, double* bytes, double* flops)
{
    if (!bytes || !flops) return Vship_BadPointer;
    return guarded([&] {
        with_object(g_batches, batch.id, "VshipAMD_SSIMU2BatchWork", [&](Ssimu2Batch& b) { b.work(bytes, flops); });
    });
}

int VshipAMD_MLPParamCount(void) { return kMlpParams; }

Vship_Exception VshipAMD_MLPInit(VshipAMD_MLP* mlp, int max_batch, int gpu_id)
{
    if (!mlp) return Vship_BadPointer;
    return guarded([&] {
        const int n = device_count();
        if (gpu_id < 0 || gpu_id >= n) VSHIP_THROW(Vship_BadDeviceArgument, "In VshipAMD_MLPInit");
        std::unique_ptr<Surrogate> m(new Surrogate(max_batch, gpu_id));
        mlp->id = g_mlps.insert(std::make_shared<Guarded<Surrogate>>(std::move(m)));
    });
}

Vship_Exception VshipAMD_MLPFree(VshipAMD_MLP mlp)
{
    std::shared_ptr<Guarded<Surrogate>> m;
    if (!g_mlps.remove(mlp.id, &m)) {
        set_last(Error(Vship_BadHandler, __FILE__, __LINE__, "invalid surrogate id in VshipAMD_MLPFree").message());
        return Vship_BadHandler;
    }
    return guarded([&] { m.reset(); });
}

Vship_Exception VshipAMD_MLPSetParams(VshipAMD_MLP mlp, const float* params, int count)
{
    return guarded([&] {
        if (!params) VSHIP_THROW(Vship_BadPointer, "null parameter array");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.set_params(params, count); });
    });
}

Vship_Exception VshipAMD_MLPGetParams(VshipAMD_MLP mlp, float* params, int count)
{
    return guarded([&] {
        if (!params) VSHIP_THROW(Vship_BadPointer, "null parameter array");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.get_params(params, count); });
    });
}

Vship_Exception VshipAMD_MLPSetNorm(VshipAMD_MLP mlp, const float* xmean, const float* xstd, float ymean, float ystd)
{
    return guarded([&] {
        if (!xmean || !xstd) VSHIP_THROW(Vship_BadPointer, "null normalisation array");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.set_norm(xmean, xstd, ymean, ystd); });
    });
}

Vship_Exception VshipAMD_MLPPredict(VshipAMD_MLP mlp, const float* dX, int n, float* dmean, float* dsigma)
{
    return guarded([&] {
        if (n > 0 && (!dX || !dmean || !dsigma)) VSHIP_THROW(Vship_BadPointer, "null sample array");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.predict(dX, n, dmean, dsigma); });
    });
}

Vship_Exception VshipAMD_MLPGradients(VshipAMD_MLP mlp, const float* dX, const float* dY, int n, float* grad,
                                      double* loss)
{
    return guarded([&] {
        if (!dX || !dY || !grad) VSHIP_THROW(Vship_BadPointer, "null argument");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.gradients(dX, dY, n, grad, loss); });
    });
}

Vship_Exception VshipAMD_MLPTrain(VshipAMD_MLP mlp, const float* dX, const float* dY, int n, int epochs, int batch,
                                  float lr, float weight_decay, uint64_t seed, float* loss_per_epoch)
{
    return guarded([&] {
        if (!dX || !dY) VSHIP_THROW(Vship_BadPointer, "null sample array");
        with_object(g_mlps, mlp.id, "surrogate", [&](Surrogate& m) { m.train(dX, dY, n, epochs, batch, lr, weight_decay, seed, loss_per_epoch); });
    });
}

Vship_Exception VshipAMD_DeviceMalloc(void** ptr, uint64_t size, int gpu_id)
{
    if (!ptr) return Vship_BadPointer;
    return guarded([&] {
        VSHIP_HIP_CHECK(hipSetDevice(gpu_id));
        if (hipMalloc(ptr, size) != hipSuccess) VSHIP_THROW(Vship_OutOfVRAM, "In VshipAMD_DeviceMalloc");
    });
}

Vship_Exception VshipAMD_DeviceFree(void* ptr)
{
    return guarded([&] { VSHIP_HIP_CHECK(hipFree(ptr)); });
}

Vship_Exception VshipAMD_MemcpyHtoD(void* dst, const void* src, uint64_t size)
{
    return guarded([&] { VSHIP_HIP_CHECK(hipMemcpy(dst, src, size, hipMemcpyHostToDevice)); });
}

Vship_Exception VshipAMD_MemcpyDtoH(void* dst, const void* src, uint64_t size)
{
    return guarded([&] { VSHIP_HIP_CHECK(hipMemcpy(dst, src, size, hipMemcpyDeviceToHost)); });
}

Vship_Exception VshipAMD_DeviceCopy(void* dst, const void* src, uint64_t bytes, int iters, double* ms)
{
    return guarded([&] {
        if (!dst || !src || (bytes % 16) != 0 || iters <= 0) VSHIP_THROW(Vship_BadPointer, "bad copy arguments");
        hipEvent_t e0, e1;
        VSHIP_HIP_CHECK(hipEventCreate(&e0));
        VSHIP_HIP_CHECK(hipEventCreate(&e1));
        VSHIP_HIP_CHECK(hipEventRecord(e0, 0));
        for (int i = 0; i < iters; i++) VSHIP_HIP_CHECK(launch_copy(dst, src, (int64_t)bytes, 0));
        VSHIP_HIP_CHECK(hipEventRecord(e1, 0));
        VSHIP_HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        VSHIP_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = t / iters;
    });
}

Vship_Exception VshipAMD_DeviceSynchronize(void)
{
    return guarded([&] { VSHIP_HIP_CHECK(hipDeviceSynchronize()); });
}

}  // extern "C"
