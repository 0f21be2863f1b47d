// vs4_abi.hpp - the parts of the VapourSynth API v4 binary interface that the
// vship plugin uses.
//
// A VapourSynth host calls a plugin through VapourSynthPluginInit2 and hands
// it two tables of function pointers (VSPLUGINAPI, VSAPI); the plugin reads
// them by position. These declarations keep the published v4 table order and
// struct layouts (VapourSynth's VapourSynth4.h, API 4.0/4.1 - vendored by
// the reference as include/VapourSynth4.h) so that this library can be
// loaded by a real VapourSynth core without its SDK headers; only the entries
// up to the end of the 4.0 table are declared (the plugin uses none of the
// 4.1 additions and never needs sizeof(VSAPI)). Entries the plugin does not
// call keep their position as untyped slots.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace vs4 {

constexpr int make_version(int major, int minor) { return (major << 16) | minor; }
constexpr int kApiVersion = make_version(4, 0);

struct VSFrame;
struct VSNode;
struct VSCore;
struct VSPlugin;
struct VSMap;
struct VSFrameContext;

enum ColorFamily { cfUndefined = 0, cfGray = 1, cfRGB = 2, cfYUV = 3 };
enum SampleType { stInteger = 0, stFloat = 1 };
constexpr uint32_t video_id(int family, int sample, int bits, int subw, int subh)
{
    return ((uint32_t)family << 28) | ((uint32_t)sample << 24) | ((uint32_t)bits << 16) | ((uint32_t)subw << 8) |
           (uint32_t)subh;
}
constexpr uint32_t pfRGBS = video_id(cfRGB, stFloat, 32, 0, 0);

enum FilterMode { fmParallel = 0 };
enum MapPropertyError { peSuccess = 0 };
enum MapAppendMode { maReplace = 0, maAppend = 1 };
enum ActivationReason { arInitial = 0, arAllFramesReady = 1, arError = -1 };
enum DataTypeHint { dtUtf8 = 1 };
enum RequestPattern { rpStrictSpatial = 2 };

struct VSVideoFormat {
    int colorFamily;
    int sampleType;
    int bitsPerSample;
    int bytesPerSample;
    int subSamplingW;
    int subSamplingH;
    int numPlanes;
};

struct VSVideoInfo {
    VSVideoFormat format;
    int64_t fpsNum;
    int64_t fpsDen;
    int width;
    int height;
    int numFrames;
};

struct VSCoreInfo {
    const char* versionString;
    int core;
    int api;
    int numThreads;
    int64_t maxFramebufferSize;
    int64_t usedFramebufferSize;
};

struct VSFilterDependency {
    VSNode* source;
    int requestPattern;
};

struct VSAPI;

typedef void (*VSPublicFunction)(const VSMap* in, VSMap* out, void* userData, VSCore* core, const VSAPI* vsapi);
typedef const VSFrame* (*VSFilterGetFrame)(int n, int activationReason, void* instanceData, void** frameData,
                                           VSFrameContext* frameCtx, VSCore* core, const VSAPI* vsapi);
typedef void (*VSFilterFree)(void* instanceData, VSCore* core, const VSAPI* vsapi);

struct VSPLUGINAPI {
    int (*getAPIVersion)(void);
    int (*configPlugin)(const char* identifier, const char* pluginNamespace, const char* name, int pluginVersion,
                        int apiVersion, int flags, VSPlugin* plugin);
    int (*registerFunction)(const char* name, const char* args, const char* returnType, VSPublicFunction argsFunc,
                            void* functionData, VSPlugin* plugin);
};

typedef void (*VSUnusedSlot)(void);

struct VSAPI {
    // nodes and filters
    void (*createVideoFilter)(VSMap* out, const char* name, const VSVideoInfo* vi, VSFilterGetFrame getFrame,
                              VSFilterFree free, int filterMode, const VSFilterDependency* dependencies, int numDeps,
                              void* instanceData, VSCore* core);
    VSUnusedSlot createVideoFilter2, createAudioFilter, createAudioFilter2, setLinearFilter, setCacheMode,
        setCacheOptions;
    void (*freeNode)(VSNode* node);
    VSUnusedSlot addNodeRef, getNodeType;
    const VSVideoInfo* (*getVideoInfo)(VSNode* node);
    VSUnusedSlot getAudioInfo;
    // frames
    VSUnusedSlot newVideoFrame, newVideoFrame2, newAudioFrame, newAudioFrame2;
    void (*freeFrame)(const VSFrame* f);
    VSUnusedSlot addFrameRef;
    VSFrame* (*copyFrame)(const VSFrame* f, VSCore* core);
    VSUnusedSlot getFramePropertiesRO;
    VSMap* (*getFramePropertiesRW)(VSFrame* f);
    ptrdiff_t (*getStride)(const VSFrame* f, int plane);
    const uint8_t* (*getReadPtr)(const VSFrame* f, int plane);
    VSUnusedSlot getWritePtr, getVideoFrameFormat, getAudioFrameFormat, getFrameType;
    int (*getFrameWidth)(const VSFrame* f, int plane);
    int (*getFrameHeight)(const VSFrame* f, int plane);
    VSUnusedSlot getFrameLength;
    // formats
    VSUnusedSlot getVideoFormatName, getAudioFormatName, queryVideoFormat, queryAudioFormat, queryVideoFormatID,
        getVideoFormatByID;
    // frame requests inside filters
    VSUnusedSlot getFrame, getFrameAsync;
    const VSFrame* (*getFrameFilter)(int n, VSNode* node, VSFrameContext* frameCtx);
    void (*requestFrameFilter)(int n, VSNode* node, VSFrameContext* frameCtx);
    VSUnusedSlot releaseFrameEarly, cacheFrame;
    void (*setFilterError)(const char* errorMessage, VSFrameContext* frameCtx);
    // external functions
    VSUnusedSlot createFunction, freeFunction, addFunctionRef, callFunction;
    // maps
    VSMap* (*createMap)(void);
    void (*freeMap)(VSMap* map);
    VSUnusedSlot clearMap, copyMap;
    void (*mapSetError)(VSMap* map, const char* errorMessage);
    const char* (*mapGetError)(const VSMap* map);
    VSUnusedSlot mapNumKeys, mapGetKey, mapDeleteKey, mapNumElements, mapGetType, mapSetEmpty;
    int64_t (*mapGetInt)(const VSMap* map, const char* key, int index, int* error);
    VSUnusedSlot mapGetIntSaturated, mapGetIntArray;
    int (*mapSetInt)(VSMap* map, const char* key, int64_t i, int append);
    VSUnusedSlot mapSetIntArray, mapGetFloat, mapGetFloatSaturated, mapGetFloatArray;
    int (*mapSetFloat)(VSMap* map, const char* key, double d, int append);
    VSUnusedSlot mapSetFloatArray, mapGetData, mapGetDataSize, mapGetDataTypeHint;
    int (*mapSetData)(VSMap* map, const char* key, const char* data, int size, int type, int append);
    VSNode* (*mapGetNode)(const VSMap* map, const char* key, int index, int* error);
    VSUnusedSlot mapSetNode;
    int (*mapConsumeNode)(VSMap* map, const char* key, VSNode* node, int append);
    VSUnusedSlot mapGetFrame, mapSetFrame, mapConsumeFrame, mapGetFunction, mapSetFunction, mapConsumeFunction;
    // plugins
    VSUnusedSlot registerFunction;
    VSPlugin* (*getPluginByID)(const char* identifier, VSCore* core);
    VSUnusedSlot getPluginByNamespace, getNextPlugin, getPluginName, getPluginID, getPluginNamespace,
        getNextPluginFunction, getPluginFunctionByName, getPluginFunctionName, getPluginFunctionArguments,
        getPluginFunctionReturnType, getPluginPath, getPluginVersion;
    VSMap* (*invoke)(VSPlugin* plugin, const char* name, const VSMap* args);
    // core
    VSUnusedSlot createCore, freeCore, setMaxCacheSize, setThreadCount;
    void (*getCoreInfo)(VSCore* core, VSCoreInfo* info);
    VSUnusedSlot getAPIVersion, logMessage, addLogHandler, removeLogHandler;
};

}  // namespace vs4
