// mock_vs.cpp - a minimal VapourSynth API v4 host for testing the vship
// plugin without VapourSynth (test infrastructure).
//
// It loads the plugin library, calls VapourSynthPluginInit2 with its own
// plugin API table, and implements the VSAPI entries the plugin uses: maps,
// source clips held in memory, filter nodes evaluated synchronously
// (arInitial, then the requested frames, then arAllFramesReady), frame and
// node reference counts, and a stand-in for resize.Bicubic that records its
// arguments and returns a clip registered by the test. Any other VSAPI entry
// aborts. A C interface (mock_*) drives it from Python through ctypes.
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../vship_amd/csrc/vs4_abi.hpp"

using namespace vs4;

struct Value {
    int type = 0;  // 1 int, 2 float, 3 data, 5 node
    int64_t i = 0;
    double f = 0.0;
    std::string data;
    VSNode* node = nullptr;
};

struct vs4::VSMap {
    std::map<std::string, std::vector<Value>> kv;
    std::string error;
};

static std::atomic<int> g_live_frames{0}, g_live_nodes{0};

struct vs4::VSFrame {
    VSVideoFormat fmt{};
    int w = 0, h = 0;
    int pw[3] = {0, 0, 0}, ph[3] = {0, 0, 0};
    ptrdiff_t stride[3] = {0, 0, 0};
    std::vector<uint8_t> plane[3];
    VSMap props;
    mutable std::atomic<int> refs{1};
};

struct vs4::VSNode {
    VSVideoInfo vi{};
    std::vector<VSFrame*> frames;  // source clip
    VSFilterGetFrame get = nullptr;
    VSFilterFree free_fn = nullptr;
    void* instance = nullptr;
    std::vector<VSFilterDependency> deps;
    std::string name;
    int mode = -1;
    std::atomic<int> refs{1};
};

struct vs4::VSFrameContext {
    std::vector<std::pair<VSNode*, int>> requested;
    std::map<std::pair<VSNode*, int>, const VSFrame*> ready;
    std::string error;
};

struct vs4::VSCore {};

struct Registered {
    std::string name, args, ret;
    VSPublicFunction fn;
    void* data;
};

struct vs4::VSPlugin {
    std::string id, ns, name;
    int version = 0, api = 0;
    std::vector<Registered> funcs;
};

static VSCore g_core;
static VSPlugin g_plugin;
static VSPlugin g_resize;  // stand-in for com.vapoursynth.resize
static std::map<VSNode*, VSNode*> g_resize_map;
static std::string g_resize_args;
static int g_core_threads = 4;
static VSAPI g_api;
static VSPLUGINAPI g_papi;

static void trap(void)
{
    fprintf(stderr, "mock_vs: the plugin called a VSAPI entry the mock does not provide\n");
    abort();
}

static void node_ref(VSNode* n) { n->refs++; }

static void free_node(VSNode* n)
{
    if (!n) return;
    if (--n->refs > 0) return;
    if (n->free_fn) n->free_fn(n->instance, &g_core, &g_api);
    for (VSFrame* f : n->frames) {
        if (--f->refs == 0) {
            delete f;
            g_live_frames--;
        }
    }
    delete n;
    g_live_nodes--;
}

static void free_frame(const VSFrame* f)
{
    if (!f) return;
    if (--f->refs == 0) {
        delete f;
        g_live_frames--;
    }
}

static VSFrame* new_frame(const VSVideoFormat& fmt, int w, int h)
{
    VSFrame* f = new VSFrame;
    f->fmt = fmt;
    f->w = w;
    f->h = h;
    for (int p = 0; p < fmt.numPlanes; p++) {
        f->pw[p] = p ? (w >> fmt.subSamplingW) : w;
        f->ph[p] = p ? (h >> fmt.subSamplingH) : h;
        f->stride[p] = ((ptrdiff_t)f->pw[p] * fmt.bytesPerSample + 63) / 64 * 64;  // 64-byte aligned rows
        f->plane[p].assign((size_t)f->stride[p] * f->ph[p], 0);
    }
    g_live_frames++;
    return f;
}

static const VSFrame* evaluate(VSNode* node, int n, std::string& err);

// ---- VSAPI entries

static void api_create_video_filter(VSMap* out, const char* name, const VSVideoInfo* vi, VSFilterGetFrame getFrame,
                                    VSFilterFree free_fn, int mode, const VSFilterDependency* deps, int ndeps,
                                    void* instance, VSCore*)
{
    VSNode* n = new VSNode;
    g_live_nodes++;
    n->vi = *vi;
    n->get = getFrame;
    n->free_fn = free_fn;
    n->instance = instance;
    n->deps.assign(deps, deps + ndeps);
    n->name = name;
    n->mode = mode;
    Value v;
    v.type = 5;
    v.node = n;  // the map holds the creation reference
    out->kv["clip"].push_back(v);
}

static const VSVideoInfo* api_get_video_info(VSNode* n) { return &n->vi; }

static VSFrame* api_copy_frame(const VSFrame* f, VSCore*)
{
    VSFrame* c = new VSFrame;
    c->fmt = f->fmt;
    c->w = f->w;
    c->h = f->h;
    for (int p = 0; p < 3; p++) {
        c->pw[p] = f->pw[p];
        c->ph[p] = f->ph[p];
        c->stride[p] = f->stride[p];
        c->plane[p] = f->plane[p];
    }
    c->props = f->props;
    g_live_frames++;
    return c;
}

static VSMap* api_props_rw(VSFrame* f) { return &f->props; }
static ptrdiff_t api_get_stride(const VSFrame* f, int p) { return f->stride[p]; }
static const uint8_t* api_get_read_ptr(const VSFrame* f, int p) { return f->plane[p].data(); }
static int api_frame_width(const VSFrame* f, int p) { return f->pw[p]; }
static int api_frame_height(const VSFrame* f, int p) { return f->ph[p]; }

static const VSFrame* api_get_frame_filter(int n, VSNode* node, VSFrameContext* ctx)
{
    auto it = ctx->ready.find({node, n});
    if (it == ctx->ready.end()) return nullptr;
    it->second->refs++;
    return it->second;
}

static void api_request_frame_filter(int n, VSNode* node, VSFrameContext* ctx)
{
    ctx->requested.push_back({node, n});
}

static void api_set_filter_error(const char* msg, VSFrameContext* ctx) { ctx->error = msg; }

static VSMap* api_create_map(void) { return new VSMap; }

static void release_values(VSMap* m)
{
    for (auto& kv : m->kv)
        for (auto& v : kv.second)
            if (v.type == 5) free_node(v.node);
    m->kv.clear();
}

static void api_free_map(VSMap* m)
{
    if (!m) return;
    release_values(m);
    delete m;
}

static void api_map_set_error(VSMap* m, const char* msg)
{
    release_values(m);
    m->error = msg ? msg : "";
}

static const char* api_map_get_error(const VSMap* m) { return m->error.empty() ? nullptr : m->error.c_str(); }

static const Value* lookup(const VSMap* m, const char* key, int index, int type, int* error)
{
    auto it = m->kv.find(key);
    int e = 0;
    const Value* v = nullptr;
    if (it == m->kv.end()) e = 1;            // peUnset
    else if (index < 0 || index >= (int)it->second.size()) e = 4;  // peIndex
    else if (it->second[index].type != type) e = 2;  // peType
    else v = &it->second[index];
    if (error) *error = e;
    return v;
}

static void set_value(VSMap* m, const char* key, const Value& v, int append)
{
    auto& vec = m->kv[key];
    if (append == maReplace) {
        for (auto& old : vec)
            if (old.type == 5) free_node(old.node);
        vec.clear();
    }
    vec.push_back(v);
}

static int64_t api_map_get_int(const VSMap* m, const char* key, int index, int* error)
{
    const Value* v = lookup(m, key, index, 1, error);
    return v ? v->i : 0;
}

static int api_map_set_int(VSMap* m, const char* key, int64_t i, int append)
{
    Value v;
    v.type = 1;
    v.i = i;
    set_value(m, key, v, append);
    return 0;
}

static int api_map_set_float(VSMap* m, const char* key, double d, int append)
{
    Value v;
    v.type = 2;
    v.f = d;
    set_value(m, key, v, append);
    return 0;
}

static int api_map_set_data(VSMap* m, const char* key, const char* data, int size, int, int append)
{
    Value v;
    v.type = 3;
    v.data.assign(data, size < 0 ? strlen(data) : (size_t)size);
    set_value(m, key, v, append);
    return 0;
}

static VSNode* api_map_get_node(const VSMap* m, const char* key, int index, int* error)
{
    const Value* v = lookup(m, key, index, 5, error);
    if (!v) return nullptr;
    node_ref(v->node);
    return v->node;
}

static int api_map_consume_node(VSMap* m, const char* key, VSNode* node, int append)
{
    Value v;
    v.type = 5;
    v.node = node;
    set_value(m, key, v, append);
    return 0;
}

static VSPlugin* api_get_plugin_by_id(const char* id, VSCore*)
{
    if (strcmp(id, "com.vapoursynth.resize") == 0) return &g_resize;
    if (g_plugin.id == id) return &g_plugin;
    return nullptr;
}

static VSMap* api_invoke(VSPlugin* plugin, const char* name, const VSMap* args)
{
    VSMap* ret = new VSMap;
    if (plugin != &g_resize || strcmp(name, "Bicubic") != 0) {
        ret->error = "mock: unknown function";
        return ret;
    }
    std::ostringstream ss;
    for (const char* k : {"matrix_in", "transfer_in", "primaries_in", "range_in", "matrix", "transfer", "primaries",
                          "range", "format"}) {
        int e = 0;
        const int64_t v = api_map_get_int(args, k, 0, &e);
        ss << k << "=" << (e ? std::string("unset") : std::to_string(v)) << ";";
    }
    g_resize_args = ss.str();
    int e = 0;
    const Value* src = lookup(args, "clip", 0, 5, &e);
    auto it = src ? g_resize_map.find(src->node) : g_resize_map.end();
    if (it == g_resize_map.end()) {
        ret->error = "mock resize: no conversion registered for this clip";
        return ret;
    }
    node_ref(it->second);
    Value v;
    v.type = 5;
    v.node = it->second;
    ret->kv["clip"].push_back(v);
    return ret;
}

static void api_get_core_info(VSCore*, VSCoreInfo* info)
{
    info->versionString = "mock";
    info->core = 0;
    info->api = kApiVersion;
    info->numThreads = g_core_threads;
    info->maxFramebufferSize = 1 << 30;
    info->usedFramebufferSize = 0;
}

// ---- plugin API

static int papi_version(void) { return kApiVersion; }

static int papi_config(const char* id, const char* ns, const char* name, int version, int api, int, VSPlugin* p)
{
    p->id = id;
    p->ns = ns;
    p->name = name;
    p->version = version;
    p->api = api;
    return 1;
}

static int papi_register(const char* name, const char* args, const char* ret, VSPublicFunction fn, void* data,
                         VSPlugin* p)
{
    p->funcs.push_back({name, args, ret, fn, data});
    return 1;
}

static void init_tables()
{
    VSUnusedSlot* slots = reinterpret_cast<VSUnusedSlot*>(&g_api);
    for (size_t i = 0; i < sizeof(VSAPI) / sizeof(VSUnusedSlot); i++) slots[i] = trap;
    g_api.createVideoFilter = api_create_video_filter;
    g_api.freeNode = free_node;
    g_api.getVideoInfo = api_get_video_info;
    g_api.freeFrame = free_frame;
    g_api.copyFrame = api_copy_frame;
    g_api.getFramePropertiesRW = api_props_rw;
    g_api.getStride = api_get_stride;
    g_api.getReadPtr = api_get_read_ptr;
    g_api.getFrameWidth = api_frame_width;
    g_api.getFrameHeight = api_frame_height;
    g_api.getFrameFilter = api_get_frame_filter;
    g_api.requestFrameFilter = api_request_frame_filter;
    g_api.setFilterError = api_set_filter_error;
    g_api.createMap = api_create_map;
    g_api.freeMap = api_free_map;
    g_api.mapSetError = api_map_set_error;
    g_api.mapGetError = api_map_get_error;
    g_api.mapGetInt = api_map_get_int;
    g_api.mapSetInt = api_map_set_int;
    g_api.mapSetFloat = api_map_set_float;
    g_api.mapSetData = api_map_set_data;
    g_api.mapGetNode = api_map_get_node;
    g_api.mapConsumeNode = api_map_consume_node;
    g_api.getPluginByID = api_get_plugin_by_id;
    g_api.invoke = api_invoke;
    g_api.getCoreInfo = api_get_core_info;
    g_papi.getAPIVersion = papi_version;
    g_papi.configPlugin = papi_config;
    g_papi.registerFunction = papi_register;
    g_resize.id = "com.vapoursynth.resize";
    g_resize.ns = "resize";
}

static const VSFrame* evaluate(VSNode* node, int n, std::string& err)
{
    if (!node->get) {
        if (n < 0 || n >= (int)node->frames.size()) {
            err = "mock: frame out of range";
            return nullptr;
        }
        node->frames[n]->refs++;
        return node->frames[n];
    }
    VSFrameContext ctx;
    void* frame_data = nullptr;
    const VSFrame* out = node->get(n, arInitial, node->instance, &frame_data, &ctx, &g_core, &g_api);
    if (!out && ctx.error.empty()) {
        for (auto& r : ctx.requested) {
            if (ctx.ready.count(r)) continue;  // the same frame requested twice (e.g. one clip as both inputs)
            std::string e2;
            const VSFrame* f = evaluate(r.first, r.second, e2);
            if (!f) {
                err = e2;
                for (auto& kv : ctx.ready) free_frame(kv.second);
                return nullptr;
            }
            ctx.ready[r] = f;
        }
        out = node->get(n, arAllFramesReady, node->instance, &frame_data, &ctx, &g_core, &g_api);
    }
    for (auto& kv : ctx.ready) free_frame(kv.second);
    if (!out) err = ctx.error.empty() ? "mock: filter returned no frame" : ctx.error;
    return out;
}

// ---- C interface for the tests

extern "C" {

typedef void (*PluginInit)(VSPlugin*, const VSPLUGINAPI*);

int mock_load(const char* path)
{
    init_tables();
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "mock_vs: %s\n", dlerror());
        return -1;
    }
    PluginInit init = (PluginInit)dlsym(h, "VapourSynthPluginInit2");
    if (!init) return -2;
    g_plugin = VSPlugin();
    init(&g_plugin, &g_papi);
    return (int)g_plugin.funcs.size();
}

const char* mock_plugin_field(int field)
{
    static std::string s;
    switch (field) {
    case 0: s = g_plugin.id; break;
    case 1: s = g_plugin.ns; break;
    case 2: s = g_plugin.name; break;
    case 3: s = std::to_string(g_plugin.version); break;
    default: s = std::to_string(g_plugin.api); break;
    }
    return s.c_str();
}

const char* mock_function_field(int i, int field)
{
    const Registered& r = g_plugin.funcs.at(i);
    return field == 0 ? r.name.c_str() : (field == 1 ? r.args.c_str() : r.ret.c_str());
}

void mock_set_threads(int n) { g_core_threads = n; }

void* mock_clip(int w, int h, int nframes, int family, int sample, int bits, int subw, int subh)
{
    VSNode* n = new VSNode;
    g_live_nodes++;
    VSVideoFormat fmt{};
    fmt.colorFamily = family;
    fmt.sampleType = sample;
    fmt.bitsPerSample = bits;
    fmt.bytesPerSample = bits <= 8 ? 1 : (bits <= 16 ? 2 : 4);
    fmt.subSamplingW = subw;
    fmt.subSamplingH = subh;
    fmt.numPlanes = family == cfGray ? 1 : 3;
    n->vi.format = fmt;
    n->vi.fpsNum = 24;
    n->vi.fpsDen = 1;
    n->vi.width = w;
    n->vi.height = h;
    n->vi.numFrames = nframes;
    for (int i = 0; i < nframes; i++) n->frames.push_back(new_frame(fmt, w, h));
    return n;
}

// Copies rows of `src` (row pitch src_stride bytes) into plane p of frame i.
int mock_fill(void* node, int i, int p, const void* src, int64_t src_stride)
{
    VSNode* n = static_cast<VSNode*>(node);
    if (i < 0 || i >= (int)n->frames.size() || p < 0 || p >= n->vi.format.numPlanes) return -1;
    VSFrame* f = n->frames[i];
    const size_t row = (size_t)f->pw[p] * f->fmt.bytesPerSample;
    for (int y = 0; y < f->ph[p]; y++)
        memcpy(f->plane[p].data() + (size_t)y * f->stride[p], (const uint8_t*)src + (size_t)y * src_stride, row);
    return 0;
}

int64_t mock_stride(void* node, int p) { return static_cast<VSNode*>(node)->frames.at(0)->stride[p]; }

void mock_free_node(void* node) { free_node(static_cast<VSNode*>(node)); }

// resize.Bicubic of `src` returns `rgbs`; the table holds a reference to both
// until mock_clear_resize.
void mock_register_resize(void* src, void* rgbs)
{
    node_ref(static_cast<VSNode*>(src));
    node_ref(static_cast<VSNode*>(rgbs));
    g_resize_map[static_cast<VSNode*>(src)] = static_cast<VSNode*>(rgbs);
}

void mock_clear_resize()
{
    for (auto& kv : g_resize_map) {
        free_node(kv.first);
        free_node(kv.second);
    }
    g_resize_map.clear();
}

const char* mock_last_resize_args() { return g_resize_args.c_str(); }

static const Registered* find(const char* name)
{
    for (const Registered& r : g_plugin.funcs)
        if (r.name == name) return &r;
    return nullptr;
}

// Calls vship.SSIMULACRA2; returns the output node or NULL with the map
// error in err. num_stream / gpu_id < 0 leave the argument unset.
void* mock_ssimu2(void* ref, void* dis, int num_stream, int gpu_id, char* err, int errlen)
{
    const Registered* r = find("SSIMULACRA2");
    if (!r) return nullptr;
    VSMap* in = api_create_map();
    VSMap* out = api_create_map();
    node_ref(static_cast<VSNode*>(ref));
    node_ref(static_cast<VSNode*>(dis));
    api_map_consume_node(in, "reference", static_cast<VSNode*>(ref), maReplace);
    api_map_consume_node(in, "distorted", static_cast<VSNode*>(dis), maReplace);
    if (num_stream >= 0) api_map_set_int(in, "numStream", num_stream, maReplace);
    if (gpu_id >= 0) api_map_set_int(in, "gpu_id", gpu_id, maReplace);
    r->fn(in, out, r->data, &g_core, &g_api);
    VSNode* node = nullptr;
    if (!out->error.empty()) {
        snprintf(err, errlen, "%s", out->error.c_str());
    } else {
        int e = 0;
        node = api_map_get_node(out, "clip", 0, &e);
    }
    api_free_map(in);
    api_free_map(out);
    return node;
}

// Output frame n of a node: the _SSIMULACRA2 property, and the filter
// mode/dependency patterns of the node (mode * 100 + pattern0 * 10 + pattern1).
int mock_frame_score(void* node, int n, double* score, char* err, int errlen)
{
    std::string e;
    const VSFrame* f = evaluate(static_cast<VSNode*>(node), n, e);
    if (!f) {
        snprintf(err, errlen, "%s", e.c_str());
        return -1;
    }
    int pe = 0;
    const Value* v = lookup(&f->props, "_SSIMULACRA2", 0, 2, &pe);
    *score = v ? v->f : -1e300;
    free_frame(f);
    return v ? 0 : -2;
}

int mock_node_shape(void* node)
{
    VSNode* n = static_cast<VSNode*>(node);
    int code = n->mode * 100;
    if (n->deps.size() == 2) code += n->deps[0].requestPattern * 10 + n->deps[1].requestPattern;
    return code;
}

int mock_gpu_info(int gpu_id, char* buf, int len)
{
    const Registered* r = find("GpuInfo");
    if (!r) return -1;
    VSMap* in = api_create_map();
    VSMap* out = api_create_map();
    if (gpu_id >= 0) api_map_set_int(in, "gpu_id", gpu_id, maReplace);
    r->fn(in, out, r->data, &g_core, &g_api);
    int rc = 0;
    if (!out->error.empty()) {
        snprintf(buf, len, "%s", out->error.c_str());
        rc = -2;
    } else {
        int e = 0;
        const Value* v = lookup(out, "gpu_human_data", 0, 3, &e);
        snprintf(buf, len, "%s", v ? v->data.c_str() : "");
    }
    api_free_map(in);
    api_free_map(out);
    return rc;
}

int mock_live_frames() { return g_live_frames.load(); }
int mock_live_nodes() { return g_live_nodes.load(); }

}  // extern "C"
