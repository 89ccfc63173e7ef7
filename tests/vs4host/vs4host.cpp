// A minimal, single-threaded VapourSynth API v4 host for the plugin tests.
//
// It loads a plugin with dlopen, calls VapourSynthPluginInit2, and implements the part of
// VSAPI the plugin uses: maps, nodes, frames, createVideoFilter and the arInitial /
// arAllFramesReady frame protocol. `invoke` checks arguments against the registered signature
// the way VapourSynth does: unknown keys, missing required ones, type mismatches, more than one
// value for a non-array argument, and int -> float conversion for float arguments. It also
// counts protocol misuse (a frame fetched without a request, a request outside a
// StrictSpatial dependency's frame n, a map getter called without an error pointer on a bad
// key) and live frame and node references, so tests can assert that the plugin is clean.
// Python drives it through the h_* functions (ctypes).
#include <dlfcn.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "vs4_abi.h"

namespace {

int g_live_frames = 0, g_live_nodes = 0, g_misuse = 0;
std::string g_misuse_log;

void misuse(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ++g_misuse;
    g_misuse_log += buf;
    g_misuse_log += "\n";
}

}  // namespace

struct VSMap {
    struct Entry {
        std::string key;
        int type = ptUnset;
        std::vector<int64_t> i;
        std::vector<double> f;
        std::vector<std::string> d;
        std::vector<VSNode *> n;
    };
    std::vector<Entry> e;
    std::string error;
    bool has_error = false;

    Entry *find(const char *k) {
        for (auto &x : e)
            if (x.key == k) return &x;
        return nullptr;
    }
    const Entry *find(const char *k) const { return const_cast<VSMap *>(this)->find(k); }
    ~VSMap();
};

struct VSFrame {
    VSVideoFormat fmt{};
    int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    ptrdiff_t stride[3] = {0, 0, 0};
    std::vector<uint8_t> data[3];
    VSMap props;
    int refs = 1;
};

struct VSNode {
    VSVideoInfo vi{};
    std::string name;
    int refs = 1;
    // source node
    std::vector<std::vector<float>> frames;  // [n] planar, planes back to back
    // filter node
    VSFilterGetFrame get = nullptr;
    VSFilterFree free_fn = nullptr;
    void *inst = nullptr;
    std::vector<VSFilterDependency> deps;
};

struct VSFrameContext {
    int n = 0;
    VSNode *node = nullptr;
    std::vector<std::pair<VSNode *, int>> requested;
    std::map<std::pair<VSNode *, int>, const VSFrame *> ready;
    std::string error;
    bool has_error = false;
};

struct VSCore {
    VSPlugin *plugin = nullptr;  // the loaded plugin (getPluginByID)
    std::string plugin_id, plugin_ns, plugin_name;
    int plugin_version = 0, plugin_api = 0;
    struct Fn {
        std::string name, args, ret;
        VSPublicFunction fn;
        void *data;
    };
    std::vector<Fn> fns;
    void *dl = nullptr;
};

struct VSPlugin {
    VSCore *core;
};

namespace {

const VSAPI *api();

void node_unref(VSNode *n);

void frame_unref(const VSFrame *cf) {
    VSFrame *f = const_cast<VSFrame *>(cf);
    if (!f) return;
    if (--f->refs == 0) {
        delete f;
        --g_live_frames;
    }
}

void node_unref(VSNode *n) {
    if (!n) return;
    if (--n->refs == 0) {
        if (n->free_fn) n->free_fn(n->inst, nullptr, api());
        delete n;
        --g_live_nodes;
    }
}

int plane_w(const VSVideoFormat &f, int w, int p) { return p == 0 ? w : w >> f.subSamplingW; }
int plane_h(const VSVideoFormat &f, int h, int p) { return p == 0 ? h : h >> f.subSamplingH; }

VSFrame *frame_new(const VSVideoFormat *fmt, int w, int h) {
    VSFrame *f = new VSFrame;
    ++g_live_frames;
    f->fmt = *fmt;
    for (int p = 0; p < fmt->numPlanes && p < 3; ++p) {
        f->w[p] = plane_w(*fmt, w, p);
        f->h[p] = plane_h(*fmt, h, p);
        f->stride[p] = (ptrdiff_t(f->w[p]) * fmt->bytesPerSample + 63) & ~ptrdiff_t(63);
        f->data[p].assign(size_t(f->stride[p]) * f->h[p], 0xCD);  // garbage fill: unwritten rows show up
    }
    return f;
}

void copy_map(const VSMap &src, VSMap &dst);

// ---- VSAPI ----------------------------------------------------------------------------------
void createVideoFilter(VSMap *out, const char *name, const VSVideoInfo *vi, VSFilterGetFrame get, VSFilterFree fr,
                       int mode, const VSFilterDependency *deps, int nd, void *inst, VSCore *) {
    if (mode != fmParallel && mode != fmParallelRequests && mode != fmUnordered && mode != fmFrameState)
        misuse("createVideoFilter: bad filter mode %d", mode);
    if (vi->numFrames <= 0 || vi->width <= 0 || vi->height <= 0) misuse("createVideoFilter: bad video info");
    VSNode *n = new VSNode;
    ++g_live_nodes;
    n->vi = *vi;
    n->name = name;
    n->get = get;
    n->free_fn = fr;
    n->inst = inst;
    for (int i = 0; i < nd; ++i) n->deps.push_back(deps[i]);
    VSMap::Entry e;
    e.key = "clip";
    e.type = ptVideoNode;
    e.n.push_back(n);  // the map owns the creation reference
    out->e.push_back(e);
}
void freeNode(VSNode *n) { node_unref(n); }
VSNode *addNodeRef(VSNode *n) {
    ++n->refs;
    return n;
}
int getNodeType(VSNode *) { return mtVideo; }
const VSVideoInfo *getVideoInfo(VSNode *n) { return &n->vi; }

VSFrame *newVideoFrame(const VSVideoFormat *fmt, int w, int h, const VSFrame *propSrc, VSCore *) {
    VSFrame *f = frame_new(fmt, w, h);
    if (propSrc) copy_map(propSrc->props, f->props);
    return f;
}
void freeFrame(const VSFrame *f) { frame_unref(f); }
const VSFrame *addFrameRef(const VSFrame *f) {
    ++const_cast<VSFrame *>(f)->refs;
    return f;
}
const VSMap *getFramePropertiesRO(const VSFrame *f) { return &f->props; }
VSMap *getFramePropertiesRW(VSFrame *f) {
    if (f->refs != 1) misuse("getFramePropertiesRW on a shared frame");
    return &f->props;
}
ptrdiff_t getStride(const VSFrame *f, int p) { return f->stride[p]; }
const uint8_t *getReadPtr(const VSFrame *f, int p) { return f->data[p].data(); }
uint8_t *getWritePtr(VSFrame *f, int p) {
    if (f->refs != 1) misuse("getWritePtr on a shared frame");
    return f->data[p].data();
}
const VSVideoFormat *getVideoFrameFormat(const VSFrame *f) { return &f->fmt; }
int getFrameWidth(const VSFrame *f, int p) { return f->w[p]; }
int getFrameHeight(const VSFrame *f, int p) { return f->h[p]; }

const VSFrame *getFrameFilter(int n, VSNode *node, VSFrameContext *ctx) {
    auto it = ctx->ready.find({node, n});
    if (it == ctx->ready.end()) {
        misuse("getFrameFilter(%d) of node %s without a request", n, node->name.c_str());
        return nullptr;
    }
    return addFrameRef(it->second);
}
void requestFrameFilter(int n, VSNode *node, VSFrameContext *ctx) {
    bool dep = false;
    for (auto &d : ctx->node->deps)
        if (d.source == node) {
            dep = true;
            if (d.requestPattern == rpStrictSpatial && n != ctx->n)
                misuse("%s requested frame %d of a StrictSpatial dependency at frame %d", ctx->node->name.c_str(), n, ctx->n);
        }
    if (!dep) misuse("%s requested a frame of an undeclared dependency", ctx->node->name.c_str());
    if (n < 0 || n >= node->vi.numFrames) misuse("request of frame %d outside 0..%d", n, node->vi.numFrames - 1);
    ctx->requested.emplace_back(node, n);
}
void setFilterError(const char *msg, VSFrameContext *ctx) {
    ctx->error = msg;
    ctx->has_error = true;
}

VSMap *createMap() { return new VSMap; }
void freeMap(VSMap *m) { delete m; }
void mapSetError(VSMap *m, const char *msg) {
    m->e.clear();
    m->error = msg;
    m->has_error = true;
}
const char *mapGetError(const VSMap *m) { return m->has_error ? m->error.c_str() : nullptr; }
int mapNumElements(const VSMap *m, const char *k) {
    const VSMap::Entry *e = m->find(k);
    if (!e) return -1;
    switch (e->type) {
        case ptInt: return int(e->i.size());
        case ptFloat: return int(e->f.size());
        case ptData: return int(e->d.size());
        case ptVideoNode: return int(e->n.size());
        default: return 0;
    }
}

const VSMap::Entry *lookup(const VSMap *m, const char *k, int type, int idx, int *err, size_t size_of_type(const VSMap::Entry &)) {
    const VSMap::Entry *e = m->find(k);
    int code = peSuccess;
    if (!e) code = peUnset;
    else if (e->type != type) code = peType;
    else if (idx < 0 || size_t(idx) >= size_of_type(*e)) code = peIndex;
    if (err) *err = code;
    else if (code != peSuccess) misuse("map getter on \"%s\" failed (%d) without an error pointer", k, code);
    return code == peSuccess ? e : nullptr;
}
size_t n_i(const VSMap::Entry &e) { return e.i.size(); }
size_t n_f(const VSMap::Entry &e) { return e.f.size(); }
size_t n_d(const VSMap::Entry &e) { return e.d.size(); }
size_t n_n(const VSMap::Entry &e) { return e.n.size(); }

int64_t mapGetInt(const VSMap *m, const char *k, int idx, int *err) {
    const VSMap::Entry *e = lookup(m, k, ptInt, idx, err, n_i);
    return e ? e->i[idx] : 0;
}
int mapGetIntSaturated(const VSMap *m, const char *k, int idx, int *err) {
    const int64_t v = mapGetInt(m, k, idx, err);
    return int(std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX));
}
double mapGetFloat(const VSMap *m, const char *k, int idx, int *err) {
    const VSMap::Entry *e = lookup(m, k, ptFloat, idx, err, n_f);
    return e ? e->f[idx] : 0.0;
}
float mapGetFloatSaturated(const VSMap *m, const char *k, int idx, int *err) { return float(mapGetFloat(m, k, idx, err)); }
const char *mapGetData(const VSMap *m, const char *k, int idx, int *err) {
    const VSMap::Entry *e = lookup(m, k, ptData, idx, err, n_d);
    return e ? e->d[idx].c_str() : nullptr;
}
int mapGetDataSize(const VSMap *m, const char *k, int idx, int *err) {
    const VSMap::Entry *e = lookup(m, k, ptData, idx, err, n_d);
    return e ? int(e->d[idx].size()) : -1;
}
VSNode *mapGetNode(const VSMap *m, const char *k, int idx, int *err) {
    const VSMap::Entry *e = lookup(m, k, ptVideoNode, idx, err, n_n);
    return e ? addNodeRef(e->n[idx]) : nullptr;
}

VSMap::Entry &slot(VSMap *m, const char *k, int type, int append) {
    VSMap::Entry *e = m->find(k);
    if (e && (!append || e->type != type)) {
        for (VSNode *n : e->n) node_unref(n);
        m->e.erase(m->e.begin() + (e - m->e.data()));
        e = nullptr;
    }
    if (!e) {
        m->e.emplace_back();
        e = &m->e.back();
        e->key = k;
        e->type = type;
    }
    return *e;
}
int mapSetInt(VSMap *m, const char *k, int64_t v, int append) {
    slot(m, k, ptInt, append).i.push_back(v);
    return 0;
}
int mapSetIntArray(VSMap *m, const char *k, const int64_t *v, int n) {
    VSMap::Entry &e = slot(m, k, ptInt, maReplace);
    e.i.assign(v, v + n);
    return 0;
}
int mapSetNode(VSMap *m, const char *k, VSNode *n, int append) {
    slot(m, k, ptVideoNode, append).n.push_back(addNodeRef(n));
    return 0;
}
int mapConsumeNode(VSMap *m, const char *k, VSNode *n, int append) {
    slot(m, k, ptVideoNode, append).n.push_back(n);
    return 0;
}
extern "C" VSMap *h_invoke(VSCore *core, const char *name, VSMap *in);
VSPlugin *getPluginByID(const char *id, VSCore *core) {
    return core && core->plugin && core->plugin_id == id ? core->plugin : nullptr;
}
// A plugin invoking one of its own functions: the same checks as h_invoke, on a copy of the
// caller's arguments (VapourSynth does not modify them).
VSMap *invoke(VSPlugin *pl, const char *name, const VSMap *args) {
    VSMap copy;
    copy_map(*args, copy);
    return h_invoke(pl->core, name, &copy);
}
int getAPIVersion() { return VSZ_VS_API_VERSION; }
void logMessage(int type, const char *msg, VSCore *) { std::fprintf(stderr, "vs4host log %d: %s\n", type, msg); }

void copy_map(const VSMap &src, VSMap &dst) {
    dst.e.clear();
    for (const auto &x : src.e) {
        dst.e.push_back(x);
        for (VSNode *n : dst.e.back().n) addNodeRef(n);
    }
}

VSAPI make_api() {
    VSAPI a;
    std::memset(&a, 0, sizeof a);
    a.createVideoFilter = createVideoFilter;
    a.freeNode = freeNode;
    a.addNodeRef = addNodeRef;
    a.getNodeType = getNodeType;
    a.getVideoInfo = getVideoInfo;
    a.newVideoFrame = newVideoFrame;
    a.freeFrame = freeFrame;
    a.addFrameRef = addFrameRef;
    a.getFramePropertiesRO = getFramePropertiesRO;
    a.getFramePropertiesRW = getFramePropertiesRW;
    a.getStride = getStride;
    a.getReadPtr = getReadPtr;
    a.getWritePtr = getWritePtr;
    a.getVideoFrameFormat = getVideoFrameFormat;
    a.getFrameWidth = getFrameWidth;
    a.getFrameHeight = getFrameHeight;
    a.getFrameFilter = getFrameFilter;
    a.requestFrameFilter = requestFrameFilter;
    a.setFilterError = setFilterError;
    a.createMap = createMap;
    a.freeMap = freeMap;
    a.mapSetError = mapSetError;
    a.mapGetError = mapGetError;
    a.mapNumElements = mapNumElements;
    a.mapGetInt = mapGetInt;
    a.mapGetIntSaturated = mapGetIntSaturated;
    a.mapSetInt = mapSetInt;
    a.mapSetIntArray = mapSetIntArray;
    a.mapGetFloat = mapGetFloat;
    a.mapGetFloatSaturated = mapGetFloatSaturated;
    a.mapGetData = mapGetData;
    a.mapGetDataSize = mapGetDataSize;
    a.mapGetNode = mapGetNode;
    a.mapSetNode = mapSetNode;
    a.mapConsumeNode = mapConsumeNode;
    a.getAPIVersion = getAPIVersion;
    a.logMessage = logMessage;
    a.getPluginByID = getPluginByID;
    a.invoke = invoke;
    return a;
}

const VSAPI *api() {
    static const VSAPI a = make_api();
    return &a;
}

// ---- VSPLUGINAPI ----------------------------------------------------------------------------
int p_getAPIVersion() { return VSZ_VS_API_VERSION; }
int p_configPlugin(const char *id, const char *ns, const char *name, int pv, int apiv, int, VSPlugin *pl) {
    pl->core->plugin_id = id;
    pl->core->plugin_ns = ns;
    pl->core->plugin_name = name;
    pl->core->plugin_version = pv;
    pl->core->plugin_api = apiv;
    return 1;
}
int p_registerFunction(const char *name, const char *args, const char *ret, VSPublicFunction fn, void *data,
                       VSPlugin *pl) {
    pl->core->fns.push_back({name, args, ret, fn, data});
    return 1;
}

// ---- frame evaluation -----------------------------------------------------------------------
VSFrame *source_frame(VSNode *node, int n) {
    const VSVideoFormat &fmt = node->vi.format;
    VSFrame *f = frame_new(&fmt, node->vi.width, node->vi.height);
    const float *src = node->frames[n].data();
    for (int p = 0; p < fmt.numPlanes; ++p) {
        for (int y = 0; y < f->h[p]; ++y) std::memcpy(f->data[p].data() + y * f->stride[p], src + size_t(y) * f->w[p], size_t(f->w[p]) * 4);
        src += size_t(f->w[p]) * f->h[p];
    }
    mapSetInt(&f->props, "_SourceFrame", n, maReplace);
    return f;
}

const VSFrame *eval(VSNode *node, int n, std::string &err) {
    if (n < 0 || n >= node->vi.numFrames) {
        err = "frame number out of range";
        return nullptr;
    }
    if (!node->get) return source_frame(node, n);
    VSFrameContext ctx;
    ctx.n = n;
    ctx.node = node;
    void *frame_data = nullptr;
    const VSFrame *r = node->get(n, arInitial, node->inst, &frame_data, &ctx, nullptr, api());
    if (r) {
        misuse("%s returned a frame in arInitial", node->name.c_str());
        frame_unref(r);
    }
    for (auto &q : ctx.requested) {
        if (ctx.ready.count(q)) continue;
        const VSFrame *f = eval(q.first, q.second, err);
        if (!f) {
            for (auto &kv : ctx.ready) frame_unref(kv.second);
            return nullptr;
        }
        ctx.ready[q] = f;
    }
    r = node->get(n, arAllFramesReady, node->inst, &frame_data, &ctx, nullptr, api());
    for (auto &kv : ctx.ready) frame_unref(kv.second);
    if (!r) err = ctx.has_error ? ctx.error : std::string("getFrame returned no frame without an error");
    else if (ctx.has_error) misuse("%s set an error and returned a frame", node->name.c_str());
    return r;
}

// signature entry: name, type, array, optional
struct Arg {
    std::string name, type;
    bool array = false, opt = false;
};

std::vector<Arg> parse_sig(const std::string &s) {
    std::vector<Arg> out;
    size_t pos = 0;
    while (pos < s.size()) {
        size_t end = s.find(';', pos);
        if (end == std::string::npos) end = s.size();
        const std::string item = s.substr(pos, end - pos);
        pos = end + 1;
        if (item.empty()) continue;
        std::vector<std::string> parts;
        size_t p = 0;
        while (true) {
            const size_t c = item.find(':', p);
            parts.push_back(item.substr(p, c == std::string::npos ? std::string::npos : c - p));
            if (c == std::string::npos) break;
            p = c + 1;
        }
        Arg a;
        a.name = parts[0];
        a.type = parts.size() > 1 ? parts[1] : "";
        if (a.type.size() > 2 && a.type.compare(a.type.size() - 2, 2, "[]") == 0) {
            a.array = true;
            a.type.resize(a.type.size() - 2);
        }
        for (size_t i = 2; i < parts.size(); ++i)
            if (parts[i] == "opt") a.opt = true;
        out.push_back(a);
    }
    return out;
}

}  // namespace

VSMap::~VSMap() {
    for (auto &x : e)
        for (VSNode *n : x.n) node_unref(n);
}

// ---- C API for the Python driver --------------------------------------------------------------
extern "C" {

VSCore *h_core_new(void) { return new VSCore; }

int h_load_plugin(VSCore *core, const char *path, char *err, int errlen) {
    core->dl = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!core->dl) {
        std::snprintf(err, size_t(errlen), "%s", dlerror());
        return -1;
    }
    auto init = reinterpret_cast<void (*)(VSPlugin *, const VSPLUGINAPI *)>(dlsym(core->dl, "VapourSynthPluginInit2"));
    if (!init) {
        std::snprintf(err, size_t(errlen), "no VapourSynthPluginInit2");
        return -2;
    }
    static const VSPLUGINAPI papi = {p_getAPIVersion, p_configPlugin, p_registerFunction};
    if (!core->plugin) core->plugin = new VSPlugin{core};
    init(core->plugin, &papi);
    return int(core->fns.size());
}

const char *h_plugin_string(VSCore *core, int which) {
    return which == 0 ? core->plugin_id.c_str() : which == 1 ? core->plugin_ns.c_str() : core->plugin_name.c_str();
}
int h_plugin_version(VSCore *core, int which) { return which == 0 ? core->plugin_version : core->plugin_api; }
const char *h_function(VSCore *core, int i, int which) {
    const auto &f = core->fns[size_t(i)];
    return which == 0 ? f.name.c_str() : which == 1 ? f.args.c_str() : f.ret.c_str();
}

VSMap *h_map_new(void) { return new VSMap; }
void h_map_free(VSMap *m) { delete m; }
void h_map_set_int(VSMap *m, const char *k, int64_t v) { mapSetInt(m, k, v, maAppend); }
void h_map_set_float(VSMap *m, const char *k, double v) { slot(m, k, ptFloat, maAppend).f.push_back(v); }
void h_map_set_data(VSMap *m, const char *k, const char *d, int size) { slot(m, k, ptData, maAppend).d.emplace_back(d, size_t(size)); }
void h_map_set_node(VSMap *m, const char *k, VSNode *n) { mapSetNode(m, k, n, maAppend); }
void h_map_set_empty(VSMap *m, const char *k, int type) { slot(m, k, type, maReplace); }  // e.g. planes=[]
const char *h_map_error(VSMap *m) { return mapGetError(m); }
VSNode *h_map_get_node(VSMap *m, const char *k) {
    int err = 0;
    return mapGetNode(m, k, 0, &err);
}

// Invokes registered function `name`: VapourSynth's argument checks, then the plugin's create.
VSMap *h_invoke(VSCore *core, const char *name, VSMap *in) {
    VSMap *out = new VSMap;
    const VSCore::Fn *fn = nullptr;
    for (auto &f : core->fns)
        if (f.name == name) fn = &f;
    if (!fn) {
        mapSetError(out, "no such function");
        return out;
    }
    const std::vector<Arg> sig = parse_sig(fn->args);
    for (auto &x : in->e) {
        const Arg *a = nullptr;
        for (auto &s : sig)
            if (s.name == x.key) a = &s;
        if (!a) {
            mapSetError(out, (std::string(name) + ": argument " + x.key + " does not exist").c_str());
            return out;
        }
        const int want = a->type == "int" ? ptInt : a->type == "float" ? ptFloat : a->type == "data" ? ptData : ptVideoNode;
        if (want == ptFloat && x.type == ptInt) {  // VapourSynth converts int arguments to float
            x.type = ptFloat;
            for (int64_t v : x.i) x.f.push_back(double(v));
            x.i.clear();
        }
        if (x.type != want) {
            mapSetError(out, (std::string(name) + ": argument " + x.key + " is not of the correct type").c_str());
            return out;
        }
        if (!a->array && mapNumElements(in, x.key.c_str()) > 1) {
            mapSetError(out, (std::string(name) + ": argument " + x.key + " is not an array").c_str());
            return out;
        }
    }
    for (auto &s : sig)
        if (!s.opt && !in->find(s.name.c_str())) {
            mapSetError(out, (std::string(name) + ": argument " + s.name + " is required").c_str());
            return out;
        }
    fn->fn(in, out, fn->data, core, api());
    return out;
}

// Source clip whose frames come from `data`: nframes x planes, each plane packed (w_p * h_p floats).
VSNode *h_source_new(int color_family, int sample_type, int bits, int ssw, int ssh, int w, int h, int nframes,
                     const float *data) {
    VSNode *n = new VSNode;
    ++g_live_nodes;
    VSVideoFormat f{color_family, sample_type, bits, (bits + 7) / 8, ssw, ssh, color_family == cfGray ? 1 : 3};
    n->vi = VSVideoInfo{f, 24000, 1001, w, h, nframes};
    n->name = "source";
    size_t per = 0;
    for (int p = 0; p < f.numPlanes; ++p) per += size_t(plane_w(f, w, p)) * plane_h(f, h, p);
    for (int i = 0; i < nframes; ++i) n->frames.emplace_back(data + per * i, data + per * (i + 1));
    return n;
}
void h_node_free(VSNode *n) { node_unref(n); }
void h_node_info(VSNode *n, int *info) {  // color, sample, bits, ssw, ssh, planes, w, h, frames
    const VSVideoInfo &v = n->vi;
    const int vals[9] = {v.format.colorFamily, v.format.sampleType, v.format.bitsPerSample, v.format.subSamplingW,
                         v.format.subSamplingH, v.format.numPlanes, v.width, v.height, v.numFrames};
    std::memcpy(info, vals, sizeof vals);
}
const char *h_node_name(VSNode *n) { return n->name.c_str(); }
int h_node_dep(VSNode *n, int i, int *pattern) {  // request pattern of dependency i, -1 if none
    if (i < 0 || size_t(i) >= n->deps.size()) return -1;
    *pattern = n->deps[size_t(i)].requestPattern;
    return 0;
}

const VSFrame *h_get_frame(VSNode *node, int n, char *err, int errlen) {
    std::string e;
    const VSFrame *f = eval(node, n, e);
    if (!f) std::snprintf(err, size_t(errlen), "%s", e.c_str());
    return f;
}
void h_frame_free(const VSFrame *f) { frame_unref(f); }
int h_frame_plane(const VSFrame *f, int p, float *dst) {  // packed copy, returns w*h
    for (int y = 0; y < f->h[p]; ++y) std::memcpy(dst + size_t(y) * f->w[p], f->data[p].data() + y * f->stride[p], size_t(f->w[p]) * 4);
    return f->w[p] * f->h[p];
}
void h_frame_dims(const VSFrame *f, int p, int *wh) {
    wh[0] = f->w[p];
    wh[1] = f->h[p];
}
int h_frame_prop_ints(const VSFrame *f, const char *key, int64_t *out, int max) {
    const VSMap::Entry *e = f->props.find(key);
    if (!e || e->type != ptInt) return -1;
    const int n = int(std::min<size_t>(e->i.size(), size_t(max)));
    for (int i = 0; i < n; ++i) out[i] = e->i[size_t(i)];
    return int(e->i.size());
}

int h_live_frames(void) { return g_live_frames; }
int h_live_nodes(void) { return g_live_nodes; }
int h_misuse(void) { return g_misuse; }
const char *h_misuse_log(void) { return g_misuse_log.c_str(); }

}  // extern "C"
