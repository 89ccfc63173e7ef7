// VapourSynth API v4 plugin entry: registers vszipcu.BM3D / BM3Dv2 / VAggregate over the C ABI
// of vszipcu_bm3d.h. This is the part of src/plugin.zig and src/bm3d.zig that talks to
// VapourSynth: argument maps in, filter nodes out, frames requested in arInitial and handed to
// the engine in arAllFramesReady.
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vs4_abi.h"
#include "vszipcu_bm3d.h"
#include "vszipcu_vs4.h"

namespace {

enum Kind { kBm3d = 0, kBm3dV2 = 1, kVAggregate = 2 };

struct Instance {
    vsz_filter *f = nullptr;
    VSNode *node[2] = {nullptr, nullptr};  // inputs in the engine's `which` order
    VSVideoInfo vi{};                      // output format
    vsz_output_info info{};
    int props_from = 0;                    // input whose frame n supplies the frame props
};

vsz_clip_info clip_info(const VSVideoInfo *vi) {
    vsz_clip_info c{};
    c.width = vi->width;
    c.height = vi->height;
    c.num_frames = vi->numFrames;
    c.color_family = vi->format.colorFamily;
    c.sample_type = vi->format.sampleType;
    c.bits_per_sample = vi->format.bitsPerSample;
    c.sub_sampling_w = vi->format.subSamplingW;
    c.sub_sampling_h = vi->format.subSamplingH;
    c.num_planes = vi->format.numPlanes;
    return c;
}

int count(const VSAPI *api, const VSMap *in, const char *key) {
    const int n = api->mapNumElements(in, key);
    return n < 0 ? 0 : n;
}

// per-plane int array (src/bm3d.zig:953-963 reads elements 0..2; the fallback is the engine's)
int int_array(const VSAPI *api, const VSMap *in, const char *key, int out[3]) {
    const int n = count(api, in, key) < 3 ? count(api, in, key) : 3;
    int err = 0;
    for (int i = 0; i < n; ++i) out[i] = api->mapGetIntSaturated(in, key, i, &err);
    return n;
}

int int_scalar(const VSAPI *api, const VSMap *in, const char *key, int *out) {
    if (count(api, in, key) == 0) return 0;
    int err = 0;
    *out = api->mapGetIntSaturated(in, key, 0, &err);
    return 1;
}

int str_array(const VSAPI *api, const VSMap *in, const char *key, std::string store[3], const char *out[3]) {
    const int n = count(api, in, key) < 3 ? count(api, in, key) : 3;
    int err = 0;
    for (int i = 0; i < n; ++i) {
        const char *p = api->mapGetData(in, key, i, &err);
        const int sz = api->mapGetDataSize(in, key, i, &err);
        store[i].assign(p ? p : "", p && sz > 0 ? size_t(sz) : 0);
        out[i] = store[i].c_str();
    }
    return n;
}

struct FrameSet {
    const VSAPI *api = nullptr;
    int lo[2] = {0, 0};
    std::vector<const VSFrame *> f[2];

    const VSFrame *get(int which, int n) const {
        if (which < 0 || which > 1) return nullptr;
        const int k = n - lo[which];
        return k >= 0 && k < int(f[which].size()) ? f[which][k] : nullptr;
    }
    ~FrameSet() {
        for (auto &v : f)
            for (const VSFrame *fr : v)
                if (fr) api->freeFrame(fr);
    }
};

const void *get_plane(void *user, int which, int n, int plane, ptrdiff_t *stride) {
    const FrameSet *fs = static_cast<const FrameSet *>(user);
    const VSFrame *fr = fs->get(which, n);
    if (!fr) return nullptr;
    *stride = fs->api->getStride(fr, plane);
    return fs->api->getReadPtr(fr, plane);
}

const VSFrame *get_frame(int n, int reason, void *inst, void ** /*frameData*/, VSFrameContext *ctx, VSCore *core,
                         const VSAPI *api) {
    Instance *d = static_cast<Instance *>(inst);
    if (reason == arInitial) {
        for (int w = 0; w < 2; ++w) {
            if (!d->node[w]) continue;
            int lo = 0, hi = -1;
            vsz_filter_frame_range(d->f, w, n, &lo, &hi);
            for (int k = lo; k <= hi; ++k) api->requestFrameFilter(k, d->node[w], ctx);
        }
        return nullptr;
    }
    if (reason != arAllFramesReady) return nullptr;

    FrameSet fs;
    fs.api = api;
    for (int w = 0; w < 2; ++w) {
        if (!d->node[w]) continue;
        int lo = 0, hi = -1;
        vsz_filter_frame_range(d->f, w, n, &lo, &hi);
        fs.lo[w] = lo;
        for (int k = lo; k <= hi; ++k) fs.f[w].push_back(api->getFrameFilter(k, d->node[w], ctx));
    }
    const VSFrame *centre = fs.get(d->props_from, n);
    VSFrame *dst = api->newVideoFrame(&d->vi.format, d->vi.width, d->vi.height, centre, core);
    if (!dst) {
        api->setFilterError("BM3D: process frame failed.", ctx);
        return nullptr;
    }
    void *ptr[3] = {nullptr, nullptr, nullptr};
    ptrdiff_t stride[3] = {0, 0, 0};
    const int np = d->vi.format.numPlanes < 3 ? d->vi.format.numPlanes : 3;
    for (int p = 0; p < np; ++p) {
        ptr[p] = api->getWritePtr(dst, p);
        stride[p] = api->getStride(dst, p);
    }
    char err[512] = {0};
    const int rc = vsz_filter_get_frame(d->f, n, &get_plane, &fs, ptr, stride, err, sizeof err);
    if (rc != VSZ_OK) {
        api->setFilterError(err, ctx);
        api->freeFrame(dst);
        return nullptr;
    }
    if (d->info.stacked) {
        // src/bm3d.zig:716-720
        VSMap *props = api->getFramePropertiesRW(dst);
        api->mapSetInt(props, "BM3D_V_radius", d->info.radius, maReplace);
        const int64_t proc[3] = {d->info.process[0] != 0, d->info.process[1] != 0, d->info.process[2] != 0};
        api->mapSetIntArray(props, "BM3D_V_process", proc, 3);
    }
    return dst;
}

void free_instance(void *inst, VSCore * /*core*/, const VSAPI *api) {
    Instance *d = static_cast<Instance *>(inst);
    vsz_filter_free(d->f);
    for (VSNode *n : d->node)
        if (n) api->freeNode(n);
    delete d;
}

// BM3Dv2 as the reference builds it (src/bm3d.zig:1432-1501): the input itself when no plane is
// processed; otherwise the registered BM3D, followed at radius > 0 by the registered VAggregate over
// BM3D's stacked output, as two nodes of the graph. Returns false for fast_fused at radius > 0,
// which is one filter of this plugin (the fused accumulator chain).
bool compose_v2(const VSMap *in, VSMap *out, VSCore *core, const VSAPI *api) {
    int err = 0;
    bool proc[3] = {true, true, true};
    const int ns = count(api, in, "sigma");
    for (int i = 0; i < 3; ++i) {
        if (i < ns) proc[i] = !(float(api->mapGetFloat(in, "sigma", i, &err)) < FLT_EPSILON);
        else if (i > 0) proc[i] = proc[i - 1];
    }
    VSNode *src = api->mapGetNode(in, "clip", 0, &err);
    if (!src) {
        api->mapSetError(out, "BM3D: missing clip.");
        return true;
    }
    const int np = api->getVideoInfo(src)->format.numPlanes;
    bool skip = true;
    for (int i = 0; i < np && i < 3; ++i)
        if (proc[i]) skip = false;
    if (skip) {
        api->mapConsumeNode(out, "clip", src, maReplace);
        return true;
    }
    VSPlugin *plugin = api->getPluginByID("com.julek.vszipcu", core);
    if (!plugin) {
        api->freeNode(src);
        api->mapSetError(out, "BM3Dv2: could not find the vszipcu plugin.");
        return true;
    }
    int radius = 0, fused = 0;
    int_scalar(api, in, "radius", &radius);
    int_scalar(api, in, "fast_fused", &fused);
    if (radius > 0 && fused != 0) {
        api->freeNode(src);
        return false;
    }
    VSMap *bret = api->invoke(plugin, "BM3D", in);
    if (const char *e = api->mapGetError(bret)) {
        api->mapSetError(out, e);
        api->freeMap(bret);
        api->freeNode(src);
        return true;
    }
    VSNode *bnode = api->mapGetNode(bret, "clip", 0, &err);
    api->freeMap(bret);
    if (radius == 0) {
        api->mapConsumeNode(out, "clip", bnode, maReplace);
        api->freeNode(src);
        return true;
    }
    VSMap *aargs = api->createMap();
    api->mapConsumeNode(aargs, "clip", bnode, maReplace);
    api->mapConsumeNode(aargs, "src", src, maReplace);
    for (int i = 0; i < 3; ++i)
        if (proc[i]) api->mapSetInt(aargs, "planes", i, maAppend);
    VSMap *aret = api->invoke(plugin, "VAggregate", aargs);
    api->freeMap(aargs);
    if (const char *e = api->mapGetError(aret)) {
        api->mapSetError(out, e);
    } else {
        api->mapConsumeNode(out, "clip", api->mapGetNode(aret, "clip", 0, &err), maReplace);
    }
    api->freeMap(aret);
    return true;
}

void create(const VSMap *in, VSMap *out, void *user, VSCore *core, const VSAPI *api) {
    const int kind = int(reinterpret_cast<intptr_t>(user));
    if (kind == kBm3dV2 && compose_v2(in, out, core, api)) return;
    int err = 0;
    Instance *d = new (std::nothrow) Instance;
    if (!d) {
        api->mapSetError(out, "BM3D: out of memory.");
        return;
    }
    auto drop = [&](const char *msg) {
        if (msg) api->mapSetError(out, msg);
        if (d->f) vsz_filter_free(d->f);
        for (VSNode *n : d->node)
            if (n) api->freeNode(n);
        delete d;
    };
    char msg[512] = {0};
    int rc;
    int dep_pattern[2] = {rpGeneral, rpGeneral};
    const char *name = "BM3D";

    if (kind == kVAggregate) {
        // clip:vnode;src:vnode;planes:int[]; (src/plugin.zig:91-98)
        d->node[0] = api->mapGetNode(in, "clip", 0, &err);
        d->node[1] = api->mapGetNode(in, "src", 0, &err);
        if (!d->node[0] || !d->node[1]) return drop("VAggregate: missing clip or src.");
        const vsz_clip_info ci = clip_info(api->getVideoInfo(d->node[0]));
        const vsz_clip_info si = clip_info(api->getVideoInfo(d->node[1]));
        std::vector<int> planes(size_t(count(api, in, "planes")));
        for (size_t i = 0; i < planes.size(); ++i) planes[i] = api->mapGetIntSaturated(in, "planes", int(i), &err);
        rc = vsz_vaggregate_create(&ci, &si, planes.data(), int(planes.size()), &d->f, msg, sizeof msg);
        if (rc != VSZ_OK) return drop(msg);
        d->vi = *api->getVideoInfo(d->node[1]);
        d->props_from = 1;
        dep_pattern[1] = rpStrictSpatial;  // src/bm3d.zig:1425-1428
        name = "VAggregate";
    } else {
        vsz_bm3d_args a{};
        std::string strs[3][3];
        d->node[0] = api->mapGetNode(in, "clip", 0, &err);
        if (!d->node[0]) return drop("BM3D: missing clip.");
        const VSVideoInfo *vi = api->getVideoInfo(d->node[0]);
        a.clip = clip_info(vi);
        if (count(api, in, "ref") > 0) {
            d->node[1] = api->mapGetNode(in, "ref", 0, &err);
            a.has_ref = 1;
            a.ref = clip_info(api->getVideoInfo(d->node[1]));
        }
        a.n_sigma = count(api, in, "sigma") < 3 ? count(api, in, "sigma") : 3;
        for (int i = 0; i < a.n_sigma; ++i) a.sigma[i] = float(api->mapGetFloat(in, "sigma", i, &err));
        a.n_block_step = int_array(api, in, "block_step", a.block_step);
        a.n_bm_range = int_array(api, in, "bm_range", a.bm_range);
        a.has_radius = int_scalar(api, in, "radius", &a.radius);
        a.n_ps_num = int_array(api, in, "ps_num", a.ps_num);
        a.n_ps_range = int_array(api, in, "ps_range", a.ps_range);
        a.has_chroma = int_scalar(api, in, "chroma", &a.chroma);
        a.has_device_id = int_scalar(api, in, "device_id", &a.device_id);
        a.has_num_streams = int_scalar(api, in, "num_streams", &a.num_streams);
        a.has_extractor_exp = int_scalar(api, in, "extractor_exp", &a.extractor_exp);
        a.n_bm_error_s = str_array(api, in, "bm_error_s", strs[0], a.bm_error_s);
        a.n_transform_2d_s = str_array(api, in, "transform_2d_s", strs[1], a.transform_2d_s);
        a.n_transform_1d_s = str_array(api, in, "transform_1d_s", strs[2], a.transform_1d_s);
        a.has_zero_init = int_scalar(api, in, "zero_init", &a.zero_init);
        a.has_fast_fused = int_scalar(api, in, "fast_fused", &a.fast_fused);

        rc = kind == kBm3dV2 ? vsz_bm3dv2_create(&a, &d->f, msg, sizeof msg) : vsz_bm3d_create(&a, &d->f, msg, sizeof msg);
        if (rc != VSZ_OK) return drop(msg);
        vsz_filter_output_info(d->f, &d->info);
        if (d->info.passthrough) {
            // BM3Dv2 with no processed plane returns its input (src/bm3d.zig:1452-1456)
            api->mapSetNode(out, "clip", d->node[0], maReplace);
            return drop(nullptr);
        }
        d->vi = *vi;
        d->vi.height = d->info.vi.height;  // stacked: H * 2 * (2r+1)
        const int radius = a.has_radius ? a.radius : 0;
        // src/bm3d.zig:1258-1266: StrictSpatial at radius 0, General otherwise
        dep_pattern[0] = dep_pattern[1] = radius > 0 ? rpGeneral : rpStrictSpatial;
        name = kind == kBm3dV2 && radius > 0 ? "BM3Dv2" : "BM3D";
    }

    VSFilterDependency deps[2];
    int nd = 0;
    for (int w = 0; w < 2; ++w)
        if (d->node[w]) deps[nd++] = VSFilterDependency{d->node[w], dep_pattern[w]};
    api->createVideoFilter(out, name, &d->vi, &get_frame, &free_instance, fmParallel, deps, nd, d, core);
}

}  // namespace

extern "C" __attribute__((visibility("default"))) void VapourSynthPluginInit2(VSPlugin *plugin,
                                                                             const VSPLUGINAPI *vspapi) {
    // src/plugin.zig:15-23
    vspapi->configPlugin("com.julek.vszipcu", "vszipcu", "VapourSynth Zig Image Process CUDA",
                         VSZ_VS_MAKE_VERSION(0, 1), VSZ_VS_API_VERSION, 0, plugin);
    // src/plugin.zig:80-98, argument strings verbatim
    static const char bm3d_sig[] =
        "clip:vnode;ref:vnode:opt;sigma:float[]:opt;block_step:int[]:opt;"
        "bm_range:int[]:opt;radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;"
        "chroma:int:opt;device_id:int:opt;num_streams:int:opt;"
        "extractor_exp:int:opt;bm_error_s:data[]:opt;transform_2d_s:data[]:opt;"
        "transform_1d_s:data[]:opt;zero_init:int:opt;fast_fused:int:opt;";
    vspapi->registerFunction("BM3D", bm3d_sig, "clip:vnode;", &create, reinterpret_cast<void *>(intptr_t(kBm3d)),
                             plugin);
    vspapi->registerFunction("BM3Dv2", bm3d_sig, "clip:vnode;", &create, reinterpret_cast<void *>(intptr_t(kBm3dV2)),
                             plugin);
    vspapi->registerFunction("VAggregate", "clip:vnode;src:vnode;planes:int[];", "clip:vnode;", &create,
                             reinterpret_cast<void *>(intptr_t(kVAggregate)), plugin);
}
