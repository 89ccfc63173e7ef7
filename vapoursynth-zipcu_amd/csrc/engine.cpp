// Host side of the MI355X BM3D engine: the BM3D / BM3Dv2 / VAggregate filter instances of
// include/vszipcu_bm3d.h. It restates the reference's filter host (src/bm3d.zig) on top of the
// HIP runtime and the kernel layer in bm3d_kernels.hip:
//   argument rules, validation and error strings ... src/bm3d.zig:953-1271, 1372-1501
//   plane entries and buffer layout ................. src/bm3d.zig:1094-1188
//   frame windows, stacked output, props ............ src/bm3d.zig:635-737
//   upload / launch / download ...................... src/bm3d.zig:274-478
//   fused accumulator chain ......................... src/bm3d.zig:481-633, 898-949
//   stream-slot pool ................................ src/pool.zig
//   device slot caches (lease table) ................ src/framecache.zig
//   CPU VAggregate .................................. src/bm3d.zig:1284-1363

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <new>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"
#include "host_copy.h"
#include "slot_cache.h"
#include "vszipcu_bm3d.h"

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

int vsz_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" const char *vsz_last_error(void) { return g_last_error.c_str(); }
extern "C" int vsz_version(void) { return 1; }

extern "C" int vsz_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

namespace {

constexpr int MAX_RADIUS = 16;
constexpr int MAX_TW = 2 * MAX_RADIUS + 1;

int set_err(char *err, size_t errlen, int code, const std::string &msg) {
    g_last_error = msg;
    if (err && errlen) {
        std::snprintf(err, errlen, "%s", msg.c_str());
    }
    return code;
}

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            std::fprintf(stderr, "vszipcu: HIP error %s at %s:%d (%s)\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__, #expr);                                                               \
            return e_ == hipErrorOutOfMemory ? VSZ_ENOMEM : VSZ_EDEVICE;                                 \
        }                                                                                                \
    } while (0)

bool eq_lower(const char *s, const char *lit) {
    if (!s) return false;
    size_t n = std::strlen(s);
    if (n != std::strlen(lit)) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[i]) != lit[i]) return false;
    return true;
}

// src/bm3d.zig:72-110
int parse_metric(const char *s) {
    static const char *names[] = {"ssd", "sad", "zssd", "zsad", "ssd/norm"};
    for (int i = 0; i < 5; ++i)
        if (eq_lower(s, names[i])) return i;
    return -1;
}
int parse_transform(const char *s) {
    static const char *names[] = {"dct", "haar", "wht", "bior1.5"};
    for (int i = 0; i < 4; ++i)
        if (eq_lower(s, names[i])) return i;
    return -1;
}

float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Device row stride: rows padded to 64 floats (256 B) for aligned, coalesced rows.
int device_stride(int w) { return (w + 63) & ~63; }

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------------------------------------
// Device slot cache: device buffers (and their fill events) managed by a LeaseTable (slot_cache.h).
// ---------------------------------------------------------------------------------------------
class SlotCache {
  public:
    struct Buffer {
        void *buf = nullptr;
        hipEvent_t ev = nullptr;  // recorded after the buffer was filled
    };
    using Lease = LeaseTable::Lease;

    ~SlotCache() { clear(); }

    // Allocates up to `want` buffers of `bytes`; on device OOM keeps what it has if that is at
    // least `min_bufs` (src/bm3d.zig:865-885, 923-935).
    int init(size_t bytes, size_t want, size_t min_bufs) {
        clear();
        for (size_t i = 0; i < want; ++i) {
            Buffer b;
            const hipError_t e = hipMalloc(&b.buf, std::max<size_t>(bytes, 256));
            if (e != hipSuccess) {
                (void)hipGetLastError();
                if (e == hipErrorOutOfMemory && bufs_.size() >= min_bufs) break;
                return e == hipErrorOutOfMemory ? VSZ_ENOMEM : VSZ_EDEVICE;
            }
            if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
                (void)hipFree(b.buf);
                return VSZ_EDEVICE;
            }
            bufs_.push_back(b);
        }
        table_.reset(bufs_.size());
        return VSZ_OK;
    }

    void clear() {
        for (auto &b : bufs_) {
            if (b.buf) (void)hipFree(b.buf);
            if (b.ev) (void)hipEventDestroy(b.ev);
        }
        bufs_.clear();
        table_.reset(0);
    }

    Lease lease(const int64_t *keys, int n) { return table_.lease(keys, n); }
    Buffer &buffer(const Lease &l, int i) { return bufs_[l.slot(i)]; }

  private:
    std::vector<Buffer> bufs_;
    LeaseTable table_;
};

// ---------------------------------------------------------------------------------------------
// per-frame stream slots (src/pool.zig) with their buffers (src/bm3d.zig:188-241)
// ---------------------------------------------------------------------------------------------
struct StreamSlot {
    hipStream_t stream = nullptr;
    float *d_src = nullptr;   // sum_src elements
    float *d_out = nullptr;   // dst (normal output) or res (stacked output)
    char *d_work = nullptr;   // kernel workspaces of all entries
    float *h_stage = nullptr; // pinned: upload region + download region
    float *d_down = nullptr;  // zero-copy output: device address of the download region

    void destroy() {
        if (stream) (void)hipStreamDestroy(stream);
        if (d_src) (void)hipFree(d_src);
        if (d_out) (void)hipFree(d_out);
        if (d_work) (void)hipFree(d_work);
        if (h_stage) (void)hipHostFree(h_stage);
        *this = StreamSlot{};
    }
};

class StreamPool {
  public:
    std::vector<std::unique_ptr<StreamSlot>> all;

    ~StreamPool() {
        for (auto &s : all) s->destroy();
    }

    StreamSlot *acquire() {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !free_.empty(); });
        StreamSlot *s = free_.back();
        free_.pop_back();
        return s;
    }

    void release(StreamSlot *s) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            free_.push_back(s);
        }
        cv_.notify_one();
    }

    void add(std::unique_ptr<StreamSlot> s) {
        free_.push_back(s.get());
        all.push_back(std::move(s));
    }

  private:
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<StreamSlot *> free_;
};

struct SlotGuard {
    StreamPool &pool;
    StreamSlot *s;
    explicit SlotGuard(StreamPool &p) : pool(p), s(p.acquire()) {}
    ~SlotGuard() { pool.release(s); }
};

// RAII: make `dev` current on this thread for the duration of a call.
struct DeviceScope {
    int prev = -1;
    bool ok = false;
    explicit DeviceScope(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Plane-0/subsampled plane geometry of a clip.
int plane_w(const vsz_clip_info &c, int p) { return p == 0 ? c.width : c.width >> c.sub_sampling_w; }
int plane_h(const vsz_clip_info &c, int p) { return p == 0 ? c.height : c.height >> c.sub_sampling_h; }

using vsz_host::copy_plane;

}  // namespace

// ---------------------------------------------------------------------------------------------
// filter base
// ---------------------------------------------------------------------------------------------
struct vsz_filter {
    virtual ~vsz_filter() = default;
    vsz_output_info out{};
    virtual int get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) = 0;
    virtual const char *frame_error() const { return "BM3D: process frame failed."; }
    // Inclusive range of frames of input `which` that get_frame(n) may pull; lo > hi if none.
    virtual void frame_range(int which, int n, int &lo, int &hi) const = 0;
};

namespace {

// ---------------------------------------------------------------------------------------------
// BM3D instance (src/bm3d.zig:117-186 Entry / Data)
// ---------------------------------------------------------------------------------------------
struct Entry {
    vsz_bm3d_kparams kp{};
    int n_planes = 1;
    int planes[3] = {0, 0, 0};
    int w = 0, h = 0, S = 0;       // plane geometry (device stride S)
    size_t plane_extent = 0;       // S * h elements
    size_t src_off = 0, src_elems = 0;
    size_t res_off = 0, res_elems = 0;
    size_t dst_off = 0, dst_elems = 0;
    size_t cache_off = 0;
    size_t work_off = 0, work_bytes = 0;
};

class Bm3d final : public vsz_filter {
  public:
    vsz_clip_info vi{};
    int radius = 0, tw = 1;
    bool final_pass = false, chroma = false, fused = false, zero_init = true;
    bool process[3] = {false, false, false};
    int device = 0;
    int num_streams = 4;
    int n_clips = 1;
    std::vector<Entry> entries;
    size_t sum_src = 0, sum_res = 0, sum_dst = 0, sum_work = 0, cache_elems = 0;
    bool scache = false;
    StreamPool pool;
    SlotCache cache;  // uploaded source frames, key = (clip << 40) | frame
    SlotCache acc;    // fused: FP32 accumulator stacks per centre frame
    std::string v2_error = "BM3D: process frame failed.";

    ~Bm3d() override {
        DeviceScope ds(device);
        for (auto &s : pool.all) {
            if (s->stream) (void)hipStreamSynchronize(s->stream);
        }
        cache.clear();
        acc.clear();
    }

    bool normal_out() const { return radius == 0 || fused; }
    // Zero-copy output (src/bm3d.zig:1203-1206): with two or more streams the output kernels
    // write the normal (dst) frame straight into mapped pinned memory, so no D2H copy runs.
    bool zc_dst() const { return normal_out() && num_streams >= 2; }
    size_t down_off() const { return std::max(sum_src, cache_elems); }
    float *out_base(StreamSlot &s) const { return zc_dst() ? s.d_down : s.d_out; }
    const char *frame_error() const override { return v2_error.c_str(); }
    // src/bm3d.zig:643-651: n-r..n+r (n-2r..n+2r when fused) of clip and ref
    void frame_range(int which, int n, int &lo, int &hi) const override {
        const int req = fused ? 2 * radius : radius;
        lo = 0;
        hi = -1;
        if (which == 0 || (which == 1 && n_clips == 2)) {
            lo = std::max(n - req, 0);
            hi = std::min(n + req, vi.num_frames - 1);
        }
    }

    int init_device(char *err, size_t errlen);
    int get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) override;

  private:
    int upload_window(StreamSlot &s, const int *idx, int count, vsz_get_plane_fn get, void *user,
                      SlotCache::Lease &lease);
    int launch_entries(StreamSlot &s, float *out_base, int out_kind);
    int download(StreamSlot &s, const float *d_base, bool stacked, void *const dst[3], const ptrdiff_t dst_stride[3]);
    int process_normal(StreamSlot &s, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                       const ptrdiff_t dst_stride[3]);
    int process_fused(StreamSlot &s, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                      const ptrdiff_t dst_stride[3]);
};

// ---- create (src/bm3d.zig:969-1271) ---------------------------------------------------------
int create_bm3d(const vsz_bm3d_args *a, bool fused_req, std::unique_ptr<Bm3d> &out, char *err, size_t errlen) {
    auto fail = [&](const char *msg) { return set_err(err, errlen, VSZ_EINVAL, msg); };
    if (!a) return fail("BM3D: missing arguments.");
    auto d = std::make_unique<Bm3d>();
    const vsz_clip_info &vi = a->clip;
    d->vi = vi;

    if (vi.width <= 0 || vi.height <= 0 || vi.sample_type != 1 || vi.bits_per_sample != 32)
        return fail("BM3D: only constant format 32 bit float input supported.");

    if (a->has_ref) {
        const vsz_clip_info &r = a->ref;
        if (r.color_family != vi.color_family || r.sample_type != vi.sample_type ||
            r.bits_per_sample != vi.bits_per_sample || r.sub_sampling_w != vi.sub_sampling_w ||
            r.sub_sampling_h != vi.sub_sampling_h)
            return fail("BM3D: \"ref\" must be of the same format as \"clip\".");
        if (r.width != vi.width || r.height != vi.height) return fail("BM3D: \"ref\" must be of the same dimensions as \"clip\".");
        if (r.num_frames != vi.num_frames) return fail("BM3D: \"ref\" must be of the same number of frames as \"clip\".");
        d->final_pass = true;
    }

    // per-plane rule: element i absent -> element i-1, element 0 absent -> default
    float sigma[3];
    for (int i = 0; i < 3; ++i) sigma[i] = i < a->n_sigma ? a->sigma[i] : (i == 0 ? 3.0f : sigma[i - 1]);
    for (float s : sigma)
        if (!(s >= 0)) return fail("BM3D: \"sigma\" must be non-negative.");
    for (int i = 0; i < 3; ++i) d->process[i] = !(sigma[i] < FLT_EPSILON);
    // f32 products (3/4)/255*64*2.7 and (3/4)/255*64*1.0, bit-exact (src/bm3d.zig:1019-1024)
    const float factor = d->final_pass ? bits_to_float(0x3e40c0c1u) : bits_to_float(0x3f021bb6u);
    for (float &s : sigma) s *= factor;

    auto per_plane_i = [&](int n, const int *v, int def, int out3[3]) {
        for (int i = 0; i < 3; ++i) out3[i] = i < n ? v[i] : (i == 0 ? def : out3[i - 1]);
    };
    int block_step[3], bm_range[3], ps_num[3], ps_range[3];
    per_plane_i(a->n_block_step, a->block_step, 8, block_step);
    for (int v : block_step)
        if (v <= 0 || v > 8) return fail("BM3D: \"block_step\" must be in range [1, 8].");
    per_plane_i(a->n_bm_range, a->bm_range, 9, bm_range);
    for (int v : bm_range)
        if (v <= 0) return fail("BM3D: \"bm_range\" must be positive.");
    per_plane_i(a->n_ps_num, a->ps_num, 2, ps_num);
    for (int v : ps_num)
        if (v <= 0 || v > 8) return fail("BM3D: \"ps_num\" must be in range [1, 8].");
    per_plane_i(a->n_ps_range, a->ps_range, 4, ps_range);
    for (int v : ps_range)
        if (v <= 0) return fail("BM3D: \"ps_range\" must be positive.");

    d->radius = a->has_radius ? a->radius : 0;
    if (d->radius < 0) return fail("BM3D: \"radius\" must be non-negative.");
    if (d->radius > MAX_RADIUS) return fail("BM3D: \"radius\" must be <= 16.");
    d->tw = 2 * d->radius + 1;
    d->fused = fused_req && d->radius > 0;

    d->chroma = a->has_chroma && a->chroma != 0;
    if (d->chroma && (vi.color_family != 3 || vi.sub_sampling_w != 0 || vi.sub_sampling_h != 0))
        return fail("BM3D: clip format must be YUV444 when \"chroma\" is true.");

    d->zero_init = !a->has_zero_init || a->zero_init != 0;
    const int extractor_exp = a->has_extractor_exp ? a->extractor_exp : 0;

    int metric[3] = {0, 0, 0}, t2d[3] = {0, 0, 0}, t1d[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        if (i < a->n_bm_error_s) {
            metric[i] = parse_metric(a->bm_error_s[i]);
            if (metric[i] < 0) return fail("BM3D: invalid \"bm_error_s\".");
        } else if (i > 0) metric[i] = metric[i - 1];
        if (i < a->n_transform_2d_s) {
            t2d[i] = parse_transform(a->transform_2d_s[i]);
            if (t2d[i] < 0) return fail("BM3D: invalid \"transform_2d_s\".");
        } else if (i > 0) t2d[i] = t2d[i - 1];
        if (i < a->n_transform_1d_s) {
            t1d[i] = parse_transform(a->transform_1d_s[i]);
            if (t1d[i] < 0) return fail("BM3D: invalid \"transform_1d_s\".");
        } else if (i > 0) t1d[i] = t1d[i - 1];
    }

    d->device = a->has_device_id ? a->device_id : 0;
    if (d->device < 0) return fail("BM3D: invalid device ID.");
    if (a->has_num_streams && (a->num_streams < 1 || a->num_streams > 32)) return fail("BM3D: num_streams must be 1..32.");
    d->num_streams = a->has_num_streams ? a->num_streams : 4;

    const int num_planes = vi.num_planes;
    bool any = false;
    for (int i = 0; i < num_planes && i < 3; ++i) any |= d->process[i];
    if (!any) return fail("BM3D: all planes have sigma < FLT_EPSILON (nothing to process).");

    // plane entries (src/bm3d.zig:1094-1188)
    d->n_clips = d->final_pass ? 2 : 1;
    auto make_kp = [&](int p, int w, int h, int S, const float sig[3], unsigned pm) {
        vsz_bm3d_kparams kp{};
        kp.width = w;
        kp.height = h;
        kp.stride = S;
        kp.block_step = block_step[p];
        kp.bm_range = bm_range[p];
        kp.radius = d->radius;
        kp.ps_num = ps_num[p];
        kp.ps_range = ps_range[p];
        kp.chroma = d->chroma ? 1 : 0;
        kp.final_pass = d->final_pass ? 1 : 0;
        for (int i = 0; i < 3; ++i) kp.sigma[i] = sig[i];
        kp.extractor_exp = extractor_exp;
        kp.metric = metric[p];
        kp.t2d = t2d[p];
        kp.t1d = t1d[p];
        kp.proc_mask = pm;
        return kp;
    };
    size_t sum_src = 0, sum_res = 0, sum_dst = 0, cache_elems = 0;
    if (d->chroma) {
        Entry e;
        e.w = vi.width;
        e.h = vi.height;
        e.S = device_stride(e.w);
        e.plane_extent = size_t(e.S) * e.h;
        unsigned pm = 0;
        for (int i = 0; i < 3; ++i) pm |= unsigned(d->process[i]) << i;
        e.kp = make_kp(0, e.w, e.h, e.S, sigma, pm);
        e.n_planes = 3;
        e.planes[0] = 0;
        e.planes[1] = 1;
        e.planes[2] = 2;
        e.src_elems = size_t(d->n_clips) * 3 * d->tw * e.plane_extent;
        e.res_elems = size_t(3) * d->tw * 2 * e.plane_extent;
        e.dst_elems = d->normal_out() ? 3 * e.plane_extent : 0;
        sum_src = e.src_elems;
        sum_res = e.res_elems;
        sum_dst = e.dst_elems;
        cache_elems = 3 * e.plane_extent;
        d->entries.push_back(e);
    } else {
        for (int p = 0; p < num_planes && p < 3; ++p) {
            if (!d->process[p]) continue;
            Entry e;
            e.w = plane_w(vi, p);
            e.h = plane_h(vi, p);
            if (e.w < 8 || e.h < 8) return fail("BM3D: every processed plane must be at least 8x8.");
            e.S = device_stride(e.w);
            e.plane_extent = size_t(e.S) * e.h;
            const float sig[3] = {sigma[p], sigma[p], sigma[p]};
            e.kp = make_kp(p, e.w, e.h, e.S, sig, 1u);
            e.n_planes = 1;
            e.planes[0] = p;
            e.cache_off = cache_elems;
            e.src_off = sum_src;
            e.src_elems = size_t(d->n_clips) * d->tw * e.plane_extent;
            e.res_off = sum_res;
            e.res_elems = size_t(d->tw) * 2 * e.plane_extent;
            e.dst_off = sum_dst;
            e.dst_elems = d->normal_out() ? e.plane_extent : 0;
            sum_src += e.src_elems;
            sum_res += e.res_elems;
            sum_dst += e.dst_elems;
            cache_elems += e.plane_extent;
            d->entries.push_back(e);
        }
    }
    if (d->chroma && (vi.width < 8 || vi.height < 8)) return fail("BM3D: every processed plane must be at least 8x8.");
    for (auto &e : d->entries) {
        // kernel-layer limits: 16-bit match coordinates, 32-bit byte offsets into a plane stack
        if (e.w > 65535 || e.h > 65535 || size_t(d->tw) * e.plane_extent * 4 >= (size_t(1) << 31))
            return fail("BM3D: plane too large for this build (width/height <= 65535, window < 2 GiB).");
        if (e.src_elems >= (size_t(1) << 31) || e.res_elems >= (size_t(1) << 31))
            return fail("BM3D: frame/radius too large (a device region exceeds 2^31 samples).");
        if (size_t(e.w) * e.h * d->tw * 2 * 4 >= (size_t(1) << 32))
            return fail("BM3D: stacked output plane too large (exceeds 4 GiB).");
        e.work_off = d->sum_work;
        e.work_bytes = align_up(vsz_bm3d_workspace_size(&e.kp), 256);
        d->sum_work += e.work_bytes;
    }
    d->sum_src = sum_src;
    d->sum_res = sum_res;
    d->sum_dst = sum_dst;
    d->cache_elems = cache_elems;
    d->scache = d->radius > 0;

    d->out.vi = vi;
    d->out.radius = d->radius;
    for (int i = 0; i < 3; ++i) d->out.process[i] = d->process[i];
    if (d->radius > 0 && !d->fused) {
        const long long mult = 2LL * (2 * d->radius + 1);
        if ((long long)vi.height * mult > INT_MAX) return fail("BM3D: clip too tall for the stacked output.");
        d->out.vi.height = int(vi.height * mult);
        d->out.stacked = 1;
    }

    const int rc = d->init_device(err, errlen);
    if (rc) return rc;
    out = std::move(d);
    return VSZ_OK;
}

// ---- device resources (src/bm3d.zig:772-951) ------------------------------------------------
int Bm3d::init_device(char *err, size_t errlen) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return set_err(err, errlen, VSZ_EDEVICE, "BM3D: HIP initialization failed (no usable GPU).");
    }
    if (device >= ndev) return set_err(err, errlen, VSZ_EINVAL, "BM3D: invalid device ID.");
    DeviceScope ds(device);
    if (!ds.ok) return set_err(err, errlen, VSZ_EDEVICE, "BM3D: HIP initialization failed (see log).");

    auto oom = [&]() { return set_err(err, errlen, VSZ_ENOMEM, "BM3D: out of device memory."); };
    if (scache) {
        const size_t win_frames = fused ? size_t(2 * tw - 1) : size_t(tw);
        const size_t min_slots = size_t(n_clips) * win_frames;
        const size_t want = min_slots + size_t(num_streams) + 2 * size_t(radius);
        if (cache.init(cache_elems * 4, want, min_slots)) return oom();
    }
    const size_t out_elems = normal_out() ? sum_dst : sum_res;
    for (int i = 0; i < num_streams; ++i) {
        auto s = std::make_unique<StreamSlot>();
        bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipMalloc(&s->d_src, std::max<size_t>(sum_src, 64) * 4) == hipSuccess;
        ok = ok && hipMalloc(&s->d_out, std::max<size_t>(zc_dst() ? 0 : out_elems, 64) * 4) == hipSuccess;
        ok = ok && hipMalloc(&s->d_work, std::max<size_t>(sum_work, 256)) == hipSuccess;
        ok = ok && hipHostMalloc(&s->h_stage, (down_off() + out_elems + 64) * 4, hipHostMallocMapped) == hipSuccess;
        if (ok && zc_dst()) {
            void *dp = nullptr;
            ok = hipHostGetDevicePointer(&dp, s->h_stage, 0) == hipSuccess && dp != nullptr;
            if (ok) s->d_down = static_cast<float *>(dp) + down_off();
        }
        if (!ok) {
            (void)hipGetLastError();
            s->destroy();
            if (!pool.all.empty()) break;  // fewer streams on OOM (src/pool.zig:35-41)
            return oom();
        }
        pool.add(std::move(s));
    }
    if (fused) {
        // accumulator cache sized against free memory minus 256 MiB (src/bm3d.zig:898-949)
        const size_t min_slots = size_t(tw);
        const size_t want = size_t(tw) + size_t(num_streams) + 2 * size_t(radius);
        const size_t slot_bytes = sum_res * 4;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return oom();
        // VSZIPCU_FREE_MEM_CAP_MIB caps the free memory this sizing sees (a testing aid that makes
        // the out-of-memory path below reachable on a large GPU)
        if (const char *cap_env = std::getenv("VSZIPCU_FREE_MEM_CAP_MIB")) {
            const long long cap_mib = std::atoll(cap_env);
            if (cap_mib >= 0) free_b = std::min(free_b, size_t(cap_mib) << 20);
        }
        const size_t margin = size_t(256) << 20;
        const size_t budget = free_b > margin ? free_b - margin : 0;
        const size_t cap = std::min(budget / std::max<size_t>(slot_bytes, 1), want);
        if (cap < min_slots || acc.init(slot_bytes, cap, min_slots)) {
            // the reference formats into a 320-byte buffer and falls back to a fixed text when the
            // message does not fit (src/bm3d.zig:1236-1246)
            char buf[320];
            const int len = std::snprintf(buf, sizeof buf,
                          "BM3Dv2: fast_fused needs at least %zu MiB of accumulator cache (%zu slots x %zu MiB) "
                          "but only %zu MiB of device memory is free. Use fast_fused=False (composed BM3D + "
                          "VAggregate, much less VRAM), or lower num_streams/radius.",
                          (min_slots * slot_bytes) >> 20, min_slots, slot_bytes >> 20, free_b >> 20);
            if (len < 0 || size_t(len) >= sizeof buf)
                return set_err(err, errlen, VSZ_ENOMEM,
                               "BM3Dv2: fast_fused: not enough device memory for the accumulator cache. Use "
                               "fast_fused=False.");
            return set_err(err, errlen, VSZ_ENOMEM, buf);
        }
    }
    return VSZ_OK;
}

// ---- upload (src/bm3d.zig:295-373) -----------------------------------------------------------
// Fills slot.d_src for the window idx[0..tw) of every clip. With the source cache, frames are
// uploaded once into cache slots and gathered device-to-device.
int Bm3d::upload_window(StreamSlot &s, const int *idx, int count, vsz_get_plane_fn get, void *user,
                        SlotCache::Lease &lease) {
    if (!scache) {
        // radius 0: stage every plane of the window; each plane's H2D copy is queued as soon as
        // it is staged, so the DMA of one plane overlaps the host copy of the next
        for (auto &e : entries) {
            for (int c = 0; c < n_clips; ++c)
                for (int i = 0; i < e.n_planes; ++i)
                    for (int t = 0; t < count; ++t) {
                        ptrdiff_t st = 0;
                        const void *src = get(user, n_clips == 2 ? (c == 0 ? 1 : 0) : 0, idx[t], e.planes[i], &st);
                        if (!src) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
                        const size_t off = e.src_off + ((size_t(c) * e.n_planes + i) * tw + t) * e.plane_extent;
                        copy_plane(s.h_stage + off, ptrdiff_t(e.S) * 4, src, st, size_t(e.w) * 4, e.h);
                        HIPCHK(hipMemcpyAsync(s.d_src + off, s.h_stage + off, e.plane_extent * 4, hipMemcpyHostToDevice,
                                              s.stream));
                    }
        }
        return VSZ_OK;
    }
    // source cache: keys (clip << 40) | frame; clip 0 = ref in the final pass
    const int nk = n_clips * count;
    std::vector<int64_t> keys(nk);
    for (int c = 0; c < n_clips; ++c)
        for (int t = 0; t < count; ++t) keys[c * count + t] = (int64_t(c) << 40) | idx[t];
    lease = cache.lease(keys.data(), nk);
    if (!lease.held()) return vsz_fail(VSZ_ENOMEM, "source cache smaller than one frame window");
    for (int k = 0; k < nk; ++k) {
        if (!lease.must_fill(k)) continue;
        const int c = k / count, t = k % count;
        SlotCache::Buffer &slot = cache.buffer(lease, k);
        // stage through pinned memory, then H2D into the slot
        for (auto &e : entries)
            for (int i = 0; i < e.n_planes; ++i) {
                ptrdiff_t st = 0;
                const void *src = get(user, n_clips == 2 ? (c == 0 ? 1 : 0) : 0, idx[t], e.planes[i], &st);
                if (!src) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
                const size_t coff = e.cache_off + size_t(i) * e.plane_extent;
                copy_plane(s.h_stage + coff, ptrdiff_t(e.S) * 4, src, st, size_t(e.w) * 4, e.h);
            }
        HIPCHK(hipMemcpyAsync(slot.buf, s.h_stage, cache_elems * 4, hipMemcpyHostToDevice, s.stream));
        HIPCHK(hipEventRecord(slot.ev, s.stream));
        // the staging buffer is reused by the next frame: wait for this copy
        HIPCHK(hipStreamSynchronize(s.stream));
        lease.filled(k);
    }
    for (int k = 0; k < nk; ++k) HIPCHK(hipStreamWaitEvent(s.stream, cache.buffer(lease, k).ev, 0));
    return VSZ_OK;
}

// Gathers window frames from source-cache slots into slot.d_src (src/bm3d.zig:327-340).
int gather_from_cache(Bm3d &b, StreamSlot &s, const SlotCache::Lease &src, int count, const int *slot_of_t) {
    for (auto &e : b.entries)
        for (int c = 0; c < b.n_clips; ++c)
            for (int t = 0; t < b.tw; ++t) {
                const SlotCache::Buffer &slot = b.cache.buffer(src, c * count + slot_of_t[t]);
                for (int i = 0; i < e.n_planes; ++i) {
                    const size_t dst_off = e.src_off + ((size_t(c) * e.n_planes + i) * b.tw + t) * e.plane_extent;
                    const size_t coff = e.cache_off + size_t(i) * e.plane_extent;
                    HIPCHK(hipMemcpyAsync(s.d_src + dst_off,
                                          static_cast<const float *>(slot.buf) + coff, e.plane_extent * 4,
                                          hipMemcpyDeviceToDevice, s.stream));
                }
            }
    return VSZ_OK;
}

int Bm3d::launch_entries(StreamSlot &s, float *out_base, int out_kind) {
    for (auto &e : entries) {
        const float *clip0 = s.d_src + e.src_off;
        const float *clip1 = final_pass ? clip0 + size_t(e.n_planes) * tw * e.plane_extent : nullptr;
        float *out = out_base + (out_kind == 0 ? e.dst_off : e.res_off);
        const int rc = vsz_bm3d_run(&e.kp, clip0, clip1, s.d_work + e.work_off, out, out_kind, s.stream);
        if (rc) return rc;
    }
    return VSZ_OK;
}

// ---- download (src/bm3d.zig:427-478) ---------------------------------------------------------
int Bm3d::download(StreamSlot &s, const float *d_base, bool stacked, void *const dst[3], const ptrdiff_t dst_stride[3]) {
    const size_t n = stacked ? sum_res : sum_dst;
    float *h = s.h_stage + down_off();
    if (d_base != s.d_down)  // zero-copy: the kernels already wrote the download region
        HIPCHK(hipMemcpyAsync(h, d_base, n * 4, hipMemcpyDeviceToHost, s.stream));
    HIPCHK(hipStreamSynchronize(s.stream));
    for (auto &e : entries)
        for (int i = 0; i < e.n_planes; ++i) {
            const int p = e.planes[i];
            if (!((e.kp.proc_mask >> i) & 1)) continue;  // skipped planes keep the passthrough copy
            if (!stacked) {
                copy_plane(dst[p], dst_stride[p], h + e.dst_off + size_t(i) * e.plane_extent, ptrdiff_t(e.S) * 4,
                           size_t(e.w) * 4, e.h);
            } else {
                // stacked plane rows: [slot][num|den][h] (src/bm3d.zig:409-419)
                const float *src = h + e.res_off + size_t(i) * tw * 2 * e.plane_extent;
                copy_plane(dst[p], dst_stride[p], src, ptrdiff_t(e.S) * 4, size_t(e.w) * 4, e.h * 2 * tw);
            }
        }
    return VSZ_OK;
}

int Bm3d::process_normal(StreamSlot &s, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                         const ptrdiff_t dst_stride[3]) {
    const int nf = vi.num_frames;
    int idx[MAX_TW];
    for (int t = 0; t < tw; ++t) idx[t] = std::min(std::max(n - radius + t, 0), nf - 1);
    SlotCache::Lease src;
    int rc = upload_window(s, idx, tw, get, user, src);
    if (rc == VSZ_OK && scache) {
        int slot_of_t[MAX_TW];
        for (int t = 0; t < tw; ++t) slot_of_t[t] = t;
        rc = gather_from_cache(*this, s, src, tw, slot_of_t);
    }
    float *out = radius == 0 ? out_base(s) : s.d_out;
    if (rc == VSZ_OK) rc = launch_entries(s, out, radius == 0 ? 0 : 1);
    if (rc == VSZ_OK) rc = download(s, out, radius > 0, dst, dst_stride);
    if (rc != VSZ_OK) (void)hipStreamSynchronize(s.stream);  // drain before the buffers are released
    src.end();
    return rc;
}

// ---- fused chain (src/bm3d.zig:481-633) --------------------------------------------------------
int Bm3d::process_fused(StreamSlot &s, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                        const ptrdiff_t dst_stride[3]) {
    const int r = radius, nf = vi.num_frames;
    const int lo = std::max(n - 2 * r, 0);
    const int span = std::min(n + 2 * r, nf - 1) - lo + 1;
    int fidx[2 * MAX_TW];
    for (int j = 0; j < span; ++j) fidx[j] = lo + j;
    SlotCache::Lease src, stacks;
    int rc = upload_window(s, fidx, span, get, user, src);

    int64_t akeys[MAX_TW];
    int centres[MAX_TW];
    if (rc == VSZ_OK) {
        for (int i = 0; i < tw; ++i) {
            centres[i] = std::min(std::max(n - r + i, 0), nf - 1);
            akeys[i] = centres[i];
        }
        stacks = acc.lease(akeys, tw);
        if (!stacks.held()) rc = vsz_fail(VSZ_ENOMEM, "accumulator cache smaller than one window");
        // one pass per centre frame whose stack is not resident (src/bm3d.zig:557-605)
        for (int i = 0; i < tw && rc == VSZ_OK; ++i) {
            if (!stacks.must_fill(i)) continue;
            const int m = centres[i];
            int slot_of_t[MAX_TW];
            for (int t = 0; t < tw; ++t) slot_of_t[t] = std::min(std::max(m - r + t, 0), nf - 1) - lo;
            rc = gather_from_cache(*this, s, src, span, slot_of_t);
            if (rc == VSZ_OK) rc = launch_entries(s, static_cast<float *>(acc.buffer(stacks, i).buf), 1);
            if (rc == VSZ_OK && hipEventRecord(acc.buffer(stacks, i).ev, s.stream) != hipSuccess) rc = VSZ_EDEVICE;
            if (rc == VSZ_OK) stacks.filled(i);
        }
    }
    if (rc == VSZ_OK) {
        for (int i = 0; i < tw; ++i)
            if (hipStreamWaitEvent(s.stream, acc.buffer(stacks, i).ev, 0) != hipSuccess) rc = VSZ_EDEVICE;
        for (auto &e : entries) {
            if (rc) break;
            for (int pi = 0; pi < e.n_planes && rc == VSZ_OK; ++pi) {
                if (!((e.kp.proc_mask >> pi) & 1)) continue;
                const float *slots[MAX_TW];
                for (int i = 0; i < tw; ++i)
                    slots[i] = static_cast<const float *>(acc.buffer(stacks, i).buf) + e.res_off +
                               size_t(pi) * tw * 2 * e.plane_extent;
                rc = vsz_vaggregate_run(e.w, e.h, e.S, r, n, nf, slots,
                                        out_base(s) + e.dst_off + size_t(pi) * e.plane_extent, s.stream);
            }
        }
    }
    if (rc == VSZ_OK) rc = download(s, out_base(s), false, dst, dst_stride);
    if (rc != VSZ_OK) (void)hipStreamSynchronize(s.stream);  // drain before the buffers are released
    stacks.end();
    src.end();
    return rc;
}

// ---- getFrame (src/bm3d.zig:635-737) ---------------------------------------------------------
int Bm3d::get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) {
    if (n < 0 || n >= vi.num_frames) return vsz_fail(VSZ_EINVAL, "frame number out of range");
    const int np = std::min(vi.num_planes, 3);
    // output frame: radius 0 / fused copies the unprocessed planes of the centre clip frame
    // (newVideoFrame2), stacked output zeroes them when zero_init.
    for (int p = 0; p < np; ++p) {
        if (process[p] || !dst[p]) continue;
        if (normal_out()) {
            ptrdiff_t st = 0;
            const void *src = get(user, 0, n, p, &st);
            if (!src) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
            copy_plane(dst[p], dst_stride[p], src, st, size_t(plane_w(vi, p)) * 4, plane_h(vi, p));
        } else if (zero_init) {
            const int h = plane_h(out.vi, p);
            for (int y = 0; y < h; ++y) std::memset(static_cast<char *>(dst[p]) + y * dst_stride[p], 0, size_t(plane_w(vi, p)) * 4);
        }
    }
    DeviceScope ds(device);
    if (!ds.ok) return vsz_fail(VSZ_EDEVICE, "hipSetDevice failed");
    SlotGuard g(pool);
    return fused ? process_fused(*g.s, n, get, user, dst, dst_stride) : process_normal(*g.s, n, get, user, dst, dst_stride);
}

// ---------------------------------------------------------------------------------------------
// VAggregate on the CPU (src/bm3d.zig:1276-1430)
// ---------------------------------------------------------------------------------------------
class VAggregate final : public vsz_filter {
  public:
    vsz_clip_info vi{}, src_vi{};
    int radius = 0;
    bool process[3] = {false, false, false};

    const char *frame_error() const override { return "VAggregate: process frame failed."; }
    // src/bm3d.zig:1328-1332: stacked frames n-r..n+r, src frame n
    void frame_range(int which, int n, int &lo, int &hi) const override {
        lo = which == 0 ? std::max(n - radius, 0) : n;
        hi = which == 0 ? std::min(n + radius, src_vi.num_frames - 1) : n;
        if (which > 1) hi = lo - 1;
    }

    // which 0 = stacked clip, 1 = src
    int get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) override {
        const int nf = src_vi.num_frames, r = radius, tw = 2 * r + 1;
        if (n < 0 || n >= nf) return vsz_fail(VSZ_EINVAL, "frame number out of range");
        const int np = std::min(src_vi.num_planes, 3);
        for (int p = 0; p < np; ++p) {
            const int w = plane_w(src_vi, p), h = plane_h(src_vi, p);
            if (!process[p]) {
                ptrdiff_t st = 0;
                const void *src = get(user, 1, n, p, &st);
                if (!src) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
                copy_plane(dst[p], dst_stride[p], src, st, size_t(w) * 4, h);
                continue;
            }
            const float *stk[MAX_TW];
            ptrdiff_t sst[MAX_TW];
            for (int i = 0; i < tw; ++i) {
                const int idx = std::min(std::max(n - r + i, 0), nf - 1);
                stk[i] = static_cast<const float *>(get(user, 0, idx, p, &sst[i]));
                if (!stk[i]) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
            }
            for (int y = 0; y < h; ++y) {
                const float *rows[MAX_TW], *wts[MAX_TW];
                for (int i = 0; i < tw; ++i) {
                    const int z = vsz_agg_z(i, n, nf, r);
                    rows[i] = reinterpret_cast<const float *>(reinterpret_cast<const char *>(stk[i]) +
                                                              (ptrdiff_t(z) * 2 * h + y) * sst[i]);
                    wts[i] = reinterpret_cast<const float *>(reinterpret_cast<const char *>(stk[i]) +
                                                             (ptrdiff_t(z) * 2 * h + h + y) * sst[i]);
                }
                float *drow = reinterpret_cast<float *>(static_cast<char *>(dst[p]) + y * dst_stride[p]);
                for (int x = 0; x < w; ++x) {
                    float acc = 0.0f, acw = 0.0f;
                    for (int i = 0; i < tw; ++i) {
                        acc = acc + rows[i][x];
                        acw = acw + wts[i][x];
                    }
                    drow[x] = acc / acw;
                }
            }
        }
        return VSZ_OK;
    }
};

// ---------------------------------------------------------------------------------------------
// BM3Dv2 composed: BM3D stacked output + VAggregate, with a small LRU of stacked frames so each
// BM3D frame is computed once per sliding window (VapourSynth's frame cache plays this role).
// ---------------------------------------------------------------------------------------------
class Composed final : public vsz_filter {
  public:
    std::unique_ptr<Bm3d> bm3d;
    std::unique_ptr<VAggregate> vagg;

    const char *frame_error() const override { return "BM3D: process frame failed."; }
    // the BM3D frames n-r..n+r of the aggregation window each read their own n-r..n+r window
    void frame_range(int which, int n, int &lo, int &hi) const override {
        const int r = vagg->radius;
        lo = 0;
        hi = -1;
        if (which == 0 || (which == 1 && bm3d->n_clips == 2)) {
            lo = std::max(n - 2 * r, 0);
            hi = std::min(n + 2 * r, vagg->src_vi.num_frames - 1);
        }
    }

    int get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) override {
        const int nf = vagg->src_vi.num_frames, r = vagg->radius, tw = 2 * r + 1;
        // BM3D stacked frames m = clamp(n-r+i)
        struct Ctx {
            Composed *self;
            vsz_get_plane_fn get;
            void *user;
            const std::vector<float> *stk[MAX_TW];
            int idx[MAX_TW];
            int tw;
            size_t plane_off[3];
            int S[3];
        } ctx{};
        ctx.self = this;
        ctx.get = get;
        ctx.user = user;
        ctx.tw = tw;
        const vsz_clip_info &svi = bm3d->out.vi;
        size_t off = 0;
        for (int p = 0; p < std::min(svi.num_planes, 3); ++p) {
            ctx.plane_off[p] = off;
            ctx.S[p] = plane_w(svi, p);
            off += size_t(ctx.S[p]) * plane_h(svi, p);
        }
        std::vector<std::shared_ptr<std::vector<float>>> keep;
        for (int i = 0; i < tw; ++i) {
            const int m = std::min(std::max(n - r + i, 0), nf - 1);
            ctx.idx[i] = m;
            std::shared_ptr<std::vector<float>> fr = lookup(m);
            if (!fr) {
                fr = std::make_shared<std::vector<float>>(off);
                void *pl[3] = {nullptr, nullptr, nullptr};
                ptrdiff_t ps[3] = {0, 0, 0};
                for (int p = 0; p < std::min(svi.num_planes, 3); ++p) {
                    pl[p] = fr->data() + ctx.plane_off[p];
                    ps[p] = ptrdiff_t(ctx.S[p]) * 4;
                }
                const int rc = bm3d->get_frame(m, get, user, pl, ps);
                if (rc) return rc;
                insert(m, fr);
            }
            keep.push_back(fr);
            ctx.stk[i] = fr.get();
        }
        // VAggregate getter: which 0 -> stacked frame from `keep`, which 1 -> src (clip)
        struct G {
            static const void *fn(void *u, int which, int fn_n, int plane, ptrdiff_t *st) {
                Ctx *c = static_cast<Ctx *>(u);
                if (which == 1) return c->get(c->user, 0, fn_n, plane, st);
                for (int i = 0; i < c->tw; ++i)
                    if (c->idx[i] == fn_n) {
                        *st = ptrdiff_t(c->S[plane]) * 4;
                        return c->stk[i]->data() + c->plane_off[plane];
                    }
                return nullptr;
            }
        };
        return vagg->get_frame(n, &G::fn, &ctx, dst, dst_stride);
    }

  private:
    std::mutex mu_;
    std::vector<std::pair<int, std::shared_ptr<std::vector<float>>>> lru_;

    std::shared_ptr<std::vector<float>> lookup(int m) {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = lru_.begin(); it != lru_.end(); ++it)
            if (it->first == m) {
                auto v = it->second;
                lru_.erase(it);
                lru_.emplace_back(m, v);
                return v;
            }
        return nullptr;
    }
    void insert(int m, std::shared_ptr<std::vector<float>> v) {
        std::lock_guard<std::mutex> lk(mu_);
        lru_.emplace_back(m, std::move(v));
        const size_t cap = size_t(2 * (2 * vagg->radius + 1)) + 4;
        while (lru_.size() > cap) lru_.erase(lru_.begin());
    }
};

class Passthrough final : public vsz_filter {
  public:
    vsz_clip_info vi{};
    void frame_range(int which, int n, int &lo, int &hi) const override {
        lo = n;
        hi = which == 0 ? n : n - 1;
    }
    int get_frame(int n, vsz_get_plane_fn get, void *user, void *const dst[3], const ptrdiff_t dst_stride[3]) override {
        for (int p = 0; p < std::min(vi.num_planes, 3); ++p) {
            ptrdiff_t st = 0;
            const void *src = get(user, 0, n, p, &st);
            if (!src) return vsz_fail(VSZ_EFRAME, "frame source returned no data");
            copy_plane(dst[p], dst_stride[p], src, st, size_t(plane_w(vi, p)) * 4, plane_h(vi, p));
        }
        return VSZ_OK;
    }
};

int create_vagg(const vsz_clip_info *clip, const vsz_clip_info *src, const int *planes, int n_planes,
                std::unique_ptr<VAggregate> &out, char *err, size_t errlen) {
    auto fail = [&](const char *m) { return set_err(err, errlen, VSZ_EINVAL, m); };
    if (!clip || !src) return fail("VAggregate: missing arguments.");
    if (src->height <= 0 || src->sample_type != 1 || src->bits_per_sample != 32)
        return fail("VAggregate: \"src\" must be 32 bit float.");
    if (clip->color_family != src->color_family || clip->sample_type != src->sample_type ||
        clip->bits_per_sample != src->bits_per_sample || clip->sub_sampling_w != src->sub_sampling_w ||
        clip->sub_sampling_h != src->sub_sampling_h || clip->num_frames != src->num_frames)
        return fail("VAggregate: \"clip\" must have the same format and frame count as \"src\".");
    const int ratio = clip->height / src->height;
    if (ratio < 6 || (ratio - 2) % 4 != 0 || clip->height % src->height != 0 || clip->width != src->width)
        return fail("VAggregate: \"clip\" is not a BM3D stacked clip of \"src\" (radius >= 1).");
    std::unique_ptr<VAggregate> v(new (std::nothrow) VAggregate());
    if (!v) return fail("VAggregate: out of memory.");  // src/bm3d.zig:1419-1421
    v->radius = (ratio - 2) / 4;
    if (v->radius > MAX_RADIUS) return fail("VAggregate: radius too large.");
    for (int i = 0; i < n_planes; ++i) {
        if (planes[i] < 0 || planes[i] > 2) return fail("VAggregate: \"planes\" must be 0..2.");
        v->process[planes[i]] = true;
    }
    if (n_planes <= 0) return fail("VAggregate: \"planes\" is required.");
    v->vi = *clip;
    v->src_vi = *src;
    v->out.vi = *src;
    for (int i = 0; i < 3; ++i) v->out.process[i] = v->process[i];
    out = std::move(v);
    return VSZ_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

int vsz_bm3d_create(const vsz_bm3d_args *args, vsz_filter **out, char *err, size_t errlen) {
    if (!out) return set_err(err, errlen, VSZ_EINVAL, "BM3D: missing output handle.");
    std::unique_ptr<Bm3d> b;
    // standalone BM3D accepts and ignores fast_fused (src/plugin.zig:83)
    const int rc = create_bm3d(args, false, b, err, errlen);
    if (rc) return rc;
    *out = b.release();
    return VSZ_OK;
}

int vsz_bm3dv2_create(const vsz_bm3d_args *a, vsz_filter **out, char *err, size_t errlen) {
    if (!a || !out) return set_err(err, errlen, VSZ_EINVAL, "BM3Dv2: missing arguments.");
    // proc from sigma with the per-plane fallback (src/bm3d.zig:1437-1444)
    bool proc[3] = {true, true, true};
    for (int i = 0; i < 3; ++i) {
        if (i < a->n_sigma) proc[i] = !(a->sigma[i] < FLT_EPSILON);
        else if (i > 0) proc[i] = proc[i - 1];
    }
    bool skip = true;
    for (int i = 0; i < std::min(a->clip.num_planes, 3); ++i)
        if (proc[i]) skip = false;
    if (skip) {
        auto p = std::make_unique<Passthrough>();
        p->vi = a->clip;
        p->out.vi = a->clip;
        p->out.passthrough = 1;
        *out = p.release();
        return VSZ_OK;
    }
    const int radius = a->has_radius ? a->radius : 0;
    if (radius > 0 && a->has_fast_fused && a->fast_fused != 0) {
        std::unique_ptr<Bm3d> b;
        const int rc = create_bm3d(a, true, b, err, errlen);
        if (rc) return rc;
        b->v2_error = "BM3Dv2: process frame failed.";
        *out = b.release();
        return VSZ_OK;
    }
    std::unique_ptr<Bm3d> b;
    int rc = create_bm3d(a, false, b, err, errlen);
    if (rc) return rc;
    if (radius == 0) {
        *out = b.release();
        return VSZ_OK;
    }
    int planes[3], np = 0;
    for (int i = 0; i < 3; ++i)
        if (proc[i]) planes[np++] = i;
    std::unique_ptr<VAggregate> v;
    rc = create_vagg(&b->out.vi, &a->clip, planes, np, v, err, errlen);
    if (rc) return rc;
    auto c = std::make_unique<Composed>();
    c->out = v->out;
    c->bm3d = std::move(b);
    c->vagg = std::move(v);
    *out = c.release();
    return VSZ_OK;
}

int vsz_vaggregate_create(const vsz_clip_info *clip, const vsz_clip_info *src, const int *planes, int n_planes,
                          vsz_filter **out, char *err, size_t errlen) {
    if (!out) return set_err(err, errlen, VSZ_EINVAL, "VAggregate: missing output handle.");
    std::unique_ptr<VAggregate> v;
    const int rc = create_vagg(clip, src, planes, n_planes, v, err, errlen);
    if (rc) return rc;
    *out = v.release();
    return VSZ_OK;
}

int vsz_filter_output_info(const vsz_filter *f, vsz_output_info *info) {
    if (!f || !info) return vsz_fail(VSZ_EINVAL, "null argument");
    *info = f->out;
    return VSZ_OK;
}

int vsz_filter_get_frame(vsz_filter *f, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                         const ptrdiff_t dst_stride[3], char *err, size_t errlen) {
    if (!f || !get || !dst || !dst_stride) return set_err(err, errlen, VSZ_EINVAL, "null argument");
    int rc;
    try {
        rc = f->get_frame(n, get, user, dst, dst_stride);
    } catch (...) {
        rc = VSZ_ENOMEM;
    }
    if (rc) {
        const std::string detail = g_last_error;
        std::fprintf(stderr, "vszipcu: %s (%s)\n", f->frame_error(), detail.c_str());
        return set_err(err, errlen, rc, f->frame_error());
    }
    return VSZ_OK;
}

int vsz_filter_frame_range(const vsz_filter *f, int which, int n, int *first, int *last) {
    if (!f || !first || !last) return vsz_fail(VSZ_EINVAL, "null argument");
    int lo = 0, hi = -1;
    f->frame_range(which, n, lo, hi);
    *first = lo;
    *last = hi;
    return VSZ_OK;
}

void vsz_filter_free(vsz_filter *f) { delete f; }

}  // extern "C"
