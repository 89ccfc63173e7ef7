/* vszipcu_bm3d.h — C ABI of the MI355X BM3D engine (libvszipcu_amd.so).
 *
 * Two layers, both plain C (no HIP, torch or VapourSynth types in any signature; device
 * pointers are `void *` / typed pointers into HIP device memory, streams are `void *`
 * hipStream_t handles, NULL = the null stream):
 *
 *  1. Filter layer: the BM3D / BM3Dv2 / VAggregate filter instances with the reference's
 *     argument rules, error strings and frame semantics. It replaces the VapourSynth create /
 *     getFrame / free callbacks of src/bm3d.zig (createInner :969-1271, getFrame :635-737,
 *     free :739-751, createVAggregate :1372-1430, aggGetFrame :1319-1363, createV2 :1432-1501),
 *     registered by src/plugin.zig:84-98. Frames are pulled through a caller callback, the way
 *     VapourSynth's getFrameFilter hands frames to getFrame.
 *
 *  2. Kernel layer: one device-resident BM3D pass over already-uploaded planes, replacing the
 *     launch sequence of src/bm3d.zig:375-401 (memset res, launch `bm3d`, launch `aggregate`)
 *     and :589-630 (fused accumulator slot + `vaggregate`).
 */
#ifndef VSZIPCU_BM3D_H
#define VSZIPCU_BM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------------------------- */
enum {
    VSZ_OK = 0,
    VSZ_EINVAL = -1,     /* argument rejected (message in the error buffer / vsz_last_error) */
    VSZ_EDEVICE = -2,    /* HIP runtime error */
    VSZ_ENOMEM = -3,     /* device or host allocation failed */
    VSZ_EFRAME = -4,     /* frame callback returned no data */
};

/* Last error message of the calling thread ("" if none). */
const char *vsz_last_error(void);

/* Library / device probes. */
int vsz_version(void);
int vsz_device_count(void); /* 0 without a usable GPU (never aborts) */

/* =============================================================================================
 * Kernel layer
 * ============================================================================================= */

/* Kernel-level parameters of one pass: the #define block src/bm3d.zig:793-823 generates for
 * src/bm3d.cu, as runtime values. `sigma` are the SCALED thresholds (SIGMA_Y/U/V). */
typedef struct vsz_bm3d_kparams {
    int width, height, stride;         /* plane geometry in elements; stride >= width */
    int block_step, bm_range;          /* 1..8, > 0 */
    int radius, ps_num, ps_range;      /* temporal radius 0..16, predictive search 1..8, > 0 */
    int chroma;                        /* 1: three planes share the plane-0 matches (CBM3D) */
    int final_pass;                    /* 1: Wiener pass (clip0 = basic estimate, clip1 = noisy) */
    float sigma[3];                    /* scaled thresholds per plane */
    int extractor_exp;                 /* 0: identity quantiser; else E = 2^exp */
    int metric;                        /* 0 ssd, 1 sad, 2 zssd, 3 zsad, 4 ssd/norm */
    int t2d, t1d;                      /* 0 dct, 1 haar, 2 wht, 3 bior1.5 */
    unsigned proc_mask;                /* planes the final division writes (src/bm3d.cu:910-914) */
    int transform_precision;           /* 0: FP32 transforms with the reference's op order (default);
                                          1: DCT stages as compensated split-BF16 MFMA products with
                                          FP32 accumulation (dct/dct only; matching stays FP32) */
} vsz_bm3d_kparams;

/* Reference blocks per axis (src/bm3d.cu:588-595): ceil((N-8)/block_step)+1. */
int vsz_bm3d_num_blocks(const vsz_bm3d_kparams *p, int *nbx, int *nby);

/* Device workspace bytes one pass needs (match lists + accumulator stack). */
size_t vsz_bm3d_workspace_size(const vsz_bm3d_kparams *p);

/* Block matching only (src/bm3d.cu:579-788). d_img: [TW][H][stride] float32, the matching
 * plane (plane 0 of clip 0). d_matches: [nby][nbx][8][4] int32 = (x, y, z, float bits of the
 * block distance), ranks ascending. */
int vsz_bm3d_match(const vsz_bm3d_kparams *p, const float *d_img, int32_t *d_matches, void *stream);

/* Diagnostic, no reference counterpart: the matcher vsz_bm3d_match would launch for these
 * parameters: "lds4" (shared-column, block_step 4), "lds_rows1/2/4" (LDS-staged, search rows split
 * over 1, 2 or 4 lanes), "lds_short" (LDS-staged 16 x 8-block tile) or "global" (raw buffer loads,
 * windows beyond the LDS budget); "+temporal" when radius > 0 (centre frame, then the predictive
 * search). "" for invalid parameters. The returned string is static. */
const char *vsz_bm3d_match_plan(const vsz_bm3d_kparams *p);

/* Diagnostic, no reference counterpart: the filter vsz_bm3d_run would launch for these
 * parameters. Radius 0 takes the LDS-tiled filter: *tbx x *tby reference blocks per 256-thread
 * workgroup, *per_cu the workgroups that registers and LDS let share one CU (at most 2 for the
 * FP32 Wiener pass, 3 for the FP32 hard pass, 4 / 5 for split-BF16 Wiener / hard). The hard pass
 * of a plane that fills one round of 768 such workgroups gets a tile shape planned for three per
 * CU. 0 x 0 tiles (and *per_cu 0) mean the direct-atomic filter: radius > 0, or a search region
 * beyond the LDS budget. Returns VSZ_OK or VSZ_EINVAL. */
int vsz_bm3d_filter_plan(const vsz_bm3d_kparams *p, int *tbx, int *tby, int *per_cu);

/* One full pass: matching, collaborative filtering and aggregation.
 *   d_clip0: [P][TW][H][stride]  basic estimate (final pass) or noisy clip (basic pass)
 *   d_clip1: [P][TW][H][stride]  noisy clip (final pass only, else ignored)
 *   d_work : vsz_bm3d_workspace_size() bytes
 *   out_kind 0: d_out = [P][H][stride] = num/den for planes in proc_mask (radius 0 only;
 *               other planes untouched)                        (src/bm3d.cu:898-922)
 *   out_kind 1: d_out = [P][TW][2][H][stride] float32 num/den accumulator stack
 *               (the `res` buffer of src/bm3d.cu:572-573, the stacked-clip payload)
 * P = 3 when chroma, else 1. Runs asynchronously on `stream`. */
int vsz_bm3d_run(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, void *d_work,
                 float *d_out, int out_kind, void *stream);

/* vsz_bm3d_run without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. The match lists and accumulators live in a static device array
 * of 128 MiB (134217728 bytes) that each GPU holds from the library's first use until the process
 * ends; a pass whose vsz_bm3d_workspace_size() exceeds it is rejected with VSZ_EINVAL. The call
 * allocates nothing. Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race, so serialise them. Inputs, outputs and results are those of
 * vsz_bm3d_run. */
int vsz_bm3d_run_noalloc(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, float *d_out,
                         int out_kind, void *stream);

/* Temporal aggregation of one plane (src/bm3d_vagg.cu:18-48): d_dst[y][x] =
 * sum_i slot_i[z_i][0] / sum_i slot_i[z_i][1], i ascending, z_i = min(max(2r-i, n-nf+1+r), n+r).
 * d_slots[i]: [TW][2][H][stride] float32 stack of centre frame clamp(n-r+i, 0, nf-1). */
int vsz_vaggregate_run(int width, int height, int stride, int radius, int n, int num_frames,
                       const float *const *d_slots, float *d_dst, void *stream);

/* ---- Batched independent planes -------------------------------------------------------------
 * Many Gray planes of different sizes and sigmas denoised together: every stage (clear, match,
 * filter, finish) is ONE launch over all planes, and a run is one hipGraph (captured on the
 * first run, replayed afterwards). No reference counterpart as one call: it batches the
 * per-entry launch loop of src/bm3d.zig:375-401 across planes. */
typedef struct vsz_bm3d_plane {
    int width, height, stride;   /* elements */
    float sigma;                 /* the filter's raw `sigma` (scaled internally per pass) */
} vsz_bm3d_plane;

typedef struct vsz_bm3d_batch_params {
    int block_step, bm_range, extractor_exp;
    int metric, t2d, t1d;        /* one transform for both stages: t2d == t1d (dct 0, haar 1, wht 2,
                                    bior1.5 3); dct with split-BF16 and the persistent schedule */
    int transform_precision;     /* as in vsz_bm3d_kparams */
    int two_pass;                /* 1: basic, then Wiener with ref = basic; 0: basic only */
    int autotune;                /* 1: on the first run, time every fitting filter tile shape per
                                    plane shape, and whole runs of every allowed (schedule, matcher)
                                    pair, and keep the fastest; 0: cost model and preferences */
    int schedule;                /* VSZ_BATCH_STAGED: one launch per stage (zeroing, matching,
                                    filtering, finishing of all planes, per pass) in one graph;
                                    VSZ_BATCH_PERSISTENT: the whole batch in one persistent launch
                                    (extractor_exp > 0, ssd or sad); VSZ_BATCH_AUTO: the autotuner
                                    times both (staged without autotune) */
} vsz_bm3d_batch_params;

#define VSZ_BATCH_STAGED 0
#define VSZ_BATCH_PERSISTENT 1
#define VSZ_BATCH_AUTO 2

typedef struct vsz_bm3d_batch vsz_bm3d_batch;

int vsz_bm3d_batch_create(const vsz_bm3d_batch_params *bp, const vsz_bm3d_plane *planes, int n,
                          vsz_bm3d_batch **out);
size_t vsz_bm3d_batch_workspace_size(const vsz_bm3d_batch *b);
/* d_noisy[i]: [height][stride] input, d_out[i]: [height][stride] result of plane i. The first
 * run tunes (when asked) and captures the graph; a run with other pointers re-captures, after
 * waiting for the previous run's replay (which may be on another stream) to finish. A batch object
 * serves one call at a time: calls from several threads must be serialised by the caller. */
int vsz_bm3d_batch_run(vsz_bm3d_batch *b, const float *const *d_noisy, float *const *d_out, void *d_work,
                       void *stream);
/* Filter tile (in reference blocks) chosen for plane i, pass 0 (basic) or 1 (final). */
int vsz_bm3d_batch_tile(const vsz_bm3d_batch *b, int plane, int pass, int *tbx, int *tby);
/* Schedule (VSZ_BATCH_STAGED or VSZ_BATCH_PERSISTENT) and matcher (1: shared-column block_step-4
 * matcher, 0: row matcher, 2: global-memory matcher for windows beyond the LDS budget) of the
 * batch; final after its first run. */
int vsz_bm3d_batch_schedule(const vsz_bm3d_batch *b, int *schedule, int *matcher);
/* Waits for the work queued on `stream`, then reports whether any dependency wait of the batch's
 * persistent launches (since its creation) ran into its time bound, i.e. ran before its inputs
 * were complete: VSZ_OK, or an error. Always VSZ_OK for the staged schedule. */
int vsz_bm3d_batch_check(vsz_bm3d_batch *b, void *stream);
void vsz_bm3d_batch_free(vsz_bm3d_batch *b);

/* =============================================================================================
 * Filter layer
 * ============================================================================================= */

/* Clip format, the fields of VSVideoInfo / VSVideoFormat the filters read. */
typedef struct vsz_clip_info {
    int width, height, num_frames;
    int color_family;    /* VS4 values: 0 undefined, 1 Gray, 2 RGB, 3 YUV */
    int sample_type;     /* 0 integer, 1 float */
    int bits_per_sample;
    int sub_sampling_w, sub_sampling_h;
    int num_planes;
} vsz_clip_info;

/* Arguments of BM3D / BM3Dv2 (src/plugin.zig:84-90). For each optional argument the `has_*` /
 * `n_*` field says whether (and how many elements) the caller gave, mirroring VSMap lookups:
 * a per-plane element i that is absent falls back to element i-1 (src/bm3d.zig:953-963). */
typedef struct vsz_bm3d_args {
    vsz_clip_info clip;
    int has_ref;
    vsz_clip_info ref;
    int n_sigma;         float sigma[3];
    int n_block_step;    int block_step[3];
    int n_bm_range;      int bm_range[3];
    int has_radius;      int radius;
    int n_ps_num;        int ps_num[3];
    int n_ps_range;      int ps_range[3];
    int has_chroma;      int chroma;
    int has_device_id;   int device_id;
    int has_num_streams; int num_streams;
    int has_extractor_exp; int extractor_exp;
    int n_bm_error_s;    const char *bm_error_s[3];
    int n_transform_2d_s; const char *transform_2d_s[3];
    int n_transform_1d_s; const char *transform_1d_s[3];
    int has_zero_init;   int zero_init;
    int has_fast_fused;  int fast_fused;
} vsz_bm3d_args;

/* Frame source: returns plane `plane` of frame `n` of input `which` and its row stride in
 * bytes, or NULL on failure. For BM3D / BM3Dv2: which 0 = clip, 1 = ref. For VAggregate:
 * which 0 = the stacked clip, 1 = src. */
typedef const void *(*vsz_get_plane_fn)(void *user, int which, int n, int plane, ptrdiff_t *stride_bytes);

typedef struct vsz_filter vsz_filter;

/* Output description: the VSVideoInfo the filter would register, plus the frame props it sets
 * (BM3D_V_radius, BM3D_V_process) on stacked radius>0 output (src/bm3d.zig:716-720). */
typedef struct vsz_output_info {
    vsz_clip_info vi;
    int passthrough;      /* 1: BM3Dv2 returned its input unchanged (no plane processed) */
    int stacked;          /* 1: BM3D radius>0 stacked num/den clip */
    int radius;
    int process[3];
} vsz_output_info;

/* BM3D (src/bm3d.zig:965): errors are the reference's verbatim strings. Returns VSZ_OK and a
 * handle, or an error code with the message copied into err (and vsz_last_error). */
int vsz_bm3d_create(const vsz_bm3d_args *args, vsz_filter **out, char *err, size_t errlen);
/* BM3Dv2 (src/bm3d.zig:1432): BM3D, then VAggregate when radius>0, or the fused accumulator
 * chain when fast_fused; passthrough when no plane is processed. */
int vsz_bm3dv2_create(const vsz_bm3d_args *args, vsz_filter **out, char *err, size_t errlen);
/* VAggregate (src/bm3d.zig:1372): clip = stacked BM3D output, src = the source clip. */
int vsz_vaggregate_create(const vsz_clip_info *clip, const vsz_clip_info *src, const int *planes, int n_planes,
                          vsz_filter **out, char *err, size_t errlen);

int vsz_filter_output_info(const vsz_filter *f, vsz_output_info *info);

/* getFrame: computes output frame n into dst[plane] (row stride dst_stride[plane] bytes).
 * Unprocessed planes are copied from the source frame (radius 0, VAggregate) or zeroed when
 * zero_init (stacked output). Thread-safe: up to num_streams calls run concurrently.
 * On failure returns an error code; the message is "BM3D: process frame failed." (or the
 * BM3Dv2 variant), as src/bm3d.zig:676, 727 sets. */
int vsz_filter_get_frame(vsz_filter *f, int n, vsz_get_plane_fn get, void *user, void *const dst[3],
                         const ptrdiff_t dst_stride[3], char *err, size_t errlen);

/* Frames of input `which` (as in vsz_get_plane_fn) that vsz_filter_get_frame(f, n, ...) may
 * pull: the inclusive range [*first, *last], already clamped to the clip; *first > *last when
 * the filter does not read that input. These are the frames the reference requests in the
 * arInitial activation (src/bm3d.zig:643-651, fused 2r window; :1328-1332 for VAggregate). */
int vsz_filter_frame_range(const vsz_filter *f, int which, int n, int *first, int *last);

void vsz_filter_free(vsz_filter *f);

#ifdef __cplusplus
}
#endif

#endif /* VSZIPCU_BM3D_H */
