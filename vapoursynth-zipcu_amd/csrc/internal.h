// Internal declarations shared by the kernels (bm3d_kernels.hip) and the host engine
// (engine.cpp). Not part of the public ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>

#ifdef __HIP_DEVICE_COMPILE__
#else
#include <cmath>
#endif

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#ifndef VSZ_MAX_TW
#define VSZ_MAX_TW 33  // 2 * max radius (16) + 1, src/bm3d.zig:52-53
#endif

namespace vszk {

#if defined(__HIPCC__)
struct MatchArgs {
    const float *img;  // [TW][H][S] matching plane
    int4 *matches;     // [nblocks][8] (x, y, z, error bits)
    int width, height, stride, nbx, nby;
    int block_step, bm_range, radius, ps_num, ps_range;
    long long ts;
    int tiles_x;     // block tiles per row
    int lds_stride;  // LDS-staged path: floats per staged row
    int no_self;     // 1: leave out the self-inclusion (a temporal stage follows)
    unsigned long long *zero;  // non-null: the spatial search also clears zero[0, zero_n64)
    long long zero_n64;
};

struct FilterArgs {
    const int4 *matches;
    const float *clip0;  // [P][TW][H][S] basic (final) or noisy (basic pass)
    const float *clip1;  // [P][TW][H][S] noisy (final pass)
    void *acc;           // [P][TW][2][H][S] int64 (exact) or float: numerator, corner-scattered weights
    long long nblocks;
    int num_planes, stride;
    long long ts, plane_elems, acc_plane_elems;
    float sigma[3];
    unsigned active_mask;
    int exact, extractor_exp;
    float extractor;
    // tiled path
    int width, height, block_step, bm_range, nbx, nby;
    int tbx, tby, guard_bits;
};

struct FinishArgs {
    const void *acc;
    float *out;
    int width, height, stride, tw, out_kind;
    long long ts;
    unsigned proc_mask;
    int extractor_exp;
};

// Passed by value, like the reference's AggSrc (src/bm3d_vagg.cu:7-16).
struct VaggArgs {
    const float *slots[VSZ_MAX_TW];
    int z[VSZ_MAX_TW];
    float *dst;
    int width, height, stride, tw;
    long long ts;
};
static_assert(sizeof(VaggArgs) < 4096, "kernel argument block must stay under 4 KiB");
#endif

}  // namespace vszk

// Records a thread-local error message and returns `code`.
int vsz_fail(int code, const char *fmt, ...);

// Exact (integer) aggregation is used for a non-zero extractor with a moderate exponent:
// contributions are multiples of 2^(e-24) and sums fit int64 while |sum| < 2^(38+e).
inline bool vsz_exact_extractor(int e) { return e != 0 && e > -40 && e < 40; }

// Slot of window position i for output frame n (src/bm3d.zig:46-48).
inline int vsz_agg_z(int i, int n, int nframes, int radius) {
    return std::min(std::max(2 * radius - i, n - nframes + 1 + radius), n + radius);
}
