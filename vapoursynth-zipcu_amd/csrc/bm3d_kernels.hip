// BM3D device path for gfx950 (CDNA4): block matching, collaborative filtering, exact
// aggregation, final division and temporal aggregation, plus the kernel-layer C ABI
// (vsz_bm3d_match / vsz_bm3d_run / vsz_vaggregate_run) declared in include/vszipcu_bm3d.h.
//
// Reference path: src/bm3d.cu:571-922 (kernels `bm3d`, `aggregate`) and src/bm3d_vagg.cu:18-48
// (`vaggregate`), launched by src/bm3d.zig:375-401 and :589-630.
//
// Numerics contract (identical in the CPU oracle): built without FP contraction and without
// fast math; every fused multiply-add is an explicit fmaf; divisions and sqrt are IEEE
// correctly rounded; reductions follow the reference's per-lane order and the 8-lane
// butterfly ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)). Aggregation with a non-zero extractor
// E = 2^e sums the quantised contributions q(u) = (u+E)-E, which all lie on the grid
// 2^(e-24), exactly in 64-bit integers (device-wide atomics commute on integers, so the
// result is deterministic); each sum is rounded once to FP32 before the IEEE division.
// With E = 0 the sums are FP32 atomics, order-dependent exactly as in the reference.

#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "vszipcu_bm3d.h"
#include "internal.h"

namespace vszk {

// ------------------------------------------------------------------------------------------------
// 8-point transforms (src/bm3d.cu:32-280). Same op sequence as the reference codelets: the
// FMA/FMS/FNMS macros become fused multiply-adds, everything else is one rounding per op.
// ------------------------------------------------------------------------------------------------
namespace kc {
constexpr float tan_pi_8 = float(0.414213562373095048801688724209698078569671875);
constexpr float two_cos_pi_8 = float(1.847759065022573512256366378793576573644833252);
constexpr float tan_pi_16 = float(0.198912367379658006911597622644676228597850501);
constexpr float two_cos_pi_16 = float(1.961570560806460898252364472268478073947867462);
constexpr float sqrt2 = float(1.414213562373095048801688724209698078569671875);
constexpr float tan_3pi_16 = float(0.668178637919298919997757686523080761552472251);
constexpr float two_cos_3pi_16 = float(1.662939224605090474157576755235811513477121624);
constexpr float inv_sqrt2 = float(0.707106781186547524400844362104849039284835938);
constexpr float two_sqrt2 = float(2.828427124746190097603377448419396157139343751);
constexpr float b_877 = float(0.877670597010003062405456290501163901200537360);
constexpr float b_1797 = float(1.797135031972863413496886690073811797696338403);
constexpr float b_2277 = float(2.277437593371384746027611934034934695963515886);
constexpr float b_1609 = float(1.609389232649111887192845766718020518480884560);
constexpr float b_334 = float(0.334024180361136429417383083658457088741315663);
constexpr float b_1495 = float(1.495435764250674860795011090214036706658653686);
constexpr float b_2058 = float(2.058234225009388964222454285384072231477027482);
constexpr float b_486 = float(0.486135912065751423025580498947083714508324707);
}  // namespace kc

__device__ __forceinline__ float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float fnms_(float a, float b, float c) { return __builtin_fmaf(-a, b, c); }
__device__ __forceinline__ float fms_(float a, float b, float c) { return __builtin_fmaf(a, b, -c); }

template <int KIND, bool FWD> __device__ __forceinline__ void xf8(float *v) {
    using namespace kc;
    if constexpr (KIND == 0 && FWD) {  // dct, forward
        const float a07 = v[0] - v[7], s07 = v[0] + v[7];
        const float a43 = v[4] - v[3], s43 = v[4] + v[3];
        const float a25 = v[2] - v[5], a16 = v[1] - v[6];
        const float p = a25 + a16, s16 = v[1] + v[6];
        const float m = a25 - a16, s25 = v[2] + v[5];
        const float tb = fnms_(inv_sqrt2, p, a07);
        const float tg = fnms_(inv_sqrt2, m, a43);
        const float o3 = two_cos_3pi_16 * fmaf_(tan_3pi_16, tg, tb);
        const float o5 = -(two_cos_3pi_16 * fnms_(tan_3pi_16, tb, tg));
        const float tp = s07 + s43, tq = s25 + s16;
        const float o4 = sqrt2 * (tp - tq);
        const float o0 = sqrt2 * (tp + tq);
        const float th = fmaf_(inv_sqrt2, p, a07);
        const float ti = fmaf_(inv_sqrt2, m, a43);
        const float o1 = two_cos_pi_16 * fnms_(tan_pi_16, ti, th);
        const float o7 = two_cos_pi_16 * fmaf_(tan_pi_16, th, ti);
        const float tl = s07 - s43, to = s25 - s16;
        const float o2 = two_cos_pi_8 * fnms_(tan_pi_8, to, tl);
        const float o6 = two_cos_pi_8 * fmaf_(tan_pi_8, tl, to);
        v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3; v[4] = o4; v[5] = o5; v[6] = o6; v[7] = o7;
    } else if constexpr (KIND == 0) {  // dct, inverse
        const float t1 = v[0] * sqrt2;
        const float t3 = fmaf_(sqrt2, v[4], t1);
        const float tj = fnms_(sqrt2, v[4], t1);
        const float t6 = fmaf_(tan_pi_8, v[6], v[2]);
        const float tk = fms_(tan_pi_8, v[2], v[6]);
        const float tb = v[5] + v[3];
        const float te = v[3] - v[5];
        const float tc = fmaf_(inv_sqrt2, tb, v[1]);
        const float tn = fnms_(inv_sqrt2, te, v[7]);
        const float tf = fmaf_(inv_sqrt2, te, v[7]);
        const float tm = fnms_(inv_sqrt2, tb, v[1]);
        const float t7 = fmaf_(two_cos_pi_8, t6, t3);
        const float tg = fmaf_(tan_pi_16, tf, tc);
        const float o7 = fnms_(two_cos_pi_16, tg, t7);
        const float o0 = fmaf_(two_cos_pi_16, tg, t7);
        const float tp = fnms_(two_cos_pi_8, tk, tj);
        const float tq = fmaf_(tan_3pi_16, tm, tn);
        const float o5 = fnms_(two_cos_3pi_16, tq, tp);
        const float o2 = fmaf_(two_cos_3pi_16, tq, tp);
        const float th = fnms_(two_cos_pi_8, t6, t3);
        const float ti = fnms_(tan_pi_16, tc, tf);
        const float o3 = fnms_(two_cos_pi_16, ti, th);
        const float o4 = fmaf_(two_cos_pi_16, ti, th);
        const float tl = fmaf_(two_cos_pi_8, tk, tj);
        const float to = fnms_(tan_3pi_16, tn, tm);
        const float o6 = fnms_(two_cos_3pi_16, to, tl);
        const float o1 = fmaf_(two_cos_3pi_16, to, tl);
        v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3; v[4] = o4; v[5] = o5; v[6] = o6; v[7] = o7;
    } else if constexpr (KIND == 1 && FWD) {  // haar, forward
        const float s01 = v[0] + v[1], d01 = v[0] - v[1];
        const float s23 = v[2] + v[3], d23 = v[2] - v[3];
        const float s45 = v[4] + v[5], d45 = v[4] - v[5];
        const float s67 = v[6] + v[7], d67 = v[6] - v[7];
        const float a = s01 + s23, b = sqrt2 * (s01 - s23);
        const float c = s45 + s67, d = sqrt2 * (s45 - s67);
        const float sc2 = sqrt2 * 2.0f;
        v[0] = sqrt2 * (a + c); v[1] = sqrt2 * (a - c); v[2] = sqrt2 * b; v[3] = sqrt2 * d;
        v[4] = sc2 * d01; v[5] = sc2 * d23; v[6] = sc2 * d45; v[7] = sc2 * d67;
    } else if constexpr (KIND == 1) {  // haar, inverse (rank 4 in the reference; kept literally)
        const float s01 = v[0] + v[1], d01 = v[0] - v[1];
        const float p2 = sqrt2 * v[2], q4 = 2.0f * v[4], n2 = (-sqrt2) * v[2];
        const float pp = p2 + q4, pm = p2 - q4, np = n2 + q4, nm = n2 - q4;
        v[0] = sqrt2 * (s01 + pp); v[1] = sqrt2 * (s01 + pm); v[2] = sqrt2 * (s01 + np); v[3] = sqrt2 * (s01 + nm);
        v[4] = sqrt2 * (d01 + pp); v[5] = sqrt2 * (d01 + pm); v[6] = sqrt2 * (d01 + np); v[7] = sqrt2 * (d01 + nm);
    } else if constexpr (KIND == 2) {  // wht (self-inverse up to 16)
        const float a = v[0] + v[1], b = v[0] - v[1], c = v[2] + v[3], d = v[2] - v[3];
        const float e = v[4] + v[5], f = v[4] - v[5], g = v[6] + v[7], h = v[6] - v[7];
        const float ac = a + c, a_c = a - c, bd = b + d, b_d = b - d;
        const float eg = e + g, e_g = e - g, fh = f + h, f_h = f - h;
        v[0] = sqrt2 * (ac + eg); v[1] = sqrt2 * (ac - eg); v[2] = sqrt2 * (a_c - e_g); v[3] = sqrt2 * (a_c + e_g);
        v[4] = sqrt2 * (b_d + f_h); v[5] = sqrt2 * (b_d - f_h); v[6] = sqrt2 * (bd - fh); v[7] = sqrt2 * (bd + fh);
    } else if constexpr (FWD) {  // bior1.5 analysis
        const float s03 = v[0] + v[3], d03 = v[0] - v[3], s12 = v[1] + v[2], d12 = v[1] - v[2];
        const float s47 = v[4] + v[7], d47 = v[4] - v[7], s56 = v[5] + v[6], d56 = v[5] - v[6];
        const float e01 = v[0] - v[1], e23 = v[2] - v[3], e45 = v[4] - v[5], e67 = v[6] - v[7];
        const float o0 = sqrt2 * (((s03 + s47) + s12) + s56);
        const float o1 = (b_877 * (s03 - s47)) + (b_1797 * (s12 - s56));
        const float o2 = ((b_2277 * d03) + (b_1609 * d12)) + (b_334 * (d56 - d47));
        const float o3 = ((b_2277 * d47) + (b_1609 * d56)) + (b_334 * (d12 - d03));
        v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
        v[4] = two_sqrt2 * e01; v[5] = two_sqrt2 * e23; v[6] = two_sqrt2 * e45; v[7] = two_sqrt2 * e67;
    } else {  // bior1.5 synthesis
        const float u0 = sqrt2 * v[0], u1 = b_1495 * v[1], u2 = b_2058 * v[2], u3 = b_2058 * v[3];
        const float a4 = two_sqrt2 * v[4], b4 = b_486 * v[4], a5 = two_sqrt2 * v[5], b5 = b_486 * v[5];
        const float a6 = two_sqrt2 * v[6], b6 = b_486 * v[6], a7 = two_sqrt2 * v[7], b7 = b_486 * v[7];
        const float p = u0 + u1, q = u0 - u1, r = b5 - b7, s = b4 - b6;
        v[0] = (p + u2) + (a4 - r); v[1] = (p + u2) - (a4 + r);
        v[2] = (p - u2) + (s + a5); v[3] = (p - u2) + (s - a5);
        v[4] = (q + u3) + (r + a6); v[5] = (q + u3) + (r - a6);
        v[6] = (q - u3) - (s - a7); v[7] = (q - u3) - (s + a7);
    }
}

__device__ __forceinline__ float tree8(const float *c) {
    return ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
}

// ------------------------------------------------------------------------------------------------
// Block matching (src/bm3d.cu:579-788). One lane per reference block; a 64-lane workgroup
// covers an 8x8 arrangement of blocks. A search row is scanned left to right with the 8x8
// candidate held in registers: each step loads one new column (8 values) into the slot the
// leftmost column vacates; the slots rotate with the step index mod 8, which a static
// unroll-by-8 turns into fixed register names. Every lane walks the same displacement grid;
// candidates outside its clipped window are skipped (their loads use clamped addresses).
// ------------------------------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): workgroup g
// runs on XCD g % 8. Remap so that XCD k processes tiles [k*n/8, (k+1)*n/8), a contiguous band
// whose search windows overlap in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int g, int n) {
    const int xcd = g & 7, slot = g >> 3;
    const int per = n >> 3, extra = n & 7;  // XCDs below `extra` take one more tile
    return xcd * per + min(xcd, extra) + slot;
}

// Per-lane reference patch a[r*8 + c] and the metric's reference-only terms.
struct RefPatch {
    float a[64];
    float mean;  // zssd / zsad: tree-summed patch mean (src/bm3d.cu:319-321)
    float norm;  // ssd/norm: sqrt of the tree-summed squares (src/bm3d.cu:363-369)
};

__device__ __forceinline__ float col_sum8(const float *p, int stride) {
    return ((p[0] + p[stride]) + (p[2 * stride] + p[3 * stride])) +
           ((p[4 * stride] + p[5 * stride]) + (p[6 * stride] + p[7 * stride]));
}

template <int M> __device__ __forceinline__ void ref_terms(RefPatch &R) {
    if constexpr (M == 2 || M == 3) {
        float ca[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) ca[c] = col_sum8(R.a + c, 8);
        R.mean = tree8(ca) * (1.0f / 64.0f);
    } else if constexpr (M == 4) {
        float qa[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
            for (int r = 0; r < 8; r += 2) {
                e0 = fmaf_(R.a[r * 8 + c], R.a[r * 8 + c], e0);
                e1 = fmaf_(R.a[(r + 1) * 8 + c], R.a[(r + 1) * 8 + c], e1);
            }
            qa[c] = e0 + e1;
        }
        R.norm = __builtin_sqrtf(tree8(qa));  // IEEE sqrt (__fsqrt_rn is the native approximation here)
    }
}

// Distance (src/bm3d.cu:290-390) of the reference patch to the candidate whose column c sits in
// window buffer (U + c) % 9; w[buffer][row].
template <int M, int U> __device__ __forceinline__ float win_dist(const RefPatch &R, const float (&w)[9][8]) {
    float col[8];
    if constexpr (M == 0 || M == 1) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float *b = w[(U + c) % 9];
            float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
            for (int r = 0; r < 8; r += 2) {
                const float d0 = R.a[r * 8 + c] - b[r];
                const float d1 = R.a[(r + 1) * 8 + c] - b[r + 1];
                if constexpr (M == 0) {
                    e0 = fmaf_(d0, d0, e0);
                    e1 = fmaf_(d1, d1, e1);
                } else {
                    e0 = e0 + fabsf(d0);
                    e1 = e1 + fabsf(d1);
                }
            }
            col[c] = e0 + e1;
        }
        return tree8(col);
    } else if constexpr (M == 2 || M == 3) {
        float cb[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) cb[c] = col_sum8(w[(U + c) % 9], 1);
        const float dm = R.mean - tree8(cb) * (1.0f / 64.0f);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float *b = w[(U + c) % 9];
            float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
            for (int r = 0; r < 8; r += 2) {
                const float d0 = (R.a[r * 8 + c] - b[r]) - dm;
                const float d1 = (R.a[(r + 1) * 8 + c] - b[r + 1]) - dm;
                if constexpr (M == 2) {
                    e0 = fmaf_(d0, d0, e0);
                    e1 = fmaf_(d1, d1, e1);
                } else {
                    e0 = e0 + fabsf(d0);
                    e1 = e1 + fabsf(d1);
                }
            }
            col[c] = e0 + e1;
        }
        return tree8(col);
    } else {  // ssd/norm
        float qb[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float *b = w[(U + c) % 9];
            float b0 = 0.0f, b1 = 0.0f, x0 = 0.0f, x1 = 0.0f;
#pragma unroll
            for (int r = 0; r < 8; r += 2) {
                b0 = fmaf_(b[r], b[r], b0);
                x0 = fmaf_(R.a[r * 8 + c], b[r], x0);
                b1 = fmaf_(b[r + 1], b[r + 1], b1);
                x1 = fmaf_(R.a[(r + 1) * 8 + c], b[r + 1], x1);
            }
            qb[c] = b0 + b1;
            col[c] = x0 + x1;
        }
        const float nb = __builtin_sqrtf(tree8(qb));
        return 2.0f - __fdiv_rn(2.0f * tree8(col), (R.norm * nb) + FLT_EPSILON);
    }
}

// Sorted 8-list of (distance, position, slot): position packs x | y << 16 (planes are below
// 65536 in both dimensions); the slot array is carried only by temporal passes.
struct List8 {
    float er[8];
    unsigned pos[8];
    int zs[8];
};

__device__ __forceinline__ unsigned pack_xy(int x, int y) { return (unsigned)x | ((unsigned)y << 16); }
__device__ __forceinline__ int pos_x(unsigned p) { return (int)(p & 0xffffu); }
__device__ __forceinline__ int pos_y(unsigned p) { return (int)(p >> 16); }

template <bool TZ> __device__ __forceinline__ void list_init(List8 &L, int z) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        L.er[k] = FLT_MAX;
        L.pos[k] = 0u;
        if constexpr (TZ) L.zs[k] = z;
    }
}

// Strict-< sorted insertion (src/bm3d.cu:641-653): a new entry goes before the first entry
// with a strictly greater distance, so ties keep the earlier scan position in front. A distance
// of FLT_MAX (how callers pass candidates outside the window) or NaN is never inserted.
template <bool TZ>
__device__ __forceinline__ void insert8(List8 &L, float e, unsigned pos, int z) {
    if (e < L.er[7]) {
#pragma unroll
        for (int k = 7; k > 0; --k) {
            const bool fk = e < L.er[k];
            const bool fp = e < L.er[k - 1];
            if (fk) {
                L.pos[k] = fp ? L.pos[k - 1] : pos;
                if constexpr (TZ) L.zs[k] = fp ? L.zs[k - 1] : z;
            }
            // er[k-1] <= er[k] and e is not NaN here, so the new entry is the median of the three
            // (the same value as the selection above in every case, ties included)
            L.er[k] = __builtin_amdgcn_fmed3f(L.er[k - 1], e, L.er[k]);
        }
        const bool f0 = e < L.er[0];
        L.pos[0] = f0 ? pos : L.pos[0];
        if constexpr (TZ) L.zs[0] = f0 ? z : L.zs[0];
        L.er[0] = f0 ? e : L.er[0];
    }
}

// Raw buffer view of one [TW][H][S] plane stack: loads take a 32-bit byte offset per lane and an
// immediate; offsets outside [0, bytes) return 0 instead of faulting, so candidates beyond the
// plane edge need no address clamping (their distances are discarded).
struct PlaneBuf {
    __amdgpu_buffer_rsrc_t rsrc;
};

__device__ __forceinline__ PlaneBuf make_plane_buf(const float *base, unsigned bytes) {
    // dword3 0x00020000: raw (untyped) 32-bit buffer access on gfx9-family targets
    return PlaneBuf{__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)bytes, 0x00020000)};
}

template <int IMM> __device__ __forceinline__ float buf_load(const PlaneBuf &b, int voff) {
    return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(b.rsrc, voff + IMM, 0, 0));
}

// Per-lane byte offset plus a wave-uniform one (an SGPR operand: no address arithmetic per load).
__device__ __forceinline__ float buf_load_s(const PlaneBuf &b, int voff, int soff) {
    return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(b.rsrc, voff, soff, 0));
}

// Loads column (step U + J) of the 8 rows whose byte offsets (at the block's first column) are vo[r].
template <int IMM> __device__ __forceinline__ void load_col(float (&col)[8], const PlaneBuf &b, const int (&vo)[8]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) col[r] = buf_load<IMM>(b, vo[r]);
}

// One step of a search row: issue the load of the column the NEXT step needs (it has the whole
// distance computation to arrive), then evaluate this step's candidate. vo[r] is the byte offset
// of row r at the group's first column xg = xt - U. With CLAMP (xg < 0) the column is addressed
// with its whole offset in the VGPR, clamped to the row start: in the first image row of the
// stack vo[r] is negative there, and a buffer load does not bring a negative VGPR offset back into
// range with its immediate offset (the load returns 0). Columns left of the row belong only to
// candidates that are never inserted.
template <int M, bool TZ, int U, bool CLAMP>
__device__ __forceinline__ void scan_step(float (&w)[9][8], const PlaneBuf &b, const int (&vo)[8], int W, int xt,
                                          bool row_ok, unsigned ybits, int z, const RefPatch &R, List8 &L) {
    if constexpr (CLAMP) {
        const int sh = 4 * max(U + 8, U - xt);
#pragma unroll
        for (int r = 0; r < 8; ++r) w[(U + 8) % 9][r] = buf_load<0>(b, vo[r] + sh);
    } else {
        load_col<(U + 8) * 4>(w[(U + 8) % 9], b, vo);
    }
    const float e = win_dist<M, U>(R, w);
    insert8<TZ>(L, row_ok && (unsigned)xt <= (unsigned)(W - 8) ? e : FLT_MAX, (unsigned)xt | ybits, z);
}

// Nine steps of a search row from column xg = x0 + tb (fewer at the row's end).
template <int M, bool TZ, bool CLAMP>
__device__ __forceinline__ void scan_group(float (&w)[9][8], const PlaneBuf &b, const int (&vo)[8], int W, int xg,
                                           int left, bool row_ok, unsigned ybits, int z, const RefPatch &R,
                                           List8 &L) {
    scan_step<M, TZ, 0, CLAMP>(w, b, vo, W, xg + 0, row_ok, ybits, z, R, L);
    if (left < 1) return;
    scan_step<M, TZ, 1, CLAMP>(w, b, vo, W, xg + 1, row_ok, ybits, z, R, L);
    if (left < 2) return;
    scan_step<M, TZ, 2, CLAMP>(w, b, vo, W, xg + 2, row_ok, ybits, z, R, L);
    if (left < 3) return;
    scan_step<M, TZ, 3, CLAMP>(w, b, vo, W, xg + 3, row_ok, ybits, z, R, L);
    if (left < 4) return;
    scan_step<M, TZ, 4, CLAMP>(w, b, vo, W, xg + 4, row_ok, ybits, z, R, L);
    if (left < 5) return;
    scan_step<M, TZ, 5, CLAMP>(w, b, vo, W, xg + 5, row_ok, ybits, z, R, L);
    if (left < 6) return;
    scan_step<M, TZ, 6, CLAMP>(w, b, vo, W, xg + 6, row_ok, ybits, z, R, L);
    if (left < 7) return;
    scan_step<M, TZ, 7, CLAMP>(w, b, vo, W, xg + 7, row_ok, ybits, z, R, L);
    if (left < 8) return;
    scan_step<M, TZ, 8, CLAMP>(w, b, vo, W, xg + 8, row_ok, ybits, z, R, L);
}

// Scans the window [cx-rng, cx+rng] x [cy-rng, cy+rng] of frame z, clipped to the plane, in the
// reference's row-major order (src/bm3d.cu:618-660), inserting into L with slot index z. Column
// j of the row lives in buffer j % 9: eight are the current candidate, one is in flight.
template <int M, bool TZ>
__device__ __forceinline__ void scan_window(const PlaneBuf &b, long long frame_off, int S, int W, int H,
                                            const RefPatch &R, int cx, int cy, int rng, int z, List8 &L) {
    const int span = 2 * rng;
    const int x0 = cx - rng;
    for (int dy = -rng; dy <= rng; ++dy) {
        const int yy = cy + dy;
        const bool row_ok = yy >= 0 && yy <= H - 8;
        const unsigned ybits = (unsigned)yy << 16;
        int vo[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) vo[r] = (int)((frame_off + (long long)(yy + r) * S + x0) * 4);
        // first 8 columns. A window starting inside the row loads them as two 16-byte vectors per
        // row (if the first candidate is valid its 8 columns lie inside the row; if not, nothing
        // in this row is inserted). A window starting left of the row uses separately clamped
        // scalar offsets, so no vector access can straddle the buffer start (a partially
        // out-of-range vector load returns all zeros).
        float w[9][8];
        if (x0 >= 0) {
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int4 v = __builtin_bit_cast(int4, __builtin_amdgcn_raw_buffer_load_b128(b.rsrc, vo[r] + 16 * q, 0, 0));
                    w[4 * q + 0][r] = __int_as_float(v.x);
                    w[4 * q + 1][r] = __int_as_float(v.y);
                    w[4 * q + 2][r] = __int_as_float(v.z);
                    w[4 * q + 3][r] = __int_as_float(v.w);
                }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int dj = 4 * (max(x0 + j, 0) - x0);
#pragma unroll
                for (int r = 0; r < 8; ++r) w[j][r] = buf_load<0>(b, vo[r] + dj);
            }
        }
        // groups that start left of the plane first (lane-dependent count), then the rest
        int tb = 0;
        for (; tb <= span && x0 + tb < 0; tb += 9) {
            scan_group<M, TZ, true>(w, b, vo, W, x0 + tb, span - tb, row_ok, ybits, z, R, L);
#pragma unroll
            for (int r = 0; r < 8; ++r) vo[r] += 36;
        }
        for (; tb <= span; tb += 9) {
            scan_group<M, TZ, false>(w, b, vo, W, x0 + tb, span - tb, row_ok, ybits, z, R, L);
#pragma unroll
            for (int r = 0; r < 8; ++r) vo[r] += 36;
        }
    }
}

// Image row y, columns x0 .. x0+15 of one frame. Inside the row: four 16-byte vector loads;
// otherwise per-column loads clamped to the row (only candidates that are never inserted read
// the clamped columns). Rows outside the frame read other frames' rows or, beyond the stack,
// zeros; they too belong to candidates that are never inserted.
__device__ __forceinline__ void load_band_row(float (&row)[16], const PlaneBuf &b, int row_byte, int x0, int W) {
    if (x0 >= 0 && x0 + 16 <= W) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 v = __builtin_bit_cast(int4, __builtin_amdgcn_raw_buffer_load_b128(b.rsrc, row_byte + 4 * x0 + 16 * q, 0, 0));
            row[4 * q + 0] = __int_as_float(v.x);
            row[4 * q + 1] = __int_as_float(v.y);
            row[4 * q + 2] = __int_as_float(v.z);
            row[4 * q + 3] = __int_as_float(v.w);
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) row[j] = buf_load<0>(b, row_byte + 4 * min(max(x0 + j, 0), W - 1));
    }
}

// Predictive-search window of range 4 (9 x 9 candidates, src/bm3d.cu:690-727) in the reference's
// row-major order, from a band of 8 image rows x 16 columns held in registers: each search row
// loads only the image row that enters the band. Same distances and insertion order as
// scan_window.
template <int M>
__device__ __forceinline__ void scan_band4(const PlaneBuf &b, long long frame_off, int S, int W, int H,
                                           const RefPatch &R, int cx, int cy, int z, List8 &F) {
    constexpr int kRng = 4;
    const int x0 = cx - kRng, y0 = cy - kRng;
    const int row_bytes = S * 4, base = (int)(frame_off * 4) + y0 * row_bytes;
    float band[8][16];
#pragma unroll
    for (int r = 0; r < 7; ++r) load_band_row(band[r], b, base + r * row_bytes, x0, W);
    for (int dy = 0; dy <= 2 * kRng; ++dy) {
        load_band_row(band[7], b, base + (dy + 7) * row_bytes, x0, W);
        const int yy = y0 + dy;
        const bool row_ok = yy >= 0 && yy <= H - 8;
        const unsigned ybits = (unsigned)yy << 16;
#pragma unroll
        for (int t = 0; t <= 2 * kRng; ++t) {
            float w[8][8];  // w[column][row] of candidate t
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) w[c][r] = band[r][t + c];
            float w9[9][8];
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) w9[c][r] = w[c][r];
            const float e = win_dist<M, 0>(R, w9);
            const int xt = x0 + t;
            insert8<false>(F, row_ok && (unsigned)xt <= (unsigned)(W - 8) ? e : FLT_MAX, (unsigned)xt | ybits, z);
        }
#pragma unroll
        for (int r = 0; r < 7; ++r)
#pragma unroll
            for (int j = 0; j < 16; ++j) band[r][j] = band[r + 1][j];
    }
}

// ---- LDS-staged spatial matching -------------------------------------------------------------
// A 256-lane workgroup owns 16x16 reference blocks (2x2 waves of 8x8). It stages the union of their
// search windows, [x_first-R, x_last+R+9) x [y_first-R, y_last+R+8) of the plane, in LDS once
// (pixels outside the plane are stored as 0; only invalid candidates read them), then every lane
// scans with LDS reads at immediate offsets. Row Y of the region starts at Y*LS + ((Y>>2)&3):
// lanes of one wave are 4 rows apart for block_step 4, so the row-dependent shift spreads a
// wave's column reads over all banks.
__device__ __forceinline__ int lds_row(int Y, int LS) { return Y * LS + ((Y >> 2) & 3); }

template <int IMM> __device__ __forceinline__ void lds_col(float (&col)[8], const char *lds, const int (&vo)[8]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) col[r] = *reinterpret_cast<const float *>(lds + vo[r] + IMM);
}

// Same, from per-row LDS pointers (the region base is folded in once per search row). The
// pointers carry the LDS address space so the loads stay ds_read instructions.
typedef const char __attribute__((address_space(3))) *lds_cptr;
typedef const float __attribute__((address_space(3))) *lds_fptr;

template <int IMM> __device__ __forceinline__ void lds_colp(float (&col)[8], const lds_cptr (&pr)[8]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) col[r] = *(lds_fptr)(pr[r] + IMM);
}

template <int M, int U>
__device__ __forceinline__ void lds_step(float (&w)[9][8], const char *lds, const int (&vo)[8], int W, int xt,
                                         bool row_ok, unsigned ybits, const RefPatch &R, List8 &L) {
    lds_col<(U + 8) * 4>(w[(U + 8) % 9], lds, vo);
    const float e = win_dist<M, U>(R, w);
    insert8<false>(L, row_ok && (unsigned)xt <= (unsigned)(W - 8) ? e : FLT_MAX, (unsigned)xt | ybits, 0);
}

// WR wave rows: a workgroup of 128 * WR lanes. SPLIT lanes share a block: each scans a
// contiguous band of its search rows, and the later bands' sorted lists are inserted into the
// earlier ones with the same strict-< rule, which reproduces one sequential scan exactly. A wave
// covers 8 x (8 / SPLIT) blocks, the workgroup 16 x WR * (8 / SPLIT).
template <int M, int WR = 2, int SPLIT = 1>
__device__ __forceinline__ void match_lds_body(const MatchArgs &a, int tile, float *region) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int W = a.width, H = a.height, S = a.stride, bs = a.block_step, Rg = a.bm_range;
    constexpr int TH = WR * 8 / SPLIT;   // tile height in blocks
    const int bx0 = tx * 16, by0 = ty * TH;
    const int x_first = min(bx0 * bs, W - 8), x_last = min(min(bx0 + 15, a.nbx - 1) * bs, W - 8);
    const int y_first = min(by0 * bs, H - 8), y_last = min(min(by0 + TH - 1, a.nby - 1) * bs, H - 8);
    const int rx0 = x_first - Rg, ry0 = y_first - Rg;
    const int RW = x_last - x_first + 2 * Rg + 9, RH = y_last - y_first + 2 * Rg + 8;
    const int LS = a.lds_stride;

    for (int i = threadIdx.x; i < RW * RH; i += 128 * WR) {
        const int Y = i / RW, X = i % RW;
        const int gy = ry0 + Y, gx = rx0 + X;
        region[lds_row(Y, LS) + X] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? a.img[(long long)gy * S + gx] : 0.0f;
    }
    __syncthreads();

    constexpr int PL = 64 / SPLIT;      // lanes per band in a wave
    const int part = lane / PL, li = lane % PL;
    const int bx = bx0 + (wave & 1) * 8 + (li & 7), by = by0 + (wave >> 1) * (8 / SPLIT) + (li >> 3);
    const bool valid = bx < a.nbx && by < a.nby;
    const int x = min(min(bx, a.nbx - 1) * bs, W - 8);
    const int y = min(min(by, a.nby - 1) * bs, H - 8);
    RefPatch R;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) R.a[r * 8 + c] = a.img[(long long)(y + r) * S + x + c];
    ref_terms<M>(R);
    List8 L;
    list_init<false>(L, 0);

    const char *lds = reinterpret_cast<const char *>(region);
    const int span = 2 * Rg;
    const int X0 = x - Rg - rx0;
    const int nrows = 2 * Rg + 1;
    const int dy_end = -Rg + (nrows * (part + 1)) / SPLIT;
    for (int dy = -Rg + (nrows * part) / SPLIT; dy < dy_end; ++dy) {
        const int yy = y + dy;
        const bool row_ok = yy >= 0 && yy <= H - 8;
        const unsigned ybits = (unsigned)yy << 16;
        int vo[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) vo[r] = (lds_row(yy + r - ry0, LS) + X0) * 4;
        float w[9][8];
        lds_col<0>(w[0], lds, vo);
        lds_col<4>(w[1], lds, vo);
        lds_col<8>(w[2], lds, vo);
        lds_col<12>(w[3], lds, vo);
        lds_col<16>(w[4], lds, vo);
        lds_col<20>(w[5], lds, vo);
        lds_col<24>(w[6], lds, vo);
        lds_col<28>(w[7], lds, vo);
        const int x0 = x - Rg;
        for (int tb = 0; tb <= span; tb += 9) {
            const int xt = x0 + tb;
            lds_step<M, 0>(w, lds, vo, W, xt + 0, row_ok, ybits, R, L);
            if (tb + 1 > span) break;
            lds_step<M, 1>(w, lds, vo, W, xt + 1, row_ok, ybits, R, L);
            if (tb + 2 > span) break;
            lds_step<M, 2>(w, lds, vo, W, xt + 2, row_ok, ybits, R, L);
            if (tb + 3 > span) break;
            lds_step<M, 3>(w, lds, vo, W, xt + 3, row_ok, ybits, R, L);
            if (tb + 4 > span) break;
            lds_step<M, 4>(w, lds, vo, W, xt + 4, row_ok, ybits, R, L);
            if (tb + 5 > span) break;
            lds_step<M, 5>(w, lds, vo, W, xt + 5, row_ok, ybits, R, L);
            if (tb + 6 > span) break;
            lds_step<M, 6>(w, lds, vo, W, xt + 6, row_ok, ybits, R, L);
            if (tb + 7 > span) break;
            lds_step<M, 7>(w, lds, vo, W, xt + 7, row_ok, ybits, R, L);
            if (tb + 8 > span) break;
            lds_step<M, 8>(w, lds, vo, W, xt + 8, row_ok, ybits, R, L);
#pragma unroll
            for (int r = 0; r < 8; ++r) vo[r] += 36;
        }
    }

    // merge the bands in scan order: (0 <- 1, 2 <- 3), then 0 <- 2
#pragma unroll
    for (int d = PL; d < 64; d *= 2) {
        float oe[8];
        unsigned op[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            oe[k] = __shfl_xor(L.er[k], d);
            op[k] = (unsigned)__shfl_xor((int)L.pos[k], d);
        }
        if ((part & ((2 * d / PL) - 1)) == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) insert8<false>(L, oe[k], op[k], 0);
        }
    }
    if (part != 0) return;

    const unsigned self = pack_xy(x, y);
    bool found = a.no_self != 0;   // a temporal stage does the self-inclusion after its merges
#pragma unroll
    for (int k = 0; k < 8; ++k) found |= L.pos[k] == self;
    if (!found) {
#pragma unroll
        for (int k = 7; k > 0; --k) {
            L.er[k] = L.er[k - 1];
            L.pos[k] = L.pos[k - 1];
        }
        L.er[0] = 0.0f;
        L.pos[0] = self;
    }
    if (!valid) return;
    int4 *out = a.matches + ((long long)by * a.nbx + bx) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = make_int4(pos_x(L.pos[k]), pos_y(L.pos[k]), 0, __float_as_int(L.er[k]));
}

// Clears this workgroup's share of the accumulator that the pass's filtering adds into (the
// filter runs after the matcher on the same stream). Issued after the search: the stores drain
// while the other waves of the CU keep the VALU busy, instead of taking a bandwidth-bound kernel
// of their own ahead of the matcher.
__device__ __forceinline__ void zero_range(unsigned long long *z, long long n64) {
    if (!z) return;
    const long long per = (n64 + gridDim.x - 1) / gridDim.x;
    const long long b = (long long)blockIdx.x * per, e = b + per < n64 ? b + per : n64;
    for (long long i = b + threadIdx.x; i < e; i += blockDim.x) z[i] = 0ull;
}

__device__ __forceinline__ void zero_share(const MatchArgs &a) { zero_range(a.zero, a.zero_n64); }

template <int M, int WR, int SPLIT>
__global__ __launch_bounds__(128 * WR) void k_match_lds(MatchArgs a) {
    extern __shared__ float region[];
    match_lds_body<M, WR, SPLIT>(a, xcd_tile(blockIdx.x, gridDim.x), region);
    zero_share(a);
}

// ---- shared-column matching for block_step 4 (ssd, sad) ---------------------------------------
// With block_step 4 the reference block right of x starts at x + 4, so at any displacement its
// columns 0..3 are the same (reference pixel, candidate pixel) pairs as columns 4..7 of the
// block at x, and the per-column sums (even/odd row chains, then e0 + e1) are bit-identical.
// Each lane therefore computes only its columns 4..7 and takes columns 0..3 from the lane on its
// left with a DPP row shift. Lanes form rows of 16: lane 0 of a row is a helper that computes
// the columns left of the row's first block (it owns no block), lanes 1..15 own 15 consecutive
// blocks of one block row. Each block's search rows are split between two lanes of the same
// wave (upper and lower half of the window), which doubles the waves in flight; a workgroup
// covers 15 x 8 blocks. Requires (W - 8) % 4 == 0, so no block position is clamped and
// neighbours are exactly 4 pixels apart.
__device__ __forceinline__ float dpp_from_left(float v) {
    // row_shr:1 within 16-lane rows; lane 0 of a row (the helper, whose distance is never used)
    // reads 0 (bound_ctrl), so no copy of the old value is needed
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}

template <int M, int U>
__device__ __forceinline__ void lds4_step(float (&w)[5][8], const lds_cptr (&pr)[8], const float (&a)[32],
                                          bool ok, unsigned pos, List8 &L) {
    lds_colp<(U + 4) * 4>(w[(U + 4) % 5], pr);
    float own[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float *b = w[(U + c) % 5];
        float e0 = 0.0f, e1 = 0.0f;
#pragma unroll
        for (int r = 0; r < 8; r += 2) {
            const float d0 = a[r * 4 + c] - b[r];
            const float d1 = a[(r + 1) * 4 + c] - b[r + 1];
            if constexpr (M == 0) {
                e0 = fmaf_(d0, d0, e0);
                e1 = fmaf_(d1, d1, e1);
            } else {
                e0 = e0 + fabsf(d0);
                e1 = e1 + fabsf(d1);
            }
        }
        own[c] = e0 + e1;
    }
    // tree8 of {left neighbour's own[0..3], own[0..3]}: the shift is a lane permutation, so the
    // left half ((c0 + c1) + (c2 + c3)) is the shifted right half, summed once
    const float half_sum = (own[0] + own[1]) + (own[2] + own[3]);
    const float e = dpp_from_left(half_sum) + half_sum;
    insert8<false>(L, ok ? e : FLT_MAX, pos, 0);  // FLT_MAX is never inserted: one branch per step
}

template <int M>
__device__ __forceinline__ void match_lds4_body(const MatchArgs &a, int tile, float *region) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int W = a.width, H = a.height, S = a.stride, Rg = a.bm_range;
    const int bx0 = tx * 15, by0 = ty * 8;
    const int x_first = bx0 * 4, x_last = min(bx0 + 14, a.nbx - 1) * 4;
    const int y_first = min(by0 * 4, H - 8), y_last = min(min(by0 + 7, a.nby - 1) * 4, H - 8);
    const int rx0 = x_first - Rg, ry0 = y_first - Rg;
    const int RW = x_last - x_first + 2 * Rg + 9, RH = y_last - y_first + 2 * Rg + 8;
    const int LS = a.lds_stride;

    for (int i = threadIdx.x; i < RW * RH; i += 256) {
        const int Y = i / RW, X = i % RW;
        const int gy = ry0 + Y, gx = rx0 + X;
        region[lds_row(Y, LS) + X] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? a.img[(long long)gy * S + gx] : 0.0f;
    }
    __syncthreads();

    // DPP row q of the wave: block row by0 + 2*wave + (q & 1), search rows of half q >> 1
    // (half 0: dy = -R..0, half 1: dy = 1..R); the halves' lists are merged at the end.
    const int q = lane >> 4, half = q >> 1;
    const int rc = (lane & 15) - 1;            // -1: the row's helper lane
    const int bx = bx0 + rc, by = by0 + 2 * wave + (q & 1);
    const bool valid = rc >= 0 && bx < a.nbx && by < a.nby;
    const int x = rc < 0 ? x_first - 4 : min(bx, a.nbx - 1) * 4;   // helper: the block left of the row
    const int y = min(min(by, a.nby - 1) * 4, H - 8);
    float ref[32];                             // this lane's columns 4..7 (x + 4 >= 0 even for the helper)
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) ref[r * 4 + c] = a.img[(long long)(y + r) * S + x + 4 + c];
    List8 L;
    list_init<false>(L, 0);

    const lds_cptr lds = (lds_cptr)(const char *)region;
    const int span = 2 * Rg;
    const int X0 = x + 4 - Rg - rx0;           // column 4 of the first candidate
    const int dy_end = half == 0 ? 0 : Rg;
    for (int dy = half == 0 ? -Rg : 1; dy <= dy_end; ++dy) {
        const int yy = y + dy;
        const bool row_ok = valid && yy >= 0 && yy <= H - 8;
        const unsigned ybits = (unsigned)yy << 16;
        lds_cptr pr[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) pr[r] = lds + (lds_row(yy + r - ry0, LS) + X0) * 4;
        float w[5][8];
        lds_colp<0>(w[0], pr);
        lds_colp<4>(w[1], pr);
        lds_colp<8>(w[2], pr);
        lds_colp<12>(w[3], pr);
        const int x0 = x - Rg;
        for (int tb = 0; tb <= span; tb += 5) {
            const int xt = x0 + tb;
            // one unsigned compare per candidate: a negative x wraps above W - 8
            const unsigned xmax = (unsigned)(W - 8);
            lds4_step<M, 0>(w, pr, ref, row_ok && (unsigned)xt <= xmax, (unsigned)xt | ybits, L);
            if (tb + 1 > span) break;
            lds4_step<M, 1>(w, pr, ref, row_ok && (unsigned)(xt + 1) <= xmax, (unsigned)(xt + 1) | ybits, L);
            if (tb + 2 > span) break;
            lds4_step<M, 2>(w, pr, ref, row_ok && (unsigned)(xt + 2) <= xmax, (unsigned)(xt + 2) | ybits, L);
            if (tb + 3 > span) break;
            lds4_step<M, 3>(w, pr, ref, row_ok && (unsigned)(xt + 3) <= xmax, (unsigned)(xt + 3) | ybits, L);
            if (tb + 4 > span) break;
            lds4_step<M, 4>(w, pr, ref, row_ok && (unsigned)(xt + 4) <= xmax, (unsigned)(xt + 4) | ybits, L);
#pragma unroll
            for (int r = 0; r < 8; ++r) pr[r] += 20;
        }
    }

    // Merge: half 1 scanned later rows, so inserting its entries in order into half 0's list
    // with the same strict-< rule gives exactly the list of one sequential scan.
    float oe[8];
    unsigned op[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        oe[k] = __shfl_xor(L.er[k], 32);
        op[k] = (unsigned)__shfl_xor((int)L.pos[k], 32);
    }
    if (half != 0 || !valid) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) insert8<false>(L, oe[k], op[k], 0);

    const unsigned self = pack_xy(x, y);
    bool found = a.no_self != 0;   // a temporal stage does the self-inclusion after its merges
#pragma unroll
    for (int k = 0; k < 8; ++k) found |= L.pos[k] == self;
    if (!found) {
#pragma unroll
        for (int k = 7; k > 0; --k) {
            L.er[k] = L.er[k - 1];
            L.pos[k] = L.pos[k - 1];
        }
        L.er[0] = 0.0f;
        L.pos[0] = self;
    }
    int4 *out = a.matches + ((long long)by * a.nbx + bx) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = make_int4(pos_x(L.pos[k]), pos_y(L.pos[k]), 0, __float_as_int(L.er[k]));
}

template <int M>
__global__ __launch_bounds__(256) void k_match_lds4(MatchArgs a) {
    extern __shared__ float region[];
    match_lds4_body<M>(a, xcd_tile(blockIdx.x, gridDim.x), region);
    zero_share(a);
}

// Spatial search on raw buffer loads, for windows whose staged region exceeds the LDS budget: one
// lane per reference block, 8 x 8 blocks per wave.
template <int M>
__device__ __forceinline__ void match_global_body(const MatchArgs &a, int tile, int lane) {
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int bx = tx * 8 + (lane & 7), by = ty * 8 + (lane >> 3);
    const bool valid = bx < a.nbx && by < a.nby;
    const int W = a.width, H = a.height, S = a.stride;
    const int x = min(min(bx, a.nbx - 1) * a.block_step, W - 8);
    const int y = min(min(by, a.nby - 1) * a.block_step, H - 8);
    const PlaneBuf pb = make_plane_buf(a.img, (unsigned)(a.ts * 4));

    RefPatch R;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) R.a[r * 8 + c] = a.img[(long long)(y + r) * S + x + c];
    ref_terms<M>(R);

    List8 L;
    list_init<false>(L, 0);
    scan_window<M, false>(pb, 0, S, W, H, R, x, y, a.bm_range, 0, L);

    // self-inclusion (src/bm3d.cu:758-788); a centre-frame search that a temporal stage follows
    // (no_self) leaves it to that stage, which applies it after its merges
    const unsigned self = pack_xy(x, y);
    bool found = a.no_self != 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) found |= L.pos[k] == self;
    if (!found) {
#pragma unroll
        for (int k = 7; k > 0; --k) {
            L.er[k] = L.er[k - 1];
            L.pos[k] = L.pos[k - 1];
        }
        L.er[0] = 0.0f;
        L.pos[0] = self;
    }
    if (!valid) return;
    int4 *out = a.matches + ((long long)by * a.nbx + bx) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = make_int4(pos_x(L.pos[k]), pos_y(L.pos[k]), 0, __float_as_int(L.er[k]));
}

template <int M>
__global__ __launch_bounds__(64) void k_match(MatchArgs a) {
    match_global_body<M>(a, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x);
    zero_share(a);
}

// Temporal matching after the centre frame's spatial search: reads the centre lists, runs the
// predictive search in the other frames (src/bm3d.cu:663-756) and the merges and the
// self-inclusion in the reference's order. PW lanes per block and direction (PW = 2 when ps_num
// is 2: one window each; PW = 1 scans the windows one after another). The window lanes' lists of
// a frame are merged in window order (the later list's entries inserted into the earlier one with
// the same strict-< rule, which equals one sequential scan). Each direction stores the first
// ps_num entries of every frame's list (frames in order) in LDS, and after the scan the block's
// first lane inserts the backward entries, then the forward ones, into the centre list: the same
// sequence of insertions as one sequential search, without keeping the centre list in registers
// during the scans. A 64-lane workgroup covers 8 x (4 / PW) blocks.
template <int M, int PW>
__global__ __launch_bounds__(64, 2) void k_match_temporal(MatchArgs a) {
    extern __shared__ int2 kept[];              // [32 / PW blocks][2 directions][radius * ps_num] (pos, error bits)
    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int win = lane % PW, dsel = (lane / PW) & 1, bl = lane / (2 * PW);
    const int dir = dsel ? 1 : -1;
    const int bx = tx * 8 + (bl & 7), by = ty * (4 / PW) + (bl >> 3);
    const bool valid = bx < a.nbx && by < a.nby;
    const int W = a.width, H = a.height, S = a.stride, Rt = a.radius;
    const int x = min(min(bx, a.nbx - 1) * a.block_step, W - 8);
    const int y = min(min(by, a.nby - 1) * a.block_step, H - 8);
    const float *centre = a.img + (long long)Rt * a.ts;
    const PlaneBuf pb = make_plane_buf(a.img, (unsigned)(a.ts * (2 * Rt + 1) * 4));

    RefPatch R;
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) R.a[r * 8 + c] = centre[(long long)(y + r) * S + x + c];
    ref_terms<M>(R);

    int4 *out = a.matches + ((long long)min(by, a.nby - 1) * a.nbx + min(bx, a.nbx - 1)) * 8;
    unsigned lpos[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int4 m = out[k];
        lpos[k] = pack_xy(m.x, m.y);
    }
    const int per_dir = Rt * a.ps_num;
    int2 *mine = kept + (bl * 2 + dsel) * per_dir;
    for (int t = 1; t <= Rt; ++t) {
        const int tz = Rt + dir * t;
        List8 F;
        list_init<false>(F, tz);
        if constexpr (PW == 2) {  // ps_num == 2: this lane's window only
            const unsigned c = win ? lpos[1] : lpos[0];
            if (a.ps_range == 4) scan_band4<M>(pb, (long long)tz * a.ts, S, W, H, R, pos_x(c), pos_y(c), tz, F);
            else scan_window<M, false>(pb, (long long)tz * a.ts, S, W, H, R, pos_x(c), pos_y(c), a.ps_range, tz, F);
            float oe[8];
            unsigned op[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                oe[k] = __shfl_xor(F.er[k], 1);
                op[k] = (unsigned)__shfl_xor((int)F.pos[k], 1);
            }
            if (win == 0) {
#pragma unroll
                for (int k = 0; k < 8; ++k) insert8<false>(F, oe[k], op[k], tz);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) F.pos[k] = (unsigned)__shfl((int)F.pos[k], lane & ~1);
        } else if (a.ps_range == 4) {
            for (int i = 0; i < a.ps_num; ++i)
                scan_band4<M>(pb, (long long)tz * a.ts, S, W, H, R, pos_x(lpos[i]), pos_y(lpos[i]), tz, F);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < a.ps_num)
                    scan_window<M, false>(pb, (long long)tz * a.ts, S, W, H, R, pos_x(lpos[i]), pos_y(lpos[i]),
                                          a.ps_range, tz, F);
        }
        if (win == 0) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < a.ps_num) mine[(t - 1) * a.ps_num + i] = make_int2((int)F.pos[i], __float_as_int(F.er[i]));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) lpos[k] = F.pos[k];
    }
    __syncthreads();
    if (dsel != 0 || win != 0 || !valid) return;

    List8 L;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int4 m = out[k];
        L.er[k] = __int_as_float(m.w);
        L.pos[k] = pack_xy(m.x, m.y);
        L.zs[k] = Rt;
    }
    for (int d = 0; d < 2; ++d) {              // backward frames first, then forward
        const int2 *e = kept + (bl * 2 + d) * per_dir;
        for (int j = 0; j < per_dir; ++j) {
            const int tz = Rt + (d == 0 ? -1 : 1) * (j / a.ps_num + 1);
            insert8<true>(L, __int_as_float(e[j].y), (unsigned)e[j].x, tz);
        }
    }

    const unsigned self = pack_xy(x, y);
    bool found = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) found |= (L.pos[k] == self && L.zs[k] == Rt);
    if (!found) {
#pragma unroll
        for (int k = 7; k > 0; --k) {
            L.er[k] = L.er[k - 1];
            L.pos[k] = L.pos[k - 1];
            L.zs[k] = L.zs[k - 1];
        }
        L.er[0] = 0.0f;
        L.pos[0] = self;
        L.zs[0] = Rt;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = make_int4(pos_x(L.pos[k]), pos_y(L.pos[k]), L.zs[k], __float_as_int(L.er[k]));
}

// ------------------------------------------------------------------------------------------------
// Collaborative filtering. 8 lanes per group, 8 groups per wave. A lane holds 64 values of its
// group: in column layout v[i*8 + j] is (rank i, row j, column = lane & 7); in row layout
// v[i*8 + c] is (rank i, row = lane & 7, column c). The vertical passes run in column layout,
// the horizontal passes, the group axis, the shrinkage and the aggregation in row layout; the
// switch is an 8x8 transpose per rank through a per-group LDS buffer. LDS instructions of one
// wave execute in order, so the transposes only need compiler barriers, not waits.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void lds_order() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Switches all 8 ranks between column and row layout (the transpose is its own inverse).
__device__ __forceinline__ void transpose_ranks(float *v, float *buf, int l8) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) buf[j * 9 + l8] = v[i * 8 + j];
        lds_order();
#pragma unroll
        for (int j = 0; j < 8; ++j) v[i * 8 + j] = buf[l8 * 9 + j];
        lds_order();
    }
}

// ---- compensated split-BF16 MFMA form of the DCT stages ---------------------------------------
// Transform kind kDctMfma (transform_precision 1): the DCT computed as y = M x with M = 4*C (C the
// orthonormal DCT-II; M^T for the inverse), which is what the reference codelets compute up to
// rounding. Every FP32 operand is split into bf16 hi + lo (hi = RNE(x), lo = RNE(x - hi)); a
// product is Ah*Bl + Al*Bh + Ah*Bh with FP32 accumulation (the lo*lo term is below 2^-16
// relative). The filter kernels run this form on group pairs (filter_pair_mfma below).
constexpr int kDctMfma = 4;
typedef short vsz_s4 __attribute__((ext_vector_type(4)));
typedef float vsz_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 vsz_b4 __attribute__((ext_vector_type(4)));

__constant__ float kDct8x4[64] = {  // 4*C[u][x], row u = frequency
    1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f,
    1.961570501e+00f, 1.662939191e+00f, 1.111140490e+00f, 3.901806474e-01f, -3.901806474e-01f, -1.111140490e+00f, -1.662939191e+00f, -1.961570501e+00f,
    1.847759008e+00f, 7.653668523e-01f, -7.653668523e-01f, -1.847759008e+00f, -1.847759008e+00f, -7.653668523e-01f, 7.653668523e-01f, 1.847759008e+00f,
    1.662939191e+00f, -3.901806474e-01f, -1.961570501e+00f, -1.111140490e+00f, 1.111140490e+00f, 1.961570501e+00f, 3.901806474e-01f, -1.662939191e+00f,
    1.414213538e+00f, -1.414213538e+00f, -1.414213538e+00f, 1.414213538e+00f, 1.414213538e+00f, -1.414213538e+00f, -1.414213538e+00f, 1.414213538e+00f,
    1.111140490e+00f, -1.961570501e+00f, 3.901806474e-01f, 1.662939191e+00f, -1.662939191e+00f, -3.901806474e-01f, 1.961570501e+00f, -1.111140490e+00f,
    7.653668523e-01f, -1.847759008e+00f, 1.847759008e+00f, -7.653668523e-01f, -7.653668523e-01f, 1.847759008e+00f, -1.847759008e+00f, 7.653668523e-01f,
    3.901806474e-01f, -1.111140490e+00f, 1.662939191e+00f, -1.961570501e+00f, 1.961570501e+00f, -1.662939191e+00f, 1.111140490e+00f, -3.901806474e-01f,
};

struct SplitOp {
    vsz_s4 hi, lo;
};

typedef __bf16 vsz_b2 __attribute__((ext_vector_type(2)));
typedef float vsz_f2 __attribute__((ext_vector_type(2)));
typedef unsigned vsz_u2 __attribute__((ext_vector_type(2)));

// Two floats rounded to bf16 (round to nearest even), packed: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((vsz_f2){a, b}, vsz_b2));
}

// A 32-bit pattern in a VGPR, so that an instruction reads it as a register operand: the
// compiler would otherwise encode a packed-bf16 constant such as (-1, 0) as an inline constant.
__device__ __forceinline__ unsigned vgpr_bits(unsigned bits) {
    unsigned v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "i"(bits));
    return v;
}

// hi = bf16(x), lo = bf16(x - hi) for a pair: two instructions per value. The residual x - hi
// comes from one bf16 dot product each, x + (-1)*hi[k] + 0*hi[1-k]: the product is exact and x - hi
// (at most 16 significant bits) is representable, so it is exact before lo rounds it to bf16.
__device__ __forceinline__ void split2(float a, float b, unsigned &hi, unsigned &lo) {
    hi = pk_bf16(a, b);
    const vsz_b2 h = __builtin_bit_cast(vsz_b2, hi);
    const vsz_b2 neg0 = __builtin_bit_cast(vsz_b2, vgpr_bits(0x0000bf80u));  // (-1, 0)
    const vsz_b2 neg1 = __builtin_bit_cast(vsz_b2, vgpr_bits(0xbf800000u));  // (0, -1)
    lo = pk_bf16(__builtin_amdgcn_fdot2_f32_bf16(h, neg0, a, false), __builtin_amdgcn_fdot2_f32_bf16(h, neg1, b, false));
}

__device__ __forceinline__ SplitOp split4(float a, float b, float c, float d) {
    unsigned h0, l0, h1, l1;
    split2(a, b, h0, l0);
    split2(c, d, h1, l1);
    return SplitOp{__builtin_bit_cast(vsz_s4, (vsz_u2){h0, h1}), __builtin_bit_cast(vsz_s4, (vsz_u2){l0, l1})};
}

template <int K, bool FWD> __device__ __forceinline__ void xform_rows(float *v) {
#pragma unroll
    for (int i = 0; i < 8; ++i) xf8<K, FWD>(v + i * 8);
}

template <int K, bool FWD> __device__ __forceinline__ void xform_group_axis(float *v) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float t[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) t[i] = v[i * 8 + c];
        xf8<K, FWD>(t);
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i * 8 + c] = t[i];
    }
}

// Forward 3-D transform (src/bm3d.cu:461-466): column layout in, row layout out.
template <int T2, int T1> __device__ __forceinline__ void forward3d(float *v, float *buf, int l8) {
    xform_rows<T2, true>(v);  // vertical pass: along j
    transpose_ranks(v, buf, l8);
    xform_rows<T2, true>(v);  // horizontal pass: along c
    xform_group_axis<T1, true>(v);
}

// Inverse (src/bm3d.cu:470-475): row layout in, row layout out.
template <int T2, int T1> __device__ __forceinline__ void inverse3d(float *v, float *buf, int l8) {
    transpose_ranks(v, buf, l8);
    xform_rows<T2, false>(v);  // vertical
    transpose_ranks(v, buf, l8);
    xform_rows<T2, false>(v);  // horizontal
    xform_group_axis<T1, false>(v);
}

__device__ __forceinline__ float bfly8(float x) {
    x += __shfl_xor(x, 1);
    x += __shfl_xor(x, 2);
    x += __shfl_xor(x, 4);
    return x;
}

// Fixed-point value of an accumulator term already scaled by 2^(24-e) (an integer, or beyond
// the clamp): the 64-bit accumulator counts units of 2^(e-24).
__device__ __forceinline__ long long fixed_from_scaled(float s) {
    s = fminf(fmaxf(s, -4.611686018427387904e18f), 4.611686018427387904e18f);
    return (long long)s;
}

__device__ __forceinline__ long long to_fixed(float q, int e) {
    return fixed_from_scaled(__builtin_ldexpf(q, 24 - e));
}

// The quantised term q = (w*v + E) - E with E = 2^e (src/bm3d.cu:884-885), computed directly
// in accumulator units as (ws*v + 2^24) - 2^24 with ws = w * 2^(24-e). Scaling by a power of two
// commutes with every rounding step here: all operands stay normal or the term rounds to zero
// in both forms, so the result is exactly q * 2^(24-e).
constexpr float kFixOne = 16777216.0f;  // 2^24 = E * 2^(24-e)
__device__ __forceinline__ float scaled_term(float ws, float v) { return (ws * v + kFixOne) - kFixOne; }

// Loads the group (and, in the final pass, the basic-estimate group) in column layout. The
// [TW][H][S] plane stack is below 2 GiB, so every offset fits the 32-bit buffer addressing.
template <bool FINAL>
__device__ __forceinline__ void gather_group(const FilterArgs &a, int p, const int *mx, const int *my, const int *mz,
                                             int c, float *v, float *r) {
    const unsigned bytes = (unsigned)(a.plane_elems * 4);
    const PlaneBuf bref = make_plane_buf(a.clip0 + (long long)p * a.plane_elems, bytes);
    const PlaneBuf bnoisy = FINAL ? make_plane_buf(a.clip1 + (long long)p * a.plane_elems, bytes) : bref;
    const int row_bytes = a.stride * 4;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int off = (int)(mz[i] * a.ts + my[i] * a.stride + mx[i] + c) * 4;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[i * 8 + j] = buf_load_s(bnoisy, off, j * row_bytes);
            if constexpr (FINAL) r[i * 8 + j] = buf_load_s(bref, off, j * row_bytes);
        }
    }
}

// 3-D transform, shrinkage (src/bm3d.cu:416-546) and inverse; returns the adaptive weight.
// Column layout in, row layout out. `r` is clobbered in the final pass.
template <bool FINAL, int T2, int T1>
__device__ __forceinline__ float filter_group(float *v, float *r, float sigma, float *buf, int l8) {
    forward3d<T2, T1>(v, buf, l8);
    float k = 0.0f;
    if constexpr (FINAL) {
        forward3d<T2, T1>(r, buf, l8);
        const float s2 = sigma * sigma;
#pragma unroll
        for (int f = 0; f < 64; ++f) {
            const float r2 = r[f] * r[f];
            float coeff = __fdiv_rn(r2, r2 + s2);
            if (f == 0 && l8 == 0) coeff = 1.0f;  // 3-D DC: rank 0, row 0, column 0
            v[f] = (v[f] * coeff) * (1.0f / 4096.0f);
            r[f] = coeff;
        }
        // k sums coeff^2 per column over (rank, row) in that order (src/bm3d.cu:491-504), so
        // the coefficients go back to column layout for the accumulation.
        transpose_ranks(r, buf, l8);
#pragma unroll
        for (int f = 0; f < 64; ++f) k = fmaf_(r[f], r[f], k);
    } else {
        // k counts the kept coefficients; as a float sum of ones it is exact, so an integer count
        // gives the same value
        int kept = 0;
#pragma unroll
        for (int f = 0; f < 64; ++f) {
            const float thr = (f == 0 && l8 == 0) ? 0.0f : sigma;
            const bool keep = fabsf(v[f]) >= thr;
            kept += keep ? 1 : 0;
            v[f] = keep ? v[f] * (1.0f / 4096.0f) : 0.0f;
        }
        k = (float)kept;
    }
    const float w = __fdiv_rn(1.0f, bfly8(k));
    inverse3d<T2, T1>(v, buf, l8);
    return w;
}

__device__ __forceinline__ void load_matches(const int4 *ml, int *mx, int *my, int *mz) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int4 m = ml[k];
        mx[k] = m.x;
        my[k] = m.y;
        mz[k] = m.z;
    }
}

// Denominator scatter: a patch adds the same quantised weight to all 64 of its pixels, so the
// kernels add it once at the patch's top-left corner; the finish kernel's exact 8x8 box sum
// turns these corner sums into the per-pixel denominator.
template <bool EXACT>
__device__ __forceinline__ void scatter_den(const FilterArgs &a, long long pb, const int *mx, const int *my,
                                            const int *mz, float qw) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const long long at = pb + (long long)mz[i] * 2 * a.ts + a.ts + (long long)my[i] * a.stride + mx[i];
        if constexpr (EXACT) atomicAdd((unsigned long long *)a.acc + at, (unsigned long long)to_fixed(qw, a.extractor_exp));
        else atomicAdd((float *)a.acc + at, qw);
    }
}

// Direct path: every numerator contribution is a device-wide atomic (temporal passes and
// search windows too large for an LDS tile).
template <bool FINAL, int T2, int T1, bool EXACT>
__device__ __forceinline__ void filter_direct_body(const FilterArgs &a, long long gbase, int lane, float (*tbuf)[72]) {
    const int g = lane >> 3, c = lane & 7;
    const long long gid = gbase + g;
    const bool valid = gid < a.nblocks;
    int mx[8], my[8], mz[8];
    load_matches(a.matches + (long long)(valid ? gid : 0) * 8, mx, my, mz);
    float *buf = tbuf[g];
    const float E = a.extractor;

    for (int p = 0; p < a.num_planes; ++p) {
        if (!((a.active_mask >> p) & 1)) continue;
        float v[64];
        float r[FINAL ? 64 : 1];
        gather_group<FINAL>(a, p, mx, my, mz, c, v, r);
        const float w = filter_group<FINAL, T2, T1>(v, r, a.sigma[p], buf, c);
        // back to column layout, so the 8 lanes of a group update 8 adjacent words of one row
        // with each atomic instruction (one cache line per group instead of eight)
        transpose_ranks(v, buf, c);
        if (!valid) continue;
        const long long pb = (long long)p * a.acc_plane_elems;
        const float ws = __builtin_ldexpf(w, 24 - a.extractor_exp);
#pragma unroll
        for (int i = 0; i < 8; ++i) {  // column layout: this lane adds column (lane & 7) of patch i
            const long long base = pb + (long long)mz[i] * 2 * a.ts + (long long)my[i] * a.stride + mx[i] + c;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if constexpr (EXACT) atomicAdd((unsigned long long *)a.acc + base + (long long)j * a.stride,
                                               (unsigned long long)fixed_from_scaled(scaled_term(ws, v[i * 8 + j])));
                else atomicAdd((float *)a.acc + base + (long long)j * a.stride, (w * v[i * 8 + j] + E) - E);
            }
        }
        if (c == 0) scatter_den<EXACT>(a, pb, mx, my, mz, (w + E) - E);
    }
}

template <bool FINAL, int T2, int T1, bool EXACT>
__global__ __launch_bounds__(64) void k_filter(FilterArgs a) {
    __shared__ float tbuf[8][72];
    filter_direct_body<FINAL, T2, T1, EXACT>(a, (long long)blockIdx.x * 8, threadIdx.x, tbuf);
}

// Adds a tile's LDS region image (rw x rh cells at rx0, ry0) to one plane's device accumulator
// and clears it, for a 256-thread workgroup: each wave takes every fourth row, its lanes
// consecutive columns, so no cell index is divided.
template <bool EXACT>
__device__ __forceinline__ void flush_region(const FilterArgs &a, long long pb, int *cell_i, float *cell_f, int rx0,
                                             int ry0, int rw, int rh) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int y = wave; y < rh; y += 4) {
        const long long row = pb + (long long)(ry0 + y) * a.stride + rx0;
        for (int x = lane; x < rw; x += 64) {
            const int i = y * rw + x;
            if constexpr (EXACT) {
                const int vv = cell_i[i];
                if (vv) atomicAdd((unsigned long long *)a.acc + row + x, (unsigned long long)(long long)vv);
                cell_i[i] = 0;
            } else {
                const float vv = cell_f[i];
                if (vv != 0.0f) atomicAdd((float *)a.acc + row + x, vv);
                cell_f[i] = 0.0f;
            }
        }
    }
}

// Tiled path (radius 0): a 256-thread workgroup owns a tbx x tby tile of reference blocks and
// accumulates the numerators of all their groups in an LDS image of the region the groups can
// reach (the tile's search windows). With a non-zero extractor the LDS cells are exact 32-bit
// fixed-point partial sums; a contribution whose magnitude could overflow them (at most
// 8*tbx*tby contributions per cell, each below 2^guard_bits) or that falls outside the region
// (fill entries of tiny windows) goes straight to the device accumulator. Each plane's region
// is flushed once with 64-bit device atomics; integer addition commutes, so the result does
// not depend on the order of any of these updates.
template <bool FINAL, int T2, int T1, bool EXACT>
__device__ __forceinline__ void filter_tile_body_vec(const FilterArgs &a, int tile_x, int tile_y, float *smem) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 3, c = lane & 7;
    float *buf = smem + (wave * 8 + g) * 72;
    int *cell_i = reinterpret_cast<int *>(smem + 4 * 8 * 72);
    float *cell_f = smem + 4 * 8 * 72;

    const int W = a.width, H = a.height, bs = a.block_step, R = a.bm_range;
    const int bx0 = tile_x * a.tbx, by0 = tile_y * a.tby;
    const int nbx_t = min(a.tbx, a.nbx - bx0), nby_t = min(a.tby, a.nby - by0);
    const int ngroups = nbx_t * nby_t;
    const int rx0 = max(min(bx0 * bs, W - 8) - R, 0);
    const int rx1 = min(min((bx0 + nbx_t - 1) * bs, W - 8) + R, W - 8) + 8;
    const int ry0 = max(min(by0 * bs, H - 8) - R, 0);
    const int ry1 = min(min((by0 + nby_t - 1) * bs, H - 8) + R, H - 8) + 8;
    const int rw = rx1 - rx0, cells = rw * (ry1 - ry0);
    const float guard_f = __builtin_ldexpf(1.0f, a.guard_bits);
    const float E = a.extractor;

    for (int i = threadIdx.x; i < cells; i += 256) cell_i[i] = 0;
    __syncthreads();

    for (int p = 0; p < a.num_planes; ++p) {
        if (!((a.active_mask >> p) & 1)) continue;
        const long long pb = (long long)p * a.acc_plane_elems;
        for (int base = wave * 8; base < ngroups; base += 32) {
            const int gi = base + g;
            const bool valid = gi < ngroups;
            const int gs = valid ? gi : 0;
            const long long gid = (long long)(by0 + gs / nbx_t) * a.nbx + bx0 + gs % nbx_t;
            int mx[8], my[8], mz[8];
            load_matches(a.matches + gid * 8, mx, my, mz);
#pragma unroll
            for (int k = 0; k < 8; ++k) mz[k] = 0;  // radius 0: every match is in slot 0
            float v[64];
            float r[FINAL ? 64 : 1];
            gather_group<FINAL>(a, p, mx, my, mz, c, v, r);
            const float w = filter_group<FINAL, T2, T1>(v, r, a.sigma[p], buf, c);
            if (!valid) continue;
            const float ws = __builtin_ldexpf(w, 24 - a.extractor_exp);
#pragma unroll
            for (int i = 0; i < 8; ++i) {  // row layout: this lane adds row (lane & 7) of patch i
                const bool inside = mx[i] >= rx0 && mx[i] + 8 <= rx1 && my[i] >= ry0 && my[i] + 8 <= ry1;
                const int cell0 = (my[i] + c - ry0) * rw + mx[i] - rx0;
                const long long gbase = pb + (long long)(my[i] + c) * a.stride + mx[i];
                if constexpr (EXACT) {
                    // Terms in accumulator units are integers; below the guard they convert
                    // exactly to int32. The common case tests the row's largest magnitude once
                    // (a NaN-propagating maximum, so a NaN term takes the per-term path).
                    float sq[8];
                    float big = 0.0f;
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        sq[cc] = scaled_term(ws, v[i * 8 + cc]);
                        big = __builtin_elementwise_maximum(big, fabsf(sq[cc]));
                    }
                    if (inside && big < guard_f) {
#pragma unroll
                        for (int cc = 0; cc < 8; ++cc) atomicAdd(cell_i + cell0 + cc, (int)sq[cc]);
                    } else {
#pragma unroll
                        for (int cc = 0; cc < 8; ++cc) {
                            if (inside && fabsf(sq[cc]) < guard_f)
                                atomicAdd(cell_i + cell0 + cc, (int)sq[cc]);
                            else
                                atomicAdd((unsigned long long *)a.acc + gbase + cc,
                                          (unsigned long long)fixed_from_scaled(sq[cc]));
                        }
                    }
                } else {
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        const float q = (w * v[i * 8 + cc] + E) - E;
                        if (inside) atomicAdd(cell_f + cell0 + cc, q);
                        else atomicAdd((float *)a.acc + gbase + cc, q);
                    }
                }
            }
            if (c == 0) scatter_den<EXACT>(a, pb, mx, my, mz, (w + E) - E);
        }
        __syncthreads();
        flush_region<EXACT>(a, pb, cell_i, cell_f, rx0, ry0, rw, ry1 - ry0);
        __syncthreads();
    }
}

// ---- group-pair filtering on MFMA (split-BF16 DCT stages) -------------------------------------
// A wave filters two groups (g = 0, 1) at once. Their 3-D DCT runs as six 8-point stages, each a
// product with the 16x16 block-diagonal matrix diag(X, X) on v_mfma_f32_16x16x16_bf16. Operand
// layouts of that instruction (checked on the GPU with small-integer operands): lane l, q = l/16,
// holds A[l%16][4q..4q+3], B[4q..4q+3][l%16] and D[4q..4q+3][l%16]. A stage multiplies the data
// from the right (data in the A layout, contraction over its columns) or from the left (data in
// the B layout, which is the D layout of the stage before, contraction over its rows), so
// consecutive stages chain in registers; where the contracted axis moves to the other side, the
// data goes once through a per-wave LDS buffer. Indices: g group, i rank, j row, c column.
//   forward: F1 right (c), F2 left (j), LDS, F3 left (i);  inverse: I3 left (i), LDS, I1 right (c),
//   I2 right (j).
// A lane's D fragment of a product P is its A fragment of P^T (D[4q+e][l%16] = P^T[l%16][4q+e]),
// so I2 multiplies I1's result from the right as the transpose: its rows become (g, c) and the
// contraction runs over (i parity, j); each lane ends with four adjacent pixels of one row.
// Each product is compensated split-BF16 (mfma_split). The lane's share of diag(X, X) is the same
// 4 values whether it is the A operand of a left
// product or the B operand of a right product: (q/2 == (l%16)/8) ? X[l%8][4(q%2) + e] : 0, with
// X = M forward and X = M^T inverse (X = M^T / 4096 in I3, which carries the 3-D round trip's
// 1/4096 of the shrinkage: a power of two, so the scaling is exact and in the same place as a
// separate multiply would be).
constexpr int kPairSJ = 8, kPairSI = 68, kPairSG = 552;  // LDS strides (floats), at most 2-way bank conflicts
constexpr int kPairBufFloats = 2 * kPairSG;               // per wave; a multiple of 4 (16-byte rows)

__device__ __forceinline__ vsz_f4 mfma16(vsz_s4 a, vsz_s4 b, vsz_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Compensated product of split operands: Al*Bh + Ah*Bl + Ah*Bh, FP32 accumulation.
__device__ __forceinline__ vsz_f4 mfma_split(const SplitOp &A, const SplitOp &B) {
    vsz_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma16(A.lo, B.hi, acc);  // small cross terms first
    acc = mfma16(A.hi, B.lo, acc);
    return mfma16(A.hi, B.hi, acc);
}

__device__ __forceinline__ SplitOp split_f4(vsz_f4 v) { return split4(v[0], v[1], v[2], v[3]); }

struct PairConsts {
    SplitOp fwd, inv, inv_s;  // M, M^T, M^T / 4096
};

__device__ __forceinline__ PairConsts pair_consts(int l) {
    const int q = l >> 4, r = l & 7, e0 = 4 * (q & 1);
    const bool same = (q >> 1) == ((l & 15) >> 3);
    float f[4], v[4], vs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        f[e] = same ? kDct8x4[r * 8 + e0 + e] : 0.0f;    // M[r][e0 + e]
        v[e] = same ? kDct8x4[(e0 + e) * 8 + r] : 0.0f;  // M^T[r][e0 + e]
        vs[e] = v[e] * (1.0f / 4096.0f);
    }
    return PairConsts{split4(f[0], f[1], f[2], f[3]), split4(v[0], v[1], v[2], v[3]), split4(vs[0], vs[1], vs[2], vs[3])};
}

// Forward 3-D DCT of a group pair. xa[t] holds this lane's row of the F1 operand for rank pair
// t: pixels X[g = q/2][i = 2t + (l%16)/8][j = l%8][c = 4(q%2) .. +3]. Returns z[s][e] =
// Z[g = q/2][i = 4(q%2) + e][j = 2s + (l%16)/8][c = l%8].
__device__ __forceinline__ void pair_forward(const SplitOp (&xa)[4], const PairConsts &K, float *tb, int l,
                                             vsz_f4 (&z)[4]) {
    const int q = l >> 4, h = (l & 15) >> 3, c = l & 7;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const vsz_f4 d1 = mfma_split(xa[t], K.fwd);         // F1: rows along c
        const vsz_f4 d2 = mfma_split(K.fwd, split_f4(d1));  // F2: columns along j
        // d2[e] = X[g = h][i = 2t + q/2][j = 4(q%2) + e][c]
        float *w = tb + h * kPairSG + (2 * t + (q >> 1)) * kPairSI + 4 * (q & 1) * kPairSJ + c;
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e * kPairSJ] = d2[e];
    }
    lds_order();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float *r = tb + (q >> 1) * kPairSG + 4 * (q & 1) * kPairSI + (2 * s + h) * kPairSJ + c;
        vsz_f4 b;
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = r[e * kPairSI];
        z[s] = mfma_split(K.fwd, split_f4(b));               // F3: along the group axis i
    }
    lds_order();
}

// Inverse of pair_forward, including the factor 1/4096. zz[s] in the layout pair_forward
// returns. ROW_OUT: out[t][e] = pixel Y[g = q/2][i = 2t + (l%16)/8][j = l%8][c = 4(q%2) + e], the
// positions pair_load read (four adjacent pixels of a row per lane, for the LDS region image);
// otherwise I2 is a left product and out[t][e] = Y[g = (l%16)/8][i = 2t + q/2][j = 4(q%2) + e]
// [c = l%8] (the 8 lanes of a row hold 8 adjacent pixels, so each device atomic instruction
// updates whole row segments).
template <bool ROW_OUT>
__device__ __forceinline__ void pair_inverse(const vsz_f4 (&zz)[4], const PairConsts &K, float *tb, int l,
                                             vsz_f4 (&out)[4]) {
    const int q = l >> 4, h = (l & 15) >> 3, c = l & 7;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const vsz_f4 v3 = mfma_split(K.inv_s, split_f4(zz[s]));  // I3: along i, with 1/4096
        float *w = tb + (q >> 1) * kPairSG + 4 * (q & 1) * kPairSI + (2 * s + h) * kPairSJ + c;
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e * kPairSI] = v3[e];
    }
    lds_order();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        // row of the I1 operand: V[g = q/2][i = 2t + (l%16)/8][j = l%8][c = 4(q%2) .. +3]
        const vsz_f4 a = *reinterpret_cast<const vsz_f4 *>(tb + (q >> 1) * kPairSG + (2 * t + h) * kPairSI +
                                                            c * kPairSJ + 4 * (q & 1));
        const vsz_f4 v1 = mfma_split(split_f4(a), K.inv);    // I1: rows along c
        if constexpr (ROW_OUT) out[t] = mfma_split(split_f4(v1), K.inv);  // I2: along j, on I1's result transposed
        else out[t] = mfma_split(K.inv, split_f4(v1));                    // I2: columns along j
    }
    lds_order();
}

// Sum over the 32 lanes of one group (lanes 32g .. 32g + 31), in a fixed order.
__device__ __forceinline__ float half_wave_sum(float x) {
#pragma unroll
    for (int d = 1; d < 32; d *= 2) x += __shfl_xor(x, d);
    return x;
}

__device__ __forceinline__ int half_wave_sum(int x) {
#pragma unroll
    for (int d = 1; d < 32; d *= 2) x += __shfl_xor(x, d);
    return x;
}

// The F1 operand rows of a group pair (and of its basic estimate in the final pass), loaded
// with one 16-byte load per rank pair from the plane's buffer: v[t] = X[g = q/2][i = 2t + (l%16)/8]
// [j = l%8][c = 4(q%2) .. +3] at patch (px[t], py[t], pz[t]).
template <bool FINAL>
__device__ __forceinline__ void pair_load(const FilterArgs &a, int p, const int (&px)[4], const int (&py)[4],
                                          const int (&pz)[4], int l, vsz_f4 (&v)[4], vsz_f4 (&r)[4]) {
    const unsigned bytes = (unsigned)(a.plane_elems * 4);
    const PlaneBuf bref = make_plane_buf(a.clip0 + (long long)p * a.plane_elems, bytes);
    const PlaneBuf bnoisy = FINAL ? make_plane_buf(a.clip1 + (long long)p * a.plane_elems, bytes) : bref;
    const int j = l & 7, c0 = 4 * ((l >> 4) & 1);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int off = (int)(pz[t] * a.ts + (long long)(py[t] + j) * a.stride + px[t] + c0) * 4;
        v[t] = __builtin_bit_cast(vsz_f4, __builtin_amdgcn_raw_buffer_load_b128(bnoisy.rsrc, off, 0, 0));
        if constexpr (FINAL) r[t] = __builtin_bit_cast(vsz_f4, __builtin_amdgcn_raw_buffer_load_b128(bref.rsrc, off, 0, 0));
    }
}

// Shrinkage of a group pair's coefficients (src/bm3d.cu:416-546): hard threshold, or Wiener
// with the basic estimate's coefficients r; the 1/4096 of the kept values is applied by
// pair_inverse. Returns the weight 1/k of the lane's group (q/2).
template <bool FINAL>
__device__ __forceinline__ float pair_shrink(vsz_f4 (&z)[4], const vsz_f4 (&r)[4], float sigma, int l) {
    // the 3-D DC (i = 0, j = 0, c = 0) sits in element 0 of N-tile 0 of lanes 0 and 32
    const bool dc_lane = (l & 31) == 0;
    if constexpr (FINAL) {
        const float s2 = sigma * sigma;
        float k = 0.0f;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float r2 = r[s][e] * r[s][e];
                // reciprocal and one Newton step (within an ulp of the division; the split form is
                // held to error bounds, not bits, and the reference itself divides with fast math)
                const float den = r2 + s2;
                float rc = __builtin_amdgcn_rcpf(den);
                rc = fmaf(fmaf(-den, rc, 1.0f), rc, rc);
                float coeff = r2 * rc;
                if (s == 0 && e == 0 && dc_lane) coeff = 1.0f;
                z[s][e] = z[s][e] * coeff;
                k = fmaf_(coeff, coeff, k);
            }
        return __fdiv_rn(1.0f, half_wave_sum(k));
    } else {
        int kept = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float thr = (s == 0 && e == 0 && dc_lane) ? 0.0f : sigma;
                const bool keep = fabsf(z[s][e]) >= thr;
                kept += keep ? 1 : 0;
                z[s][e] = keep ? z[s][e] : 0.0f;
            }
        return __fdiv_rn(1.0f, (float)half_wave_sum(kept));
    }
}

// One group pair, loaded pixels (pair_load) to filtered pixels out[t][e] in pair_inverse's
// ROW_OUT layout; returns the weight of the group of the lane's output pixels (q/2 with ROW_OUT,
// else (l%16)/8).
template <bool FINAL, bool ROW_OUT>
__device__ __forceinline__ float pair_filter(const vsz_f4 (&v)[4], const vsz_f4 (&vr)[4], float sigma,
                                             const PairConsts &K, float *tb, int l, vsz_f4 (&out)[4]) {
    SplitOp xv[4], xr[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        xv[t] = split_f4(v[t]);
        if constexpr (FINAL) xr[t] = split_f4(vr[t]);
    }
    vsz_f4 z[4], r[4];
    pair_forward(xv, K, tb, l, z);
    if constexpr (FINAL) pair_forward(xr, K, tb, l, r);
    const float w = pair_shrink<FINAL>(z, r, sigma, l);
    pair_inverse<ROW_OUT>(z, K, tb, l, out);
    if constexpr (ROW_OUT) return w;
    else return __shfl(w, ((l & 15) >> 3) * 32);
}

// Positions of a pair's patches in the lane's layout: (g = q/2, i = 2t + (l%16)/8).
__device__ __forceinline__ void pair_positions(const int4 *m0, const int4 *m1, int l, int (&px)[4], int (&py)[4],
                                               int (&pz)[4]) {
    const int q = l >> 4, h = (l & 15) >> 3;
    const int4 *ma = (q >> 1) ? m1 : m0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int4 m = ma[2 * t + h];
        px[t] = m.x;
        py[t] = m.y;
        pz[t] = m.z;
    }
}

// The same for the tiled path (radius 0, every match in slot 0): x and y only, one 8-byte load each.
__device__ __forceinline__ void pair_positions_xy(const int4 *m0, const int4 *m1, int l, int (&px)[4], int (&py)[4]) {
    const int q = l >> 4, h = (l & 15) >> 3;
    const int4 *ma = (q >> 1) ? m1 : m0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int2 m = *reinterpret_cast<const int2 *>(ma + 2 * t + h);
        px[t] = m.x;
        py[t] = m.y;
    }
}

// Tiled path on MFMA: the tiles, LDS region image, guard and flush of filter_tile_body_vec; four
// waves per workgroup, each filtering two groups at a time.
template <bool FINAL, bool EXACT>
__device__ __forceinline__ void filter_tile_body_mfma(const FilterArgs &a, int tile_x, int tile_y, float *smem) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, q = l >> 4, j = l & 7, c0 = 4 * (q & 1);
    float *tb = smem + wave * kPairBufFloats;
    int *cell_i = reinterpret_cast<int *>(smem + 4 * kPairBufFloats);
    float *cell_f = smem + 4 * kPairBufFloats;

    const int W = a.width, H = a.height, bs = a.block_step, R = a.bm_range;
    const int bx0 = tile_x * a.tbx, by0 = tile_y * a.tby;
    const int nbx_t = min(a.tbx, a.nbx - bx0), nby_t = min(a.tby, a.nby - by0);
    const int ngroups = nbx_t * nby_t;
    const int rx0 = max(min(bx0 * bs, W - 8) - R, 0);
    const int rx1 = min(min((bx0 + nbx_t - 1) * bs, W - 8) + R, W - 8) + 8;
    const int ry0 = max(min(by0 * bs, H - 8) - R, 0);
    const int ry1 = min(min((by0 + nby_t - 1) * bs, H - 8) + R, H - 8) + 8;
    const int rw = rx1 - rx0, cells = rw * (ry1 - ry0);
    const float guard_f = __builtin_ldexpf(1.0f, a.guard_bits);
    const float E = a.extractor;
    const PairConsts K = pair_consts(l);

    for (int i = threadIdx.x; i < cells; i += 256) cell_i[i] = 0;
    __syncthreads();

    // positions of the pair at group index `base` of the tile (its second group clamped to the
    // tile, and a pair past the tile reads the last group: in bounds, never used)
    const auto positions = [&](int base, int (&px)[4], int (&py)[4]) {
        const int g0 = min(base, ngroups - 1), g1 = min(base + 1, ngroups - 1);
        const long long gid0 = (long long)(by0 + g0 / nbx_t) * a.nbx + bx0 + g0 % nbx_t;
        const long long gid1 = (long long)(by0 + g1 / nbx_t) * a.nbx + bx0 + g1 % nbx_t;
        pair_positions_xy(a.matches + gid0 * 8, a.matches + gid1 * 8, l, px, py);
    };
    const int zero4[4] = {0, 0, 0, 0};  // radius 0: every match is in slot 0

    for (int p = 0; p < a.num_planes; ++p) {
        if (!((a.active_mask >> p) & 1)) continue;
        const long long pb = (long long)p * a.acc_plane_elems;
        int base = wave * 2;
        if (base < ngroups) {  // (no early exit: every wave reaches the barriers below)
            // software pipeline: the positions of the next pair are in flight while this pair is
            // filtered (pixel prefetch as well costs more occupancy than it hides)
            int cx[4], cy[4], nx[4], ny[4];
            positions(base, cx, cy);
            for (;;) {
                const bool more = base + 8 < ngroups;
                vsz_f4 cv[4], cr[4];
                pair_load<FINAL>(a, p, cx, cy, zero4, l, cv, cr);
                if (more) positions(base + 8, nx, ny);
                vsz_f4 out[4];
                const float w = pair_filter<FINAL, true>(cv, cr, a.sigma[p], K, tb, l, out);
                if (base + (q >> 1) < ngroups) {  // else the pair's second group is past the tile
                    const float ws = __builtin_ldexpf(w, 24 - a.extractor_exp);
                    const float wq = (w + E) - E;
                    const unsigned long long wfix = EXACT ? (unsigned long long)to_fixed(wq, a.extractor_exp) : 0ull;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int mx = cx[t], my = cy[t];
                        const bool inside = mx >= rx0 && mx + 8 <= rx1 && my >= ry0 && my + 8 <= ry1;
                        const int cell0 = (my + j - ry0) * rw + mx + c0 - rx0;
                        const long long at0 = pb + (long long)(my + j) * a.stride + mx + c0;
                        if constexpr (EXACT) {
                            // the common case tests the four terms' largest magnitude once (a
                            // NaN-propagating maximum, so a NaN term takes the per-term path)
                            float sq[4];
                            float big = 0.0f;
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                sq[e] = scaled_term(ws, out[t][e]);
                                big = __builtin_elementwise_maximum(big, fabsf(sq[e]));
                            }
                            if (inside && big < guard_f) {
#pragma unroll
                                for (int e = 0; e < 4; ++e) atomicAdd(cell_i + cell0 + e, (int)sq[e]);
                            } else {
#pragma unroll
                                for (int e = 0; e < 4; ++e) {
                                    if (inside && fabsf(sq[e]) < guard_f) atomicAdd(cell_i + cell0 + e, (int)sq[e]);
                                    else atomicAdd((unsigned long long *)a.acc + at0 + e, (unsigned long long)fixed_from_scaled(sq[e]));
                                }
                            }
                        } else {
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const float qv = (w * out[t][e] + E) - E;
                                if (inside) atomicAdd(cell_f + cell0 + e, qv);
                                else atomicAdd((float *)a.acc + at0 + e, qv);
                            }
                        }
                        // weights: one lane per patch adds it at the top-left corner
                        if (j == 0 && (q & 1) == 0) {
                            const long long at = pb + a.ts + (long long)my * a.stride + mx;
                            if constexpr (EXACT) atomicAdd((unsigned long long *)a.acc + at, wfix);
                            else atomicAdd((float *)a.acc + at, wq);
                        }
                    }
                }
                if (!more) break;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    cx[t] = nx[t];
                    cy[t] = ny[t];
                }
                base += 8;
            }
        }
        __syncthreads();
        flush_region<EXACT>(a, pb, cell_i, cell_f, rx0, ry0, rw, ry1 - ry0);
        __syncthreads();
    }
}

template <bool FINAL, int T2, int T1, bool EXACT>
__device__ __forceinline__ void filter_tile_body(const FilterArgs &a, int tile_x, int tile_y, float *smem) {
    if constexpr (T2 == kDctMfma) filter_tile_body_mfma<FINAL, EXACT>(a, tile_x, tile_y, smem);
    else filter_tile_body_vec<FINAL, T2, T1, EXACT>(a, tile_x, tile_y, smem);
}

// Direct path on MFMA (temporal passes, regions above the LDS budget): two groups per wave,
// device-wide atomics; the 8 lanes of a row update 8 adjacent words.
template <bool FINAL, bool EXACT>
__global__ __launch_bounds__(64) void k_filter_mfma(FilterArgs a) {
    __shared__ __align__(16) float tbuf[kPairBufFloats];
    const int l = threadIdx.x, q = l >> 4, h = (l & 15) >> 3, c = l & 7;
    const long long gid0 = min((long long)blockIdx.x * 2, a.nblocks - 1);
    const long long gid1 = min((long long)blockIdx.x * 2 + 1, a.nblocks - 1);
    const bool valid = (long long)blockIdx.x * 2 + h < a.nblocks;  // the output group is (l%16)/8
    int px[4], py[4], pz[4];
    pair_positions(a.matches + gid0 * 8, a.matches + gid1 * 8, l, px, py, pz);
    int4 po[4];  // output patches: (g = (l%16)/8, i = 2t + q/2)
    {
        const int4 *mb = h ? a.matches + gid1 * 8 : a.matches + gid0 * 8;
#pragma unroll
        for (int t = 0; t < 4; ++t) po[t] = mb[2 * t + (q >> 1)];
    }
    const PairConsts K = pair_consts(l);
    const float E = a.extractor;
    for (int p = 0; p < a.num_planes; ++p) {
        if (!((a.active_mask >> p) & 1)) continue;
        vsz_f4 v[4], vr[4], out[4];
        pair_load<FINAL>(a, p, px, py, pz, l, v, vr);
        const float w = pair_filter<FINAL, false>(v, vr, a.sigma[p], K, tbuf, l, out);
        if (!valid) continue;
        const long long pb = (long long)p * a.acc_plane_elems;
        const float ws = __builtin_ldexpf(w, 24 - a.extractor_exp);
        const float wq = (w + E) - E;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const long long pbase = pb + (long long)po[t].z * 2 * a.ts + (long long)po[t].y * a.stride + po[t].x;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const long long at = pbase + (long long)(4 * (q & 1) + e) * a.stride + c;
                if constexpr (EXACT)
                    atomicAdd((unsigned long long *)a.acc + at, (unsigned long long)fixed_from_scaled(scaled_term(ws, out[t][e])));
                else atomicAdd((float *)a.acc + at, (w * out[t][e] + E) - E);
            }
            if (c == 0 && (q & 1) == 0) {
                if constexpr (EXACT)
                    atomicAdd((unsigned long long *)a.acc + pbase + a.ts, (unsigned long long)to_fixed(wq, a.extractor_exp));
                else atomicAdd((float *)a.acc + pbase + a.ts, wq);
            }
        }
    }
}

// Waves per SIMD the registers must allow: four for split-BF16 tiles, three for the FP32 hard
// pass (168 VGPRs, so three workgroups share a CU), two for the FP32 Wiener pass (two groups of
// 64 values per lane).
template <bool FINAL, int T2, int T1, bool EXACT>
__global__ __launch_bounds__(256, T2 == kDctMfma ? 4 : (FINAL ? 2 : 3)) void k_filter_tile(FilterArgs a) {
    extern __shared__ float smem[];
    // Wiener tiles in XCD-contiguous order (consecutive tiles on one XCD share their search halos
    // of the ref and noisy planes in that XCD's L2): 0.772 -> 0.765 ms for the 1080p bench frame's
    // Wiener pass; the hard pass measured no difference and keeps row-major order.
    const int tiles_x = (a.nbx + a.tbx - 1) / a.tbx;
    const int tile = FINAL ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    filter_tile_body<FINAL, T2, T1, EXACT>(a, tile % tiles_x, tile / tiles_x, smem);
}

// ------------------------------------------------------------------------------------------------
// Finishing: numerator as accumulated, denominator as the exact 8x8 box sum of the corner
// scatter, each rounded once to FP32; then either the IEEE division (src/bm3d.cu:898-922) or
// the stacked [P][TW][2][H][S] payload. One workgroup per 64x16 output tile.
// ------------------------------------------------------------------------------------------------
template <bool EXACT>
__device__ __forceinline__ void finish_body(const FinishArgs &a, int tile_x, int tile_y, int pz,
                                            typename std::conditional<EXACT, long long, float>::type (*sD)[72],
                                            typename std::conditional<EXACT, long long, float>::type (*sH)[64]) {
    using T = typename std::conditional<EXACT, long long, float>::type;
    const int x0 = tile_x * 64, y0 = tile_y * 16;
    const int p = pz / a.tw;
    if (a.out_kind == 0 && !((a.proc_mask >> p) & 1)) return;
    const long long base = (long long)pz * 2 * a.ts;
    const T *num = reinterpret_cast<const T *>(a.acc) + base;
    const T *D = num + a.ts;
    for (int idx = threadIdx.x; idx < 23 * 71; idx += 256) {
        const int rr = idx / 71, cc = idx % 71;
        const int y = y0 - 7 + rr, x = x0 - 7 + cc;
        sD[rr][cc] = (y >= 0 && y < a.height && x >= 0 && x < a.width) ? D[(long long)y * a.stride + x] : T(0);
    }
    __syncthreads();
    // separable 8x8 box sum: horizontal 8-sums of the 23 staged rows, then vertical 8-sums
    for (int idx = threadIdx.x; idx < 23 * 64; idx += 256) {
        const int rr = idx >> 6, cc = idx & 63;
        T h = T(0);
#pragma unroll
        for (int dx = 0; dx < 8; ++dx) h += sD[rr][cc + dx];
        sH[rr][cc] = h;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int px = threadIdx.x & 63, py = (threadIdx.x >> 6) + 4 * k;
        const int x = x0 + px, y = y0 + py;
        if (x >= a.width || y >= a.height) continue;
        T s = T(0);
#pragma unroll
        for (int dy = 0; dy < 8; ++dy) s += sH[py + dy][px];
        const long long i = (long long)y * a.stride + x;
        float fn, fd;
        if constexpr (EXACT) {
            fn = __builtin_ldexpf((float)num[i], a.extractor_exp - 24);
            fd = __builtin_ldexpf((float)s, a.extractor_exp - 24);
        } else {
            fn = num[i];
            fd = s;
        }
        if (a.out_kind == 0) {
            a.out[(long long)p * a.ts + i] = __fdiv_rn(fn, fd);
        } else {
            a.out[base + i] = fn;
            a.out[base + a.ts + i] = fd;
        }
    }
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_finish(FinishArgs a) {
    __shared__ typename std::conditional<EXACT, long long, float>::type sD[23][72];
    __shared__ typename std::conditional<EXACT, long long, float>::type sH[23][64];
    finish_body<EXACT>(a, blockIdx.x, blockIdx.y, blockIdx.z, sD, sH);
}

// Temporal aggregation (src/bm3d_vagg.cu:18-48): ascending FP32 sums, one IEEE division.
__global__ __launch_bounds__(256) void k_vaggregate(VaggArgs a) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.width || y >= a.height) return;
    const long long i = (long long)y * a.stride + x;
    float sv = 0.0f, sw = 0.0f;
    for (int t = 0; t < a.tw; ++t) {
        const float *base = a.slots[t] + (long long)a.z[t] * 2 * a.ts;
        sv = sv + base[i];
        sw = sw + base[a.ts + i];
    }
    a.dst[i] = __fdiv_rn(sv, sw);
}

// ------------------------------------------------------------------------------------------------
// Batched planes: one launch per stage covers every plane of a batch. Workgroup g belongs to
// the plane p with off[p] <= g < off[p+1]; its arguments come from a device-side table.
// ------------------------------------------------------------------------------------------------
constexpr int kMaxBatch = 256;

struct BatchTable {
    int n;
    int match_off[kMaxBatch + 1];
    int filter_off[kMaxBatch + 1];
    int finish_off[kMaxBatch + 1];
    int filter_tiles_x[kMaxBatch];
    int finish_tiles_x[kMaxBatch];
    int finish_tiles_y[kMaxBatch];
    MatchArgs m[kMaxBatch];
    FilterArgs f[kMaxBatch];
    FinishArgs fin[kMaxBatch];
};

// Largest p with off[p] <= g (off ascending, off[0] = 0).
__device__ __forceinline__ int batch_plane(const int *off, int n, int g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The batch matchers also clear the pass's accumulator span, zero[0, zero_n64) (see zero_range).
template <int M>
__global__ __launch_bounds__(256) void k_match_lds_batch(const BatchTable *__restrict__ t, unsigned long long *zero,
                                                         long long zero_n64) {
    extern __shared__ float region[];
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    const int p = batch_plane(t->match_off, t->n, g);
    const MatchArgs &a = t->m[p];
    match_lds_body<M>(a, g - t->match_off[p], region);
    zero_range(zero, zero_n64);
}

template <int M>
__global__ __launch_bounds__(256) void k_match_lds4_batch(const BatchTable *__restrict__ t, unsigned long long *zero,
                                                          long long zero_n64) {
    extern __shared__ float region[];
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    const int p = batch_plane(t->match_off, t->n, g);
    const MatchArgs &a = t->m[p];
    match_lds4_body<M>(a, g - t->match_off[p], region);
    zero_range(zero, zero_n64);
}

template <bool FINAL, int T2, int T1, bool EXACT>
__global__ __launch_bounds__(256, T2 == kDctMfma ? 4 : 2) void k_filter_tile_batch(const BatchTable *__restrict__ t) {
    extern __shared__ float smem[];
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    const int p = batch_plane(t->filter_off, t->n, g);
    const FilterArgs &a = t->f[p];
    const int local = g - t->filter_off[p], tx = t->filter_tiles_x[p];
    filter_tile_body<FINAL, T2, T1, EXACT>(a, local % tx, local / tx, smem);
}

// Windows beyond the LDS budgets: the global-memory matcher (each wave one 8 x 8-block tile) and
// the direct-atomic filter (each wave 8 groups), over a batch.
template <int M>
__global__ __launch_bounds__(256) void k_match_global_batch(const BatchTable *__restrict__ t,
                                                            unsigned long long *zero, long long zero_n64) {
    const int g = blockIdx.x;
    const int p = batch_plane(t->match_off, t->n, g);
    const MatchArgs &a = t->m[p];
    const int tile = (g - t->match_off[p]) * 4 + (threadIdx.x >> 6);
    if (tile < a.tiles_x * ((a.nby + 7) / 8)) match_global_body<M>(a, tile, threadIdx.x & 63);
    zero_range(zero, zero_n64);
}

template <bool FINAL, int T2, int T1, bool EXACT>
__global__ __launch_bounds__(256) void k_filter_direct_batch(const BatchTable *__restrict__ t) {
    __shared__ float tbuf[4][8][72];
    const int g = blockIdx.x, wave = threadIdx.x >> 6;
    const int p = batch_plane(t->filter_off, t->n, g);
    filter_direct_body<FINAL, T2, T1, EXACT>(t->f[p], ((long long)(g - t->filter_off[p]) * 4 + wave) * 8,
                                            threadIdx.x & 63, tbuf[wave]);
}

template <bool EXACT>
__global__ __launch_bounds__(256) void k_finish_batch(const BatchTable *__restrict__ t) {
    __shared__ typename std::conditional<EXACT, long long, float>::type sD[23][72];
    __shared__ typename std::conditional<EXACT, long long, float>::type sH[23][64];
    const int g = blockIdx.x;
    const int p = batch_plane(t->finish_off, t->n, g);
    const FinishArgs &a = t->fin[p];
    const int local = g - t->finish_off[p], tx = t->finish_tiles_x[p], ty = t->finish_tiles_y[p];
    finish_body<EXACT>(a, local % tx, (local / tx) % ty, local / (tx * ty), sD, sH);
}

// ---- the whole batch in one persistent launch --------------------------------------------------
// Work items (accumulator zeroing chunks, matcher tiles, filter tiles and finishing tiles of every
// plane and pass) are handed out in one fixed order from a device counter. Before it runs, an item
// waits until the items whose results it reads are complete, counted per (pass, stage, plane):
//   filter <- this pass's match and zeroing of the plane;  finish <- this pass's filtering;
//   pass 1 match and zeroing <- pass 0 finishing of the plane (its basic estimate, its
//   accumulator). Those items always come earlier in the order, so the earliest unfinished item can
// run and the launch drains with any number of resident workgroups; a plane's stages overlap with
// other planes' earlier stages. Completion: every wave waits for its own writes before the
// item's closing barrier, then lane 0 publishes with an agent-scope release increment (which writes
// back the XCD's L2); a wait polls with relaxed coherent loads and lane 0 acquires once (its CU's
// caches serve the whole workgroup), so neither polling nor the other waves invalidate caches.
enum : int { kStageZero = 0, kStageMatch = 1, kStageFilter = 2, kStageFinish = 3 };
constexpr int kZeroChunk = 256 * 128;  // 64-bit words per zeroing item (256 KB)
constexpr int kFinishPer = 8;          // finishing tiles (64 x 16 pixels) per item
constexpr int kMaxSegs = 2 * 4 * kMaxBatch;
constexpr int kSpinLimit = 1 << 24;  // polls of 256 cycles: a bound far above any legitimate wait

struct BatchSeg {
    int begin;  // first item of the segment (items of one pass, stage and plane)
    int pass, stage, plane;
};

struct PersistTable {
    int nseg, total;
    BatchSeg seg[kMaxSegs];
    int count[2][4][kMaxBatch];  // items per (pass, stage, plane)
    unsigned long long *acc[2][kMaxBatch];  // per pass and plane
    long long acc_words[kMaxBatch];
};

struct BatchQueue {
    unsigned next, exited;
    unsigned timeouts;  // waits that reached kSpinLimit (never reset by the kernel)
    unsigned done[2][4][kMaxBatch];
};

// Polls with relaxed coherent loads (no cache invalidation per poll); the caller acquires once.
__device__ __forceinline__ void await_count(const unsigned *cnt, int need, unsigned *timeouts) {
    for (int spin = 0; spin < kSpinLimit; ++spin) {
        if (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned)need) return;
        __builtin_amdgcn_s_sleep(4);
    }
    atomicAdd(timeouts, 1u);
}

// Loop shape: one `threadIdx.x == 0` block per iteration (publish the previous item, claim the
// next), and the item index read back as a wave-uniform scalar, so that the loop exit and the
// stage dispatch are uniform branches and every barrier is reached by all lanes of every wave.
// (With the publication at the end of an iteration and the claim at its start, the compiler
// joined the two lane-0 blocks across the back edge into an outer loop of their own, and the
// other lanes re-entered the item loop without a new claim.)
template <int M, bool SC, int T>
__global__ __launch_bounds__(256, 2) void k_batch_persistent(const BatchTable *__restrict__ tabs,
                                                          const PersistTable *__restrict__ pt, BatchQueue *q) {
    // one dynamic LDS block: 16 bytes of loop control, then the stages' 16-byte aligned memory
    extern __shared__ __align__(16) float lds[];
    int &s_item = reinterpret_cast<int *>(lds)[0];
    int &s_last = reinterpret_cast<int *>(lds)[1];
    float *smem = lds + 4;
    int prev_pass = -1, prev_stage = 0, prev_plane = 0;  // the item this workgroup finished last
    for (;;) {
        if (threadIdx.x == 0) {
            if (prev_pass >= 0)  // every thread released its writes before the barrier that ended it
                __hip_atomic_fetch_add(&q->done[prev_pass][prev_stage][prev_plane], 1u, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
            s_item = (int)atomicAdd(&q->next, 1u);
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= pt->total) break;
        int lo = 0, hi = pt->nseg - 1;  // last segment with begin <= it
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pt->seg[mid].begin <= it) lo = mid;
            else hi = mid - 1;
        }
        const int local = it - __builtin_amdgcn_readfirstlane(pt->seg[lo].begin);
        const int pass = __builtin_amdgcn_readfirstlane(pt->seg[lo].pass);
        const int stage = __builtin_amdgcn_readfirstlane(pt->seg[lo].stage);
        const int p = __builtin_amdgcn_readfirstlane(pt->seg[lo].plane);
        if (threadIdx.x == 0) {
            if (stage == kStageFilter) {
                await_count(&q->done[pass][kStageMatch][p], pt->count[pass][kStageMatch][p], &q->timeouts);
                await_count(&q->done[pass][kStageZero][p], pt->count[pass][kStageZero][p], &q->timeouts);
            } else if (stage == kStageFinish) {
                await_count(&q->done[pass][kStageFilter][p], pt->count[pass][kStageFilter][p], &q->timeouts);
            } else if (pass == 1) {
                await_count(&q->done[0][kStageFinish][p], pt->count[0][kStageFinish][p], &q->timeouts);
            }
            // one acquire per workgroup: its waves share this CU's caches and wait at the barrier
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        const BatchTable &t = tabs[pass];
        if (stage == kStageZero) {
            unsigned long long *z = pt->acc[pass][p];
            const long long b0 = (long long)local * kZeroChunk;
            const long long e0 = min(b0 + kZeroChunk, pt->acc_words[p]);
            for (long long i = b0 + threadIdx.x; i < e0; i += 256) z[i] = 0ull;
        } else if (stage == kStageMatch) {
            if constexpr (SC) match_lds4_body<M>(t.m[p], local, smem);
            else match_lds_body<M>(t.m[p], local, smem);
        } else if (stage == kStageFilter) {
            const int tx = t.filter_tiles_x[p];
            if (pass == 0) filter_tile_body<false, T, T, true>(t.f[p], local % tx, local / tx, smem);
            else filter_tile_body<true, T, T, true>(t.f[p], local % tx, local / tx, smem);
        } else {
            long long (*sD)[72] = reinterpret_cast<long long (*)[72]>(smem);
            long long (*sH)[64] = reinterpret_cast<long long (*)[64]>(smem + 23 * 72 * 2);
            const int tx = t.finish_tiles_x[p], ty = t.finish_tiles_y[p];
            const int nfin = t.finish_off[p + 1] - t.finish_off[p];
            for (int k = 0; k < kFinishPer; ++k) {
                const int ft = local * kFinishPer + k;
                if (ft >= nfin) break;  // uniform: every wave leaves together
                if (k > 0) __syncthreads();  // the previous tile's reads of sD / sH are done
                finish_body<true>(t.fin[p], ft % tx, (ft / tx) % ty, ft / (tx * ty), sD, sH);
            }
        }
        // every wave's writes have left it before the barrier; lane 0 then releases them for the
        // workgroup (cache write-back and the counter update) when it publishes the item
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        prev_pass = pass;
        prev_stage = stage;
        prev_plane = p;
    }
    // (the last item was published at the top of the iteration that left the loop) the last
    // workgroup to leave resets the queue for the next launch
    if (threadIdx.x == 0) s_last = atomicAdd(&q->exited, 1u) == gridDim.x - 1;
    __syncthreads();
    if (s_last) {
        unsigned *d = &q->done[0][0][0];
        for (int i = threadIdx.x; i < 2 * 4 * kMaxBatch; i += 256) d[i] = 0u;
        if (threadIdx.x == 0) {
            q->next = 0u;
            q->exited = 0u;
        }
    }
}

}  // namespace vszk

// ================================================================================================
// Kernel-layer C ABI
// ================================================================================================
using namespace vszk;

namespace {

bool kparams_ok(const vsz_bm3d_kparams *p) {
    if (!p) return false;
    if (p->width < 8 || p->height < 8 || p->stride < p->width) return false;
    if (p->block_step < 1 || p->block_step > 8 || p->bm_range < 1) return false;
    if (p->radius < 0 || p->radius > 16 || p->ps_num < 1 || p->ps_num > 8 || p->ps_range < 1) return false;
    if (p->metric < 0 || p->metric > 4 || p->t2d < 0 || p->t2d > 3 || p->t1d < 0 || p->t1d > 3) return false;
    // the split-BF16 MFMA form exists for the DCT pair only
    if (p->transform_precision < 0 || p->transform_precision > 1) return false;
    if (p->transform_precision == 1 && (p->t2d != 0 || p->t1d != 0)) return false;
    // the matcher addresses a [TW][H][S] plane stack with 32-bit byte offsets and packs x, y in 16 bits
    if ((long long)(2 * p->radius + 1) * p->height * p->stride * 4 >= (1LL << 31)) return false;
    if (p->width > 65535 || p->height > 65535) return false;
    return true;
}

struct Layout {
    int nbx, nby, P, TW;
    long long ts, nblocks;
    size_t match_bytes, acc_bytes;
    bool exact;
};

Layout layout_of(const vsz_bm3d_kparams *p) {
    Layout L{};
    L.nbx = (p->width - 8 + p->block_step - 1) / p->block_step + 1;
    L.nby = (p->height - 8 + p->block_step - 1) / p->block_step + 1;
    L.P = p->chroma ? 3 : 1;
    L.TW = 2 * p->radius + 1;
    L.ts = (long long)p->height * p->stride;
    L.nblocks = (long long)L.nbx * L.nby;
    L.exact = vsz_exact_extractor(p->extractor_exp);
    L.match_bytes = ((size_t)L.nblocks * 8 * sizeof(int4) + 255) & ~size_t(255);
    L.acc_bytes = (size_t)L.P * L.TW * 2 * L.ts * (L.exact ? 8 : 4);
    return L;
}

// LDS bytes the staged matcher needs for a full 16 x 8*WR block tile (0: does not fit).
constexpr size_t kMatchLdsBudget = 80 * 1024;  // two 256-lane workgroups per CU

size_t match_lds_bytes(int block_step, int bm_range, int *lds_stride, int wr = 2, int split = 1) {
    const long long rw = 15LL * block_step + 2LL * bm_range + 9;
    const long long rh = (8LL * wr / split - 1) * block_step + 2LL * bm_range + 8;
    const long long ls = (rw + 3 + 31) / 32 * 32;
    const size_t bytes = (size_t)(rh * ls * 4);
    if (lds_stride) *lds_stride = (int)ls;
    return bytes <= kMatchLdsBudget ? bytes : 0;
}

// LDS bytes of the shared-column matcher's 15 x 16 block tile (0: not usable).
size_t match_lds4_bytes(const MatchArgs &a, int *lds_stride) {
    if (a.block_step != 4 || (a.width - 8) % 4 != 0) return 0;
    const long long rw = 14LL * 4 + 2LL * a.bm_range + 9;
    const long long rh = 7LL * 4 + 2LL * a.bm_range + 8;
    const long long ls = (rw + 3 + 31) / 32 * 32;
    const size_t bytes = (size_t)(rh * ls * 4);
    if (lds_stride) *lds_stride = (int)ls;
    return bytes <= kMatchLdsBudget ? bytes : 0;
}

// Which kernel runs a spatial (or centre-frame) search.
enum MatchKernel { kMkLds4 = 0, kMkLds1, kMkLds2, kMkLds4Bands, kMkLdsShort, kMkGlobal };

struct MatchPlan {
    MatchKernel kernel;
    size_t lds;
    int lds_stride;
};

// The shared-column matcher for block_step 4 with ssd/sad; otherwise the LDS-staged matcher with
// each block's search rows split over 2 or 4 lanes when the plane has too few blocks to fill the
// chip (about 2048 waves), then its 16 x 8-block tile; the global-memory matcher when no staged
// region fits the LDS budget.
MatchPlan plan_match(const MatchArgs &a, int metric) {
    MatchPlan m{kMkGlobal, 0, 0};
    if (metric == 0 || metric == 1) {
        m.lds = match_lds4_bytes(a, &m.lds_stride);
        if (m.lds) {
            m.kernel = kMkLds4;
            return m;
        }
    }
    const long long nb = (long long)a.nbx * a.nby;
    const int split = nb >= 2048LL * 64 ? 1 : nb >= 1024LL * 64 ? 2 : 4;
    m.lds = match_lds_bytes(a.block_step, a.bm_range, &m.lds_stride, 2, split);
    if (m.lds) {
        m.kernel = split == 1 ? kMkLds1 : split == 2 ? kMkLds2 : kMkLds4Bands;
        return m;
    }
    m.lds = match_lds_bytes(a.block_step, a.bm_range, &m.lds_stride, 1, 1);
    if (m.lds) {
        m.kernel = kMkLdsShort;
        return m;
    }
    m.lds = 0;
    m.lds_stride = 0;
    return m;
}

template <int M> void launch_spatial(const MatchArgs &a, const MatchPlan &pl, hipStream_t s) {
    MatchArgs b = a;
    b.lds_stride = pl.lds_stride;
    switch (pl.kernel) {
    case kMkLds4:
        if constexpr (M == 0 || M == 1) {
            b.tiles_x = (a.nbx + 14) / 15;
            hipLaunchKernelGGL((k_match_lds4<M>), dim3((unsigned)(b.tiles_x * ((a.nby + 7) / 8))), dim3(256), pl.lds, s, b);
        }
        break;
    case kMkLds1:
        b.tiles_x = (a.nbx + 15) / 16;
        hipLaunchKernelGGL((k_match_lds<M, 2, 1>), dim3((unsigned)(b.tiles_x * ((a.nby + 15) / 16))), dim3(256), pl.lds, s, b);
        break;
    case kMkLds2:
        b.tiles_x = (a.nbx + 15) / 16;
        hipLaunchKernelGGL((k_match_lds<M, 2, 2>), dim3((unsigned)(b.tiles_x * ((a.nby + 7) / 8))), dim3(256), pl.lds, s, b);
        break;
    case kMkLds4Bands:
        b.tiles_x = (a.nbx + 15) / 16;
        hipLaunchKernelGGL((k_match_lds<M, 2, 4>), dim3((unsigned)(b.tiles_x * ((a.nby + 3) / 4))), dim3(256), pl.lds, s, b);
        break;
    case kMkLdsShort:
        b.tiles_x = (a.nbx + 15) / 16;
        hipLaunchKernelGGL((k_match_lds<M, 1, 1>), dim3((unsigned)(b.tiles_x * ((a.nby + 7) / 8))), dim3(128), pl.lds, s, b);
        break;
    default:
        b.tiles_x = (a.nbx + 7) / 8;
        hipLaunchKernelGGL((k_match<M>), dim3((unsigned)(b.tiles_x * ((a.nby + 7) / 8))), dim3(64), 0, s, b);
        break;
    }
}

// radius 0: the spatial search with its self-inclusion. radius > 0: the centre frame's spatial
// search with the self-inclusion held back, then the temporal stage (predictive search, merges,
// self-inclusion).
template <int M> void launch_match_m(const MatchArgs &a, hipStream_t s) {
    if (a.radius == 0) {
        launch_spatial<M>(a, plan_match(a, M), s);
        return;
    }
    MatchArgs c = a;
    c.img = a.img + (long long)a.radius * a.ts;
    c.no_self = 1;
    launch_spatial<M>(c, plan_match(c, M), s);
    MatchArgs t = a;
    t.zero = nullptr;  // the centre frame's spatial search cleared it
    t.tiles_x = (a.nbx + 7) / 8;
    if (a.ps_num == 2) {  // one lane per window
        const dim3 grid((unsigned)(t.tiles_x * ((a.nby + 1) / 2)), 1, 1);
        const size_t lds = (size_t)16 * 2 * a.radius * a.ps_num * sizeof(int2);
        hipLaunchKernelGGL((k_match_temporal<M, 2>), grid, dim3(64), lds, s, t);
    } else {
        const dim3 grid((unsigned)(t.tiles_x * ((a.nby + 3) / 4)), 1, 1);
        const size_t lds = (size_t)32 * 2 * a.radius * a.ps_num * sizeof(int2);
        hipLaunchKernelGGL((k_match_temporal<M, 1>), grid, dim3(64), lds, s, t);
    }
}

void launch_match(const MatchArgs &a, int metric, hipStream_t s) {
    switch (metric) {
    case 0: launch_match_m<0>(a, s); break;
    case 1: launch_match_m<1>(a, s); break;
    case 2: launch_match_m<2>(a, s); break;
    case 3: launch_match_m<3>(a, s); break;
    default: launch_match_m<4>(a, s); break;
    }
}

// LDS budget of the tiled filter kernel: per-wave transpose buffers + the region image.
constexpr int kTileThreads = 256;
constexpr size_t kTransposeBytes = 4 * 8 * 72 * sizeof(float);
constexpr size_t kRegionBudgetBytes = 64 * 1024;
constexpr size_t kPairTransposeBytes = 4 * kPairBufFloats * sizeof(float);  // split-BF16 form

size_t transpose_bytes(const vsz_bm3d_kparams *p) {
    return p->transform_precision == 1 ? kPairTransposeBytes : kTransposeBytes;
}

// LDS per workgroup: two FP32 workgroups per CU (the single-frame hard pass may take three, see
// plan_tiles_frame); the split-BF16 kernel (fewer registers, longer MFMA dependency chains) runs
// five: a measured sweep of 28, 32, 40 and 48 KB on the 1080p bench frame put 32 KB first (1.56 ms
// per two-pass step, against 1.62, 1.59 and 1.74).
size_t region_budget(const vsz_bm3d_kparams *p) {
    // split-BF16 tiles: as many workgroups per CU as their registers allow (five for the hard
    // pass, four for Wiener) share the CU's 160 KB
    if (p->transform_precision == 1) return size_t(p->final_pass ? 40 : 32) * 1024;
    return kRegionBudgetBytes;
}

struct TilePlan {
    bool ok = false;
    int tbx = 0, tby = 0, guard_bits = 0;
    size_t lds_bytes = 0;
};

// Exact int32 LDS cells take at most 8*tbx*tby contributions per plane; each must stay below
// 2^guard_bits so that the cell cannot overflow: (8*tbx*tby) * 2^guard_bits <= 2^31.
int guard_bits_for(int tbx, int tby) {
    int lg = 0;
    while ((1LL << lg) < 8LL * tbx * tby) ++lg;
    return 31 - lg;
}

// Picks the block tile of the LDS-accumulating filter kernel: the region a tile's groups can
// reach must fit the LDS budget; among those, the tiling with the least flush traffic (tiles x
// region cells) that still yields enough workgroups to fill the chip once (two workgroups per CU
// on 256 CUs; more, smaller tiles measured slower on 1080p planes at block_step 4 and 8).
TilePlan plan_tiles(const vsz_bm3d_kparams *p, const Layout &L, int min_wgs = 512) {
    TilePlan best;
    if (p->radius != 0) return best;
    const int opts[] = {32, 16, 8, 4, 2, 1};
    long long best_cost = -1;
    for (int tbx : opts)
        for (int tby : opts) {
            if (tbx > L.nbx * 2 && tbx > 1) continue;
            if (tby > L.nby * 2 && tby > 1) continue;
            if (tbx * tby > 512) continue;
            const long long rw = std::min<long long>((long long)(tbx - 1) * p->block_step + 8 + 2LL * p->bm_range, p->width);
            const long long rh = std::min<long long>((long long)(tby - 1) * p->block_step + 8 + 2LL * p->bm_range, p->height);
            const size_t bytes = transpose_bytes(p) + (size_t)(rw * rh) * 4;
            if (bytes > region_budget(p)) continue;
            const long long wgs = (long long)((L.nbx + tbx - 1) / tbx) * ((L.nby + tby - 1) / tby);
            long long cost = wgs * rw * rh;
            if (wgs < min_wgs) cost += (min_wgs - wgs) * rw * rh * 4;  // under-filled chip
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best.ok = true;
                best.tbx = tbx;
                best.tby = tby;
                best.lds_bytes = bytes;
            }
        }
    if (best.ok) best.guard_bits = guard_bits_for(best.tbx, best.tby);
    return best;
}

// Tile plan of the single-frame FP32 hard pass, whose kernel can run three workgroups per CU
// (768 on 256 CUs) when a tile's region fits a third of the CU's 160 KB. Used when the plane
// splits into one near-full round of such workgroups (691-768) with at least 128 groups per tile;
// otherwise plan_tiles decides (small planes, wide windows, large planes). Among those shapes:
// the fewest 32-group iterations per workgroup (a wave filters 8 groups at a time), then the
// most workgroups (the fullest round), then the fewest flushed cells. On the 1080p bench frame
// (block_step 4, bm_range 16) this picks 10x17 tiles: basic pass 0.659 ms against 0.678 with
// two 16x16 tiles per CU (three interleaved runs each; 16x11 0.656, 12x15 0.665, 12x16 0.666).
TilePlan plan_tiles_frame(const vsz_bm3d_kparams *p, const Layout &L) {
    if (p->transform_precision != 0 || p->final_pass || p->radius != 0) return plan_tiles(p, L);
    constexpr long long kSlots = 3 * 256, kMinWgs = 691;
    constexpr size_t kBudget = 53 * 1024;
    TilePlan best;
    int best_iters = 0;
    long long best_wgs = 0, best_cells = 0;
    for (int tbx = 1; tbx <= 32; ++tbx)
        for (int tby = 1; tby <= 32; ++tby) {
            const int groups = tbx * tby;
            if (groups < 128 || groups > 512 || tbx > L.nbx || tby > L.nby) continue;
            const long long rw = std::min<long long>((long long)(tbx - 1) * p->block_step + 8 + 2LL * p->bm_range, p->width);
            const long long rh = std::min<long long>((long long)(tby - 1) * p->block_step + 8 + 2LL * p->bm_range, p->height);
            const size_t bytes = transpose_bytes(p) + (size_t)(rw * rh) * 4;
            if (bytes > kBudget) continue;
            const long long wgs = (long long)((L.nbx + tbx - 1) / tbx) * ((L.nby + tby - 1) / tby);
            if (wgs < kMinWgs || wgs > kSlots) continue;
            const int iters = (groups + 31) / 32;
            const long long cells = wgs * rw * rh;
            const bool better = !best.ok || iters < best_iters ||
                                (iters == best_iters && (wgs > best_wgs || (wgs == best_wgs && cells < best_cells)));
            if (better) {
                best.ok = true;
                best.tbx = tbx;
                best.tby = tby;
                best.lds_bytes = bytes;
                best_iters = iters;
                best_wgs = wgs;
                best_cells = cells;
            }
        }
    if (!best.ok) return plan_tiles(p, L);
    best.guard_bits = guard_bits_for(best.tbx, best.tby);
    return best;
}

// Every candidate tile shape that fits the LDS budget (the autotuner's search space).
std::vector<TilePlan> tile_candidates(const vsz_bm3d_kparams *p, const Layout &L) {
    std::vector<TilePlan> out;
    if (p->radius != 0) return out;
    const int opts[] = {32, 16, 8, 4, 2, 1};
    for (int tbx : opts)
        for (int tby : opts) {
            if ((tbx > L.nbx * 2 && tbx > 1) || (tby > L.nby * 2 && tby > 1) || tbx * tby > 512 || tbx * tby < 16)
                continue;
            const long long rw = std::min<long long>((long long)(tbx - 1) * p->block_step + 8 + 2LL * p->bm_range, p->width);
            const long long rh = std::min<long long>((long long)(tby - 1) * p->block_step + 8 + 2LL * p->bm_range, p->height);
            const size_t bytes = transpose_bytes(p) + (size_t)(rw * rh) * 4;
            if (bytes > region_budget(p)) continue;
            TilePlan t;
            t.ok = true;
            t.tbx = tbx;
            t.tby = tby;
            t.lds_bytes = bytes;
            t.guard_bits = guard_bits_for(tbx, tby);
            out.push_back(t);
        }
    return out;
}

template <bool FINAL, int T2, int T1, bool EXACT>
void launch_filter_4(const FilterArgs &a, const TilePlan &tp, hipStream_t s) {
    if (tp.ok) {
        const dim3 grid((unsigned)(((a.nbx + tp.tbx - 1) / tp.tbx) * ((a.nby + tp.tby - 1) / tp.tby)), 1, 1);
        hipLaunchKernelGGL((k_filter_tile<FINAL, T2, T1, EXACT>), grid, dim3(kTileThreads), tp.lds_bytes, s, a);
    } else if constexpr (T2 == kDctMfma) {
        const unsigned grid = (unsigned)((a.nblocks + 1) / 2);  // two groups per wave
        hipLaunchKernelGGL((k_filter_mfma<FINAL, EXACT>), dim3(grid), dim3(64), 0, s, a);
    } else {
        const unsigned grid = (unsigned)((a.nblocks + 7) / 8);
        hipLaunchKernelGGL((k_filter<FINAL, T2, T1, EXACT>), dim3(grid), dim3(64), 0, s, a);
    }
}

template <bool FINAL, int T2, int T1> void launch_filter_3(const FilterArgs &a, const TilePlan &tp, hipStream_t s) {
    if (a.exact) launch_filter_4<FINAL, T2, T1, true>(a, tp, s);
    else launch_filter_4<FINAL, T2, T1, false>(a, tp, s);
}

template <bool FINAL, int T2> void launch_filter_2(const FilterArgs &a, const TilePlan &tp, int t1, hipStream_t s) {
    switch (t1) {
    case 0: launch_filter_3<FINAL, T2, 0>(a, tp, s); break;
    case 1: launch_filter_3<FINAL, T2, 1>(a, tp, s); break;
    case 2: launch_filter_3<FINAL, T2, 2>(a, tp, s); break;
    default: launch_filter_3<FINAL, T2, 3>(a, tp, s); break;
    }
}

template <bool FINAL> void launch_filter_1(const FilterArgs &a, const TilePlan &tp, int t2, int t1, hipStream_t s) {
    switch (t2) {
    case 0: launch_filter_2<FINAL, 0>(a, tp, t1, s); break;
    case 1: launch_filter_2<FINAL, 1>(a, tp, t1, s); break;
    case 2: launch_filter_2<FINAL, 2>(a, tp, t1, s); break;
    default: launch_filter_2<FINAL, 3>(a, tp, t1, s); break;
    }
}

int check_launch(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return vsz_fail(VSZ_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    return VSZ_OK;
}

// The search of vsz_bm3d_match; with zero non-null its spatial kernel also clears
// zero[0, zero_n64) (the accumulator of the pass that follows).
int match_and_clear(const vsz_bm3d_kparams *p, const float *d_img, int32_t *d_matches, unsigned long long *zero,
                    long long zero_n64, hipStream_t s) {
    const Layout L = layout_of(p);
    MatchArgs m{};
    m.img = d_img;
    m.matches = reinterpret_cast<int4 *>(d_matches);
    m.width = p->width;
    m.height = p->height;
    m.stride = p->stride;
    m.nbx = L.nbx;
    m.nby = L.nby;
    m.block_step = p->block_step;
    m.bm_range = p->bm_range;
    m.radius = p->radius;
    m.ps_num = p->ps_num;
    m.ps_range = p->ps_range;
    m.ts = L.ts;
    m.tiles_x = (L.nbx + 7) / 8;
    m.zero = zero;
    m.zero_n64 = zero_n64;
    launch_match(m, p->metric, s);
    return check_launch("vsz_bm3d_match");
}

}  // namespace

extern "C" {

int vsz_bm3d_num_blocks(const vsz_bm3d_kparams *p, int *nbx, int *nby) {
    if (!kparams_ok(p)) return vsz_fail(VSZ_EINVAL, "invalid kernel parameters");
    const Layout L = layout_of(p);
    if (nbx) *nbx = L.nbx;
    if (nby) *nby = L.nby;
    return VSZ_OK;
}

size_t vsz_bm3d_workspace_size(const vsz_bm3d_kparams *p) {
    if (!kparams_ok(p)) return 0;
    const Layout L = layout_of(p);
    return L.match_bytes + L.acc_bytes;
}

int vsz_bm3d_match(const vsz_bm3d_kparams *p, const float *d_img, int32_t *d_matches, void *stream) {
    if (!kparams_ok(p) || !d_img || !d_matches) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_match: invalid arguments");
    return match_and_clear(p, d_img, d_matches, nullptr, 0, (hipStream_t)stream);
}

const char *vsz_bm3d_match_plan(const vsz_bm3d_kparams *p) {
    static const char *const names[] = {"lds4", "lds_rows1", "lds_rows2", "lds_rows4", "lds_short", "global"};
    static const char *const temporal_names[] = {"lds4+temporal", "lds_rows1+temporal", "lds_rows2+temporal",
                                                 "lds_rows4+temporal", "lds_short+temporal", "global+temporal"};
    if (!kparams_ok(p)) {
        vsz_fail(VSZ_EINVAL, "invalid kernel parameters");
        return "";
    }
    const Layout L = layout_of(p);
    MatchArgs m{};
    m.width = p->width;
    m.height = p->height;
    m.stride = p->stride;
    m.nbx = L.nbx;
    m.nby = L.nby;
    m.block_step = p->block_step;
    m.bm_range = p->bm_range;
    const MatchPlan pl = plan_match(m, p->metric);
    return p->radius > 0 ? temporal_names[pl.kernel] : names[pl.kernel];
}

int vsz_bm3d_filter_plan(const vsz_bm3d_kparams *p, int *tbx, int *tby, int *per_cu) {
    if (!kparams_ok(p) || !tbx || !tby || !per_cu) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_filter_plan: invalid arguments");
    const Layout L = layout_of(p);
    const TilePlan tp = plan_tiles_frame(p, L);
    *tbx = tp.ok ? tp.tbx : 0;
    *tby = tp.ok ? tp.tby : 0;
    int n = 0;
    if (tp.ok) {
        const int by_regs = p->transform_precision == 1 ? (p->final_pass ? 4 : 5) : (p->final_pass ? 2 : 3);
        n = std::min<int>(by_regs, (int)((160 * 1024) / tp.lds_bytes));
    }
    *per_cu = n;
    return VSZ_OK;
}

}  // extern "C"

namespace {

// Scratch of vsz_bm3d_run_noalloc: the match lists and the accumulator stack of one pass. Each GPU
// holds its own copy from the library's first use on it until the process ends. Host code must
// not pass `scratch` itself to a kernel (on the host it names a stand-in, not device memory); it
// uses the device address scratch_ptr() looked up once per device.
constexpr size_t kScratchBytes = size_t(128) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// Device address of the current device's scratch, or nullptr (error set) on failure.
void *scratch_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        vsz_fail(VSZ_EDEVICE, "vsz_bm3d_run_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            vsz_fail(VSZ_EDEVICE, "vsz_bm3d_run_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

int run_pass(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, void *d_work, float *d_out,
             int out_kind, void *stream);

}  // namespace

extern "C" {

int vsz_bm3d_run(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, void *d_work, float *d_out,
                 int out_kind, void *stream) {
    if (!kparams_ok(p) || !d_clip0 || !d_work || !d_out || (p->final_pass && !d_clip1) || out_kind < 0 ||
        out_kind > 1 || (out_kind == 0 && p->radius != 0))
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_run: invalid arguments");
    return run_pass(p, d_clip0, d_clip1, d_work, d_out, out_kind, stream);
}

int vsz_bm3d_run_noalloc(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, float *d_out,
                         int out_kind, void *stream) {
    if (!kparams_ok(p) || !d_clip0 || !d_out || (p->final_pass && !d_clip1) || out_kind < 0 || out_kind > 1 ||
        (out_kind == 0 && p->radius != 0))
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_run_noalloc: invalid arguments");
    const size_t need = vsz_bm3d_workspace_size(p);
    if (need > kScratchBytes)
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_run_noalloc: the pass needs %zu bytes of workspace, above the %zu-byte "
                        "static scratch; use vsz_bm3d_run with a workspace", need, kScratchBytes);
    void *work = scratch_ptr();
    if (!work) return VSZ_EDEVICE;
    return run_pass(p, d_clip0, d_clip1, work, d_out, out_kind, stream);
}

}  // extern "C"

namespace {

int run_pass(const vsz_bm3d_kparams *p, const float *d_clip0, const float *d_clip1, void *d_work, float *d_out,
             int out_kind, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const Layout L = layout_of(p);
    int4 *matches = reinterpret_cast<int4 *>(d_work);
    void *acc = static_cast<char *>(d_work) + L.match_bytes;

    // the spatial search also clears the accumulators (kernel stores: graph-capture safe, no
    // memset node); the filter below runs after it on the same stream
    int rc = match_and_clear(p, d_clip0, reinterpret_cast<int32_t *>(matches), (unsigned long long *)acc,
                             (long long)(L.acc_bytes / 8), s);
    if (rc) return rc;

    FilterArgs f{};
    f.matches = matches;
    f.clip0 = d_clip0;
    f.clip1 = p->final_pass ? d_clip1 : d_clip0;
    f.acc = acc;
    f.nblocks = L.nblocks;
    f.num_planes = L.P;
    f.stride = p->stride;
    f.ts = L.ts;
    f.plane_elems = (long long)L.TW * L.ts;
    f.acc_plane_elems = (long long)L.TW * 2 * L.ts;
    unsigned active = 0;
    for (int i = 0; i < L.P; ++i) {
        f.sigma[i] = p->sigma[i];
        // chroma planes whose scaled sigma is below FLT_EPSILON are skipped (src/bm3d.cu:804-810)
        if (!(p->chroma && p->sigma[i] < FLT_EPSILON)) active |= 1u << i;
    }
    f.active_mask = active;
    f.exact = L.exact ? 1 : 0;
    f.extractor_exp = p->extractor_exp;
    f.extractor = p->extractor_exp != 0 ? ldexpf(1.0f, p->extractor_exp) : 0.0f;
    f.width = p->width;
    f.height = p->height;
    f.block_step = p->block_step;
    f.bm_range = p->bm_range;
    f.nbx = L.nbx;
    f.nby = L.nby;
    const TilePlan tp = plan_tiles_frame(p, L);
    f.tbx = tp.tbx;
    f.tby = tp.tby;
    f.guard_bits = tp.guard_bits;
    if (p->transform_precision == 1) {
        if (p->final_pass) launch_filter_3<true, kDctMfma, kDctMfma>(f, tp, s);
        else launch_filter_3<false, kDctMfma, kDctMfma>(f, tp, s);
    } else if (p->final_pass) {
        launch_filter_1<true>(f, tp, p->t2d, p->t1d, s);
    } else {
        launch_filter_1<false>(f, tp, p->t2d, p->t1d, s);
    }
    rc = check_launch("vsz_bm3d_run: filter");
    if (rc) return rc;

    FinishArgs fa{};
    fa.acc = acc;
    fa.out = d_out;
    fa.width = p->width;
    fa.height = p->height;
    fa.stride = p->stride;
    fa.tw = L.TW;
    fa.out_kind = out_kind;
    fa.ts = L.ts;
    fa.proc_mask = p->proc_mask;
    fa.extractor_exp = p->extractor_exp;
    const dim3 grid((p->width + 63) / 64, (p->height + 15) / 16, L.P * L.TW);
    if (L.exact) hipLaunchKernelGGL(k_finish<true>, grid, dim3(256), 0, s, fa);
    else hipLaunchKernelGGL(k_finish<false>, grid, dim3(256), 0, s, fa);
    return check_launch("vsz_bm3d_run: finish");
}

}  // namespace

extern "C" {

int vsz_vaggregate_run(int width, int height, int stride, int radius, int n, int num_frames,
                       const float *const *d_slots, float *d_dst, void *stream) {
    if (width < 1 || height < 1 || stride < width || radius < 1 || radius > 16 || n < 0 || n >= num_frames ||
        !d_slots || !d_dst)
        return vsz_fail(VSZ_EINVAL, "vsz_vaggregate_run: invalid arguments");
    VaggArgs a{};
    a.width = width;
    a.height = height;
    a.stride = stride;
    a.ts = (long long)height * stride;
    a.tw = 2 * radius + 1;
    for (int i = 0; i < a.tw; ++i) {
        if (!d_slots[i]) return vsz_fail(VSZ_EINVAL, "vsz_vaggregate_run: null slot");
        a.slots[i] = d_slots[i];
        a.z[i] = vsz_agg_z(i, n, num_frames, radius);
    }
    a.dst = d_dst;
    const dim3 grid((width + 63) / 64, (height + 3) / 4, 1);
    hipLaunchKernelGGL(k_vaggregate, grid, dim3(256), 0, (hipStream_t)stream, a);
    return check_launch("vsz_vaggregate_run");
}

}  // extern "C"

// ================================================================================================
// Batched planes: plan, autotuner, graph
// ================================================================================================
namespace {

struct BatchPlane {
    vsz_bm3d_kparams k[2]{};  // basic, final
    Layout L{};
    TilePlan tile[2];
    // per pass: accumulators and match lists of each pass have addresses of their own, so that no
    // memory is rewritten within a persistent launch after another workgroup may have cached it
    size_t acc_off[2] = {0, 0}, match_off[2] = {0, 0}, basic_off = 0;
};

float bits_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// One tiled filter launch of the uniform transform `kind` (both stages), for the autotuner.
void launch_filter_tile_kind(bool final_pass, int kind, const FilterArgs &a, const TilePlan &tp, hipStream_t s) {
    switch (kind) {
    case 0: final_pass ? launch_filter_3<true, 0, 0>(a, tp, s) : launch_filter_3<false, 0, 0>(a, tp, s); break;
    case 1: final_pass ? launch_filter_3<true, 1, 1>(a, tp, s) : launch_filter_3<false, 1, 1>(a, tp, s); break;
    case 2: final_pass ? launch_filter_3<true, 2, 2>(a, tp, s) : launch_filter_3<false, 2, 2>(a, tp, s); break;
    default: final_pass ? launch_filter_3<true, 3, 3>(a, tp, s) : launch_filter_3<false, 3, 3>(a, tp, s); break;
    }
}

// Autotuner results, shared by every plan of the process: (shape, pass, kernel variant) -> tile.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int, int, int>, std::pair<int, int>> g_tune;

template <bool FINAL, int T, bool EXACT>
void launch_filter_batch(const BatchTable *d, int wgs, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_filter_tile_batch<FINAL, T, T, EXACT>), dim3(wgs), dim3(256), lds, s, d);
}

// kind: the batch's transform (dct 0, haar 1, wht 2, bior1.5 3, both stages) or kDctMfma.
template <bool FINAL, bool EXACT>
void launch_filter_batch_kind(int kind, const BatchTable *d, int wgs, size_t lds, hipStream_t s) {
    switch (kind) {
    case 0: launch_filter_batch<FINAL, 0, EXACT>(d, wgs, lds, s); break;
    case 1: launch_filter_batch<FINAL, 1, EXACT>(d, wgs, lds, s); break;
    case 2: launch_filter_batch<FINAL, 2, EXACT>(d, wgs, lds, s); break;
    case 3: launch_filter_batch<FINAL, 3, EXACT>(d, wgs, lds, s); break;
    default: launch_filter_batch<FINAL, kDctMfma, EXACT>(d, wgs, lds, s); break;
    }
}

void launch_match_batch(int metric, bool shared_cols, const BatchTable *d, int wgs, size_t lds, unsigned long long *zero,
                        long long n64, hipStream_t s) {
    if (shared_cols) {
        if (metric == 0) hipLaunchKernelGGL((k_match_lds4_batch<0>), dim3(wgs), dim3(256), lds, s, d, zero, n64);
        else hipLaunchKernelGGL((k_match_lds4_batch<1>), dim3(wgs), dim3(256), lds, s, d, zero, n64);
        return;
    }
    switch (metric) {
    case 0: hipLaunchKernelGGL((k_match_lds_batch<0>), dim3(wgs), dim3(256), lds, s, d, zero, n64); break;
    case 1: hipLaunchKernelGGL((k_match_lds_batch<1>), dim3(wgs), dim3(256), lds, s, d, zero, n64); break;
    case 2: hipLaunchKernelGGL((k_match_lds_batch<2>), dim3(wgs), dim3(256), lds, s, d, zero, n64); break;
    case 3: hipLaunchKernelGGL((k_match_lds_batch<3>), dim3(wgs), dim3(256), lds, s, d, zero, n64); break;
    default: hipLaunchKernelGGL((k_match_lds_batch<4>), dim3(wgs), dim3(256), lds, s, d, zero, n64); break;
    }
}

void launch_match_global_batch(int metric, const BatchTable *d, int wgs, unsigned long long *zero, long long n64,
                               hipStream_t s) {
    switch (metric) {
    case 0: hipLaunchKernelGGL((k_match_global_batch<0>), dim3(wgs), dim3(256), 0, s, d, zero, n64); break;
    case 1: hipLaunchKernelGGL((k_match_global_batch<1>), dim3(wgs), dim3(256), 0, s, d, zero, n64); break;
    case 2: hipLaunchKernelGGL((k_match_global_batch<2>), dim3(wgs), dim3(256), 0, s, d, zero, n64); break;
    case 3: hipLaunchKernelGGL((k_match_global_batch<3>), dim3(wgs), dim3(256), 0, s, d, zero, n64); break;
    default: hipLaunchKernelGGL((k_match_global_batch<4>), dim3(wgs), dim3(256), 0, s, d, zero, n64); break;
    }
}

template <bool FINAL, bool EXACT>
void launch_filter_direct_batch(int kind, const BatchTable *d, int wgs, hipStream_t s) {
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_filter_direct_batch<FINAL, 0, 0, EXACT>), dim3(wgs), dim3(256), 0, s, d); break;
    case 1: hipLaunchKernelGGL((k_filter_direct_batch<FINAL, 1, 1, EXACT>), dim3(wgs), dim3(256), 0, s, d); break;
    case 2: hipLaunchKernelGGL((k_filter_direct_batch<FINAL, 2, 2, EXACT>), dim3(wgs), dim3(256), 0, s, d); break;
    default: hipLaunchKernelGGL((k_filter_direct_batch<FINAL, 3, 3, EXACT>), dim3(wgs), dim3(256), 0, s, d); break;
    }
}

// Sizes the grid (as many workgroups as can be resident at once) when *grid is 0, then launches
// on s unless size_only (s may be the null stream).
template <int M, bool SC, int T>
int launch_persistent_t(const BatchTable *tabs, const PersistTable *pt, BatchQueue *q, size_t lds, int *grid,
                        bool size_only, hipStream_t s) {
    auto k = k_batch_persistent<M, SC, T>;
    if (*grid == 0) {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds) != hipSuccess || per_cu < 1 ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: no occupancy for the persistent kernel");
        *grid = per_cu * cus;
    }
    if (!size_only) hipLaunchKernelGGL(k, dim3(*grid), dim3(256), lds, s, tabs, pt, q);
    return VSZ_OK;
}

int launch_persistent(int metric, bool sc, bool mfma, const BatchTable *tabs, const PersistTable *pt, BatchQueue *q,
                      size_t lds, int *grid, bool size_only, hipStream_t s) {
    if (metric == 0) {
        if (sc) return mfma ? launch_persistent_t<0, true, kDctMfma>(tabs, pt, q, lds, grid, size_only, s)
                            : launch_persistent_t<0, true, 0>(tabs, pt, q, lds, grid, size_only, s);
        return mfma ? launch_persistent_t<0, false, kDctMfma>(tabs, pt, q, lds, grid, size_only, s)
                    : launch_persistent_t<0, false, 0>(tabs, pt, q, lds, grid, size_only, s);
    }
    if (sc) return mfma ? launch_persistent_t<1, true, kDctMfma>(tabs, pt, q, lds, grid, size_only, s)
                        : launch_persistent_t<1, true, 0>(tabs, pt, q, lds, grid, size_only, s);
    return mfma ? launch_persistent_t<1, false, kDctMfma>(tabs, pt, q, lds, grid, size_only, s)
                : launch_persistent_t<1, false, 0>(tabs, pt, q, lds, grid, size_only, s);
}

}  // namespace

struct vsz_bm3d_batch {
    vsz_bm3d_batch_params bp{};
    std::vector<BatchPlane> planes;
    bool exact = false;
    size_t acc_total = 0, acc_span = 0, work_bytes = 0;  // acc_span: the accumulators of one pass
    size_t match_lds = 0;
    int match_stride = 0;
    bool shared_cols = false;  // every plane takes the shared-column block_step-4 matcher
    bool sc_allowed = false;   // ... may take it (the autotuner then times both matchers)
    bool global_match = false;  // windows beyond the LDS matchers' budget: the global-memory matcher
    bool direct_filter = false; // a region beyond the filter tiles' LDS budget: direct-atomic filtering
    size_t lds_rows = 0, lds_sc = 0;
    int stride_rows = 0, stride_sc = 0;
    int schedule = VSZ_BATCH_STAGED;  // chosen schedule
    bool persist_ok = false;          // the persistent form supports these parameters
    PersistTable *h_pt = nullptr, *d_pt = nullptr;
    BatchQueue *d_q = nullptr;
    int persist_grid = 0;
    bool tuned = false;
    BatchTable *h_tab[2] = {nullptr, nullptr};
    BatchTable *d_tab = nullptr;  // [2]
    const void *last_in = nullptr, *last_out = nullptr, *last_work = nullptr;
    std::vector<const float *> last_noisy;
    std::vector<float *> last_dst;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t replayed = nullptr;  // recorded after every graph launch; the tables wait for it
    bool launched = false;

    ~vsz_bm3d_batch() {
        if (launched) (void)hipEventSynchronize(replayed);
        if (replayed) (void)hipEventDestroy(replayed);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (d_tab) (void)hipFree(d_tab);
        if (d_pt) (void)hipFree(d_pt);
        if (d_q) (void)hipFree(d_q);
        delete h_pt;
        delete h_tab[0];
        delete h_tab[1];
    }

    int passes() const { return bp.two_pass ? 2 : 1; }
    // workgroups of one plane's filtering: tiles, or 32 groups each on the direct path
    int filter_wgs_of(const BatchPlane &pl, int pass) const {
        if (direct_filter) return (int)((pl.L.nblocks + 31) / 32);
        return ((pl.L.nbx + pl.tile[pass].tbx - 1) / pl.tile[pass].tbx) * ((pl.L.nby + pl.tile[pass].tby - 1) / pl.tile[pass].tby);
    }
    int filter_wgs(int pass) const {
        int n = 0;
        for (const auto &pl : planes) n += filter_wgs_of(pl, pass);
        return n;
    }
    size_t filter_lds(int pass) const {
        size_t m = 0;
        if (!direct_filter)
            for (const auto &pl : planes) m = std::max(m, pl.tile[pass].lds_bytes);
        return m;
    }
    int match_tiles_x(const BatchPlane &pl) const {
        return global_match ? (pl.L.nbx + 7) / 8 : shared_cols ? (pl.L.nbx + 14) / 15 : (pl.L.nbx + 15) / 16;
    }
    int match_tiles_y(const BatchPlane &pl) const {
        return global_match ? (pl.L.nby + 7) / 8 : shared_cols ? (pl.L.nby + 7) / 8 : (pl.L.nby + 15) / 16;
    }
    // workgroups of one plane's matching (the global matcher: four 8 x 8-block tiles each)
    int match_wgs_of(const BatchPlane &pl) const {
        const int tiles = match_tiles_x(pl) * match_tiles_y(pl);
        return global_match ? (tiles + 3) / 4 : tiles;
    }
    int match_wgs() const {
        int n = 0;
        for (const auto &pl : planes) n += match_wgs_of(pl);
        return n;
    }
    int finish_wgs() const {
        int n = 0;
        for (const auto &pl : planes) n += ((pl.k[0].width + 63) / 64) * ((pl.k[0].height + 15) / 16);
        return n;
    }

    void fill_tables(const float *const *noisy, float *const *dst, char *work) {
        for (int pass = 0; pass < passes(); ++pass) {
            BatchTable &t = *h_tab[pass];
            t.n = (int)planes.size();
            int mo = 0, fo = 0, so = 0;
            for (size_t i = 0; i < planes.size(); ++i) {
                const BatchPlane &pl = planes[i];
                const vsz_bm3d_kparams &k = pl.k[pass];
                const bool last = pass == passes() - 1;
                const float *clip0 = pass == 0 ? noisy[i] : reinterpret_cast<const float *>(work + pl.basic_off);
                float *out = last ? dst[i] : reinterpret_cast<float *>(work + pl.basic_off);
                int4 *matches = reinterpret_cast<int4 *>(work + pl.match_off[pass]);
                void *acc = work + pl.acc_off[pass];
                t.match_off[i] = mo;
                t.filter_off[i] = fo;
                t.finish_off[i] = so;
                MatchArgs &m = t.m[i];
                m = MatchArgs{};
                m.img = clip0;
                m.matches = matches;
                m.width = k.width;
                m.height = k.height;
                m.stride = k.stride;
                m.nbx = pl.L.nbx;
                m.nby = pl.L.nby;
                m.block_step = k.block_step;
                m.bm_range = k.bm_range;
                m.ps_num = k.ps_num;
                m.ps_range = k.ps_range;
                m.ts = pl.L.ts;
                m.tiles_x = match_tiles_x(pl);
                m.lds_stride = match_stride;
                mo += match_wgs_of(pl);
                FilterArgs &f = t.f[i];
                f = FilterArgs{};
                f.matches = matches;
                f.clip0 = clip0;
                f.clip1 = noisy[i];
                f.acc = acc;
                f.nblocks = pl.L.nblocks;
                f.num_planes = 1;
                f.stride = k.stride;
                f.ts = pl.L.ts;
                f.plane_elems = pl.L.ts;
                f.acc_plane_elems = 2 * pl.L.ts;
                f.sigma[0] = f.sigma[1] = f.sigma[2] = k.sigma[0];
                f.active_mask = 1;
                f.exact = exact ? 1 : 0;
                f.extractor_exp = k.extractor_exp;
                f.extractor = k.extractor_exp != 0 ? ldexpf(1.0f, k.extractor_exp) : 0.0f;
                f.width = k.width;
                f.height = k.height;
                f.block_step = k.block_step;
                f.bm_range = k.bm_range;
                f.nbx = pl.L.nbx;
                f.nby = pl.L.nby;
                f.tbx = pl.tile[pass].tbx;
                f.tby = pl.tile[pass].tby;
                f.guard_bits = pl.tile[pass].guard_bits;
                t.filter_tiles_x[i] = direct_filter ? 1 : (pl.L.nbx + f.tbx - 1) / f.tbx;
                fo += filter_wgs_of(pl, pass);
                FinishArgs &fa = t.fin[i];
                fa = FinishArgs{};
                fa.acc = acc;
                fa.out = out;
                fa.width = k.width;
                fa.height = k.height;
                fa.stride = k.stride;
                fa.tw = 1;
                fa.out_kind = 0;
                fa.ts = pl.L.ts;
                fa.proc_mask = 1;
                fa.extractor_exp = k.extractor_exp;
                t.finish_tiles_x[i] = (k.width + 63) / 64;
                t.finish_tiles_y[i] = (k.height + 15) / 16;
                so += t.finish_tiles_x[i] * t.finish_tiles_y[i];
            }
            t.match_off[planes.size()] = mo;
            t.filter_off[planes.size()] = fo;
            t.finish_off[planes.size()] = so;
        }
    }

    // All launches of one run, on stream s (captured into the graph); each pass's matcher also
    // clears that pass's accumulator span.
    void enqueue(char *work, hipStream_t s) {
        const long long n64 = (long long)(acc_span / 8);
        for (int pass = 0; pass < passes(); ++pass) {
            const BatchTable *d = d_tab + pass;
            unsigned long long *acc = (unsigned long long *)(work + pass * acc_span);
            if (global_match) launch_match_global_batch(bp.metric, d, match_wgs(), acc, n64, s);
            else launch_match_batch(bp.metric, shared_cols, d, match_wgs(), match_lds, acc, n64, s);
            const int fw = filter_wgs(pass);
            const size_t fl = filter_lds(pass);
            const bool mf = bp.transform_precision == 1;
            const int kind = mf ? kDctMfma : bp.t2d;
            if (direct_filter) {  // FP32 only (create rejects split-BF16 without tiles)
                if (pass == 0) exact ? launch_filter_direct_batch<false, true>(kind, d, fw, s)
                                     : launch_filter_direct_batch<false, false>(kind, d, fw, s);
                else exact ? launch_filter_direct_batch<true, true>(kind, d, fw, s)
                           : launch_filter_direct_batch<true, false>(kind, d, fw, s);
            } else if (pass == 0) {
                if (exact) launch_filter_batch_kind<false, true>(kind, d, fw, fl, s);
                else launch_filter_batch_kind<false, false>(kind, d, fw, fl, s);
            } else {
                if (exact) launch_filter_batch_kind<true, true>(kind, d, fw, fl, s);
                else launch_filter_batch_kind<true, false>(kind, d, fw, fl, s);
            }
            if (exact) hipLaunchKernelGGL(k_finish_batch<true>, dim3(finish_wgs()), dim3(256), 0, s, d);
            else hipLaunchKernelGGL(k_finish_batch<false>, dim3(finish_wgs()), dim3(256), 0, s, d);
        }
    }

    void use_matcher(bool sc) {
        shared_cols = sc;
        match_lds = sc ? lds_sc : lds_rows;
        match_stride = sc ? stride_sc : stride_rows;
        persist_grid = 0;  // the persistent kernel's LDS, so its occupancy, follows the matcher
    }

    // the persistent kernel's dynamic LDS: 16 bytes of loop control and the largest stage's need
    size_t persist_lds() const {
        size_t m = std::max(match_lds, (size_t)(23 * 72 + 23 * 64) * 8);  // finishing: sD, sH (int64)
        for (int pass = 0; pass < passes(); ++pass) m = std::max(m, filter_lds(pass));
        return 16 + m;
    }

    // Items of the persistent launch: per pass, the zeroing, matching, filtering and finishing
    // segments of every plane, each stage's planes largest first (from the tables' tile counts).
    void fill_persist(char *work) {
        PersistTable &t = *h_pt;
        std::vector<int> order(planes.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return (long long)planes[x].k[0].width * planes[x].k[0].height >
                   (long long)planes[y].k[0].width * planes[y].k[0].height;
        });
        memset(&t.count[0][0][0], 0, sizeof(t.count));
        t.nseg = 0;
        int item = 0;
        for (int pass = 0; pass < passes(); ++pass) {
            const BatchTable &bt = *h_tab[pass];
            for (int stage = 0; stage < 4; ++stage)
                for (int i : order) {
                    int n = 0;
                    if (stage == kStageZero) n = (int)((planes[i].L.acc_bytes / 8 + kZeroChunk - 1) / kZeroChunk);
                    else if (stage == kStageMatch) n = bt.match_off[i + 1] - bt.match_off[i];
                    else if (stage == kStageFilter) n = bt.filter_off[i + 1] - bt.filter_off[i];
                    else n = (bt.finish_off[i + 1] - bt.finish_off[i] + kFinishPer - 1) / kFinishPer;
                    t.count[pass][stage][i] = n;
                    if (n == 0) continue;
                    t.seg[t.nseg++] = BatchSeg{item, pass, stage, i};
                    item += n;
                }
        }
        t.total = item;
        for (size_t i = 0; i < planes.size(); ++i) {
            for (int pass = 0; pass < 2; ++pass)
                t.acc[pass][i] = reinterpret_cast<unsigned long long *>(work + planes[i].acc_off[pass]);
            t.acc_words[i] = (long long)(planes[i].L.acc_bytes / 8);
        }
    }

    // size_only: only sizes the grid (before a capture)
    int enqueue_persistent(hipStream_t s, bool size_only = false) {
        return launch_persistent(bp.metric, shared_cols, bp.transform_precision == 1, d_tab, d_pt, d_q, persist_lds(),
                                 &persist_grid, size_only, s);
    }

    // Uploads the tables of the current matcher choice for these pointers.
    int upload(const float *const *noisy, float *const *dst, char *work, hipStream_t s) {
        fill_tables(noisy, dst, work);
        for (int pass = 0; pass < passes(); ++pass)
            if (hipMemcpyAsync(d_tab + pass, h_tab[pass], sizeof(BatchTable), hipMemcpyHostToDevice, s) != hipSuccess)
                return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: table upload failed");
        if (persist_ok) {
            fill_persist(work);
            if (hipMemcpyAsync(d_pt, h_pt, sizeof(PersistTable), hipMemcpyHostToDevice, s) != hipSuccess)
                return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: table upload failed");
        }
        // the pageable sources may be reused as soon as the copies are done
        if (hipStreamSynchronize(s) != hipSuccess) return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: table upload failed");
        return VSZ_OK;
    }

    // One whole run of the current schedule and matcher, uncaptured, on stream s.
    int run_direct(char *work, hipStream_t s) {
        if (schedule == VSZ_BATCH_PERSISTENT) return enqueue_persistent(s);
        enqueue(work, s);
        return VSZ_OK;
    }

    int tune(const float *const *noisy, char *work, hipStream_t s);
    int tune_schedule(const float *const *noisy, float *const *dst, char *work, hipStream_t s);
};

// Times every candidate tile shape of the filter kernel on each distinct plane shape (both
// passes, the plane's own noisy data as input, its own workspace) and keeps the fastest.
int vsz_bm3d_batch::tune(const float *const *noisy, char *work, hipStream_t s) {
    if (direct_filter) return VSZ_OK;  // no tile fits: nothing to tune
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipEventCreate failed");
    int rc = VSZ_OK;
    for (size_t i = 0; i < planes.size() && rc == VSZ_OK; ++i) {
        BatchPlane &pl = planes[i];
        for (int pass = 0; pass < passes() && rc == VSZ_OK; ++pass) {
            const vsz_bm3d_kparams &k = pl.k[pass];
            const auto key = std::make_tuple(k.width, k.height, k.block_step, k.bm_range, pass, bp.metric,
                                             bp.transform_precision * 4 + bp.t2d, exact ? 1 : 0);  // kernel variant
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune.find(key);
                if (it != g_tune.end()) {
                    for (const TilePlan &c : tile_candidates(&k, pl.L))
                        if (c.tbx == it->second.first && c.tby == it->second.second) pl.tile[pass] = c;
                    continue;
                }
            }
            // matches for this plane (the final pass is timed on the noisy plane as its ref)
            int4 *matches = reinterpret_cast<int4 *>(work + pl.match_off[pass]);
            rc = vsz_bm3d_match(&k, noisy[i], reinterpret_cast<int32_t *>(matches), s);
            if (rc) break;
            FilterArgs f{};
            f.matches = matches;
            f.clip0 = noisy[i];
            f.clip1 = noisy[i];
            f.acc = work + pl.acc_off[pass];
            f.nblocks = pl.L.nblocks;
            f.num_planes = 1;
            f.stride = k.stride;
            f.ts = pl.L.ts;
            f.plane_elems = pl.L.ts;
            f.acc_plane_elems = 2 * pl.L.ts;
            f.sigma[0] = f.sigma[1] = f.sigma[2] = k.sigma[0];
            f.active_mask = 1;
            f.exact = exact ? 1 : 0;
            f.extractor_exp = k.extractor_exp;
            f.extractor = k.extractor_exp != 0 ? ldexpf(1.0f, k.extractor_exp) : 0.0f;
            f.width = k.width;
            f.height = k.height;
            f.block_step = k.block_step;
            f.bm_range = k.bm_range;
            f.nbx = pl.L.nbx;
            f.nby = pl.L.nby;
            float best_ms = -1.0f;
            TilePlan best = pl.tile[pass];
            for (const TilePlan &c : tile_candidates(&k, pl.L)) {
                f.tbx = c.tbx;
                f.tby = c.tby;
                f.guard_bits = c.guard_bits;
                const bool mf = bp.transform_precision == 1;
                float ms = 0.0f;
                for (int rep = 0; rep < 3 && rc == VSZ_OK; ++rep) {
                    (void)hipEventRecord(e0, s);
                    if (mf) {
                        if (pass == 0) launch_filter_3<false, kDctMfma, kDctMfma>(f, c, s);
                        else launch_filter_3<true, kDctMfma, kDctMfma>(f, c, s);
                    } else {
                        launch_filter_tile_kind(pass == 1, bp.t2d, f, c, s);
                    }
                    (void)hipEventRecord(e1, s);
                    if (hipEventSynchronize(e1) != hipSuccess) rc = vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: tuning launch failed");
                    float t = 0.0f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    if (rep > 0) ms += t;  // the first launch warms caches
                }
                if (rc) break;
                if (best_ms < 0.0f || ms < best_ms) {
                    best_ms = ms;
                    best = c;
                }
            }
            pl.tile[pass] = best;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[key] = std::make_pair(best.tbx, best.tby);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Chooses the schedule (one persistent launch or one launch per stage) and, where every plane
// allows the shared-column matcher, the matcher: with autotune, by timing whole runs of every
// allowed pair on this data (one warm-up, then the best of two); otherwise by preference.
int vsz_bm3d_batch::tune_schedule(const float *const *noisy, float *const *dst, char *work, hipStream_t s) {
    struct Cand {
        int sched;
        bool sc;
    };
    std::vector<Cand> cands;
    if (global_match) cands.push_back({VSZ_BATCH_STAGED, false});  // the only form for such windows
    for (int sc = 1; sc >= 0 && !global_match; --sc) {
        if (sc ? !sc_allowed : lds_rows == 0) continue;
        if (bp.schedule != VSZ_BATCH_STAGED && persist_ok) cands.push_back({VSZ_BATCH_PERSISTENT, sc != 0});
        if (bp.schedule != VSZ_BATCH_PERSISTENT) cands.push_back({VSZ_BATCH_STAGED, sc != 0});
    }
    if (cands.empty()) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_run: no schedule fits these parameters");
    if (!bp.autotune || cands.size() == 1) {
        // preference: the shared-column matcher, the persistent launch only where asked for
        Cand pick = cands[0];
        for (const Cand &c : cands)
            if (c.sc == cands[0].sc && c.sched == (bp.schedule == VSZ_BATCH_PERSISTENT ? VSZ_BATCH_PERSISTENT : VSZ_BATCH_STAGED))
                pick = c;
        use_matcher(pick.sc);
        schedule = pick.sched;
        return VSZ_OK;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipEventCreate failed");
    int rc = VSZ_OK;
    float best_ms = -1.0f;
    Cand best = cands[0];
    for (const Cand &c : cands) {
        use_matcher(c.sc);
        schedule = c.sched;
        rc = upload(noisy, dst, work, s);
        float ms = -1.0f;
        for (int rep = 0; rep < 3 && rc == VSZ_OK; ++rep) {
            (void)hipEventRecord(e0, s);
            rc = run_direct(work, s);
            (void)hipEventRecord(e1, s);
            if (rc == VSZ_OK && hipEventSynchronize(e1) != hipSuccess)
                rc = vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: tuning run failed");
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (rep > 0 && (ms < 0.0f || t < ms)) ms = t;
        }
        if (rc) break;
        if (best_ms < 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    use_matcher(best.sc);
    schedule = best.sched;
    return rc;
}

extern "C" {

int vsz_bm3d_batch_create(const vsz_bm3d_batch_params *bp, const vsz_bm3d_plane *planes, int n, vsz_bm3d_batch **out) {
    if (!bp || !planes || !out || n < 1 || n > kMaxBatch)
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: need 1..%d planes", kMaxBatch);
    if (bp->t2d < 0 || bp->t2d > 3 || bp->t1d != bp->t2d || bp->transform_precision < 0 || bp->transform_precision > 1 ||
        (bp->transform_precision == 1 && bp->t2d != 0))
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: a batch uses one transform for both stages (t2d == t1d; "
                                    "dct for split-BF16)");
    if (bp->schedule < VSZ_BATCH_STAGED || bp->schedule > VSZ_BATCH_AUTO)
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: schedule must be staged, persistent or auto");
    auto b = std::make_unique<vsz_bm3d_batch>();
    b->bp = *bp;
    b->exact = vsz_exact_extractor(bp->extractor_exp);
    int ls = 0;
    b->lds_rows = match_lds_bytes(bp->block_step, bp->bm_range, &ls);
    b->stride_rows = ls;
    // the shared-column matcher where every plane allows it (block_step 4, (W-8) % 4 == 0, ssd/sad)
    b->sc_allowed = bp->block_step == 4 && (bp->metric == 0 || bp->metric == 1);
    for (int i = 0; i < n && b->sc_allowed; ++i)
        if ((planes[i].width - 8) % 4 != 0) b->sc_allowed = false;
    if (b->sc_allowed) {
        MatchArgs probe{};
        probe.block_step = 4;
        probe.width = 8;
        probe.bm_range = bp->bm_range;
        b->lds_sc = match_lds4_bytes(probe, &ls);
        b->stride_sc = ls;
        b->sc_allowed = b->lds_sc != 0;
    }
    b->global_match = !b->lds_rows && !b->sc_allowed;  // windows beyond the LDS matchers' budget
    b->use_matcher(b->sc_allowed);
    const float fac[2] = {bits_float(0x3f021bb6u), bits_float(0x3e40c0c1u)};  // src/bm3d.zig:1013-1024
    size_t acc = 0, rest = 0;
    for (int i = 0; i < n; ++i) {
        BatchPlane pl;
        if (!(planes[i].sigma >= FLT_EPSILON))
            return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: plane %d: sigma must be >= FLT_EPSILON", i);
        for (int pass = 0; pass < 2; ++pass) {
            vsz_bm3d_kparams &k = pl.k[pass];
            k.width = planes[i].width;
            k.height = planes[i].height;
            k.stride = planes[i].stride;
            k.block_step = bp->block_step;
            k.bm_range = bp->bm_range;
            k.radius = 0;
            k.ps_num = 2;
            k.ps_range = 4;
            k.final_pass = pass;
            k.sigma[0] = k.sigma[1] = k.sigma[2] = planes[i].sigma * fac[pass];
            k.extractor_exp = bp->extractor_exp;
            k.metric = bp->metric;
            k.t2d = k.t1d = bp->t2d;
            k.proc_mask = 1;
            k.transform_precision = bp->transform_precision;
            if (!kparams_ok(&k)) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: plane %d: invalid geometry", i);
            pl.L = layout_of(&k);
            pl.tile[pass] = plan_tiles(&k, pl.L, 0);
            if (!pl.tile[pass].ok) {  // a search region beyond the tiles' LDS budget: direct atomics
                if (bp->transform_precision == 1)
                    return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: plane %d: no filter tile fits (split-BF16 "
                                                "batches need LDS tiles)", i);
                b->direct_filter = true;
            }
        }
        pl.acc_off[0] = acc;
        acc += (pl.L.acc_bytes + 255) & ~size_t(255);
        b->planes.push_back(pl);
    }
    // the persistent form: exact aggregation, ssd or sad, dct, the LDS matchers and filter tiles
    b->persist_ok = b->exact && (bp->metric == 0 || bp->metric == 1) && bp->t2d == 0 && !b->global_match &&
                    !b->direct_filter;
    if (bp->schedule == VSZ_BATCH_PERSISTENT && !b->persist_ok)
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_create: the persistent schedule needs extractor_exp > 0, the ssd or sad "
                                    "metric, the dct transform and windows within the LDS budgets");
    // [pass 0 accumulators][pass 1 accumulators] then per plane [matches per pass][basic], every
    // region 256-byte aligned (no cache line holds data of two regions)
    b->acc_span = acc;
    b->acc_total = acc * b->passes();
    rest = b->acc_total;
    for (auto &pl : b->planes) {
        pl.acc_off[1] = pl.acc_off[0] + b->acc_span;
        for (int pass = 0; pass < b->passes(); ++pass) {
            pl.match_off[pass] = rest;
            rest += (pl.L.match_bytes + 255) & ~size_t(255);
        }
        pl.basic_off = rest;
        rest += ((size_t)pl.L.ts * 4 + 255) & ~size_t(255);
    }
    b->work_bytes = rest;
    b->h_tab[0] = new BatchTable();
    b->h_tab[1] = new BatchTable();
    if (hipMalloc(&b->d_tab, 2 * sizeof(BatchTable)) != hipSuccess) {
        b->d_tab = nullptr;
        return vsz_fail(VSZ_ENOMEM, "vsz_bm3d_batch_create: hipMalloc failed");
    }
    if (b->persist_ok) {
        b->h_pt = new PersistTable();
        if (hipMalloc(&b->d_pt, sizeof(PersistTable)) != hipSuccess) {
            b->d_pt = nullptr;
            return vsz_fail(VSZ_ENOMEM, "vsz_bm3d_batch_create: hipMalloc failed");
        }
        if (hipMalloc(&b->d_q, sizeof(BatchQueue)) != hipSuccess) {
            b->d_q = nullptr;
            return vsz_fail(VSZ_ENOMEM, "vsz_bm3d_batch_create: hipMalloc failed");
        }
        // the queue starts empty; every persistent launch leaves it empty again
        if (hipMemset(b->d_q, 0, sizeof(BatchQueue)) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_create: hipMemset failed");
    }
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
        b->cap = nullptr;
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_create: hipStreamCreate failed");
    }
    if (hipEventCreateWithFlags(&b->replayed, hipEventDisableTiming) != hipSuccess) {
        b->replayed = nullptr;
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_create: hipEventCreate failed");
    }
    b->tuned = false;  // the first run tunes (when asked) and picks the schedule
    *out = b.release();
    return VSZ_OK;
}

size_t vsz_bm3d_batch_workspace_size(const vsz_bm3d_batch *b) { return b ? b->work_bytes : 0; }

int vsz_bm3d_batch_run(vsz_bm3d_batch *b, const float *const *d_noisy, float *const *d_out, void *d_work, void *stream) {
    if (!b || !d_noisy || !d_out || !d_work) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_run: invalid arguments");
    const size_t n = b->planes.size();
    for (size_t i = 0; i < n; ++i)
        if (!d_noisy[i] || !d_out[i]) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_run: null plane %zu", i);
    hipStream_t s = (hipStream_t)stream;
    char *work = static_cast<char *>(d_work);
    if (b->launched && hipStreamWaitEvent(s, b->replayed, 0) != hipSuccess)  // runs never overlap
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipStreamWaitEvent failed");
    if (!b->tuned) {
        if (b->launched && hipEventSynchronize(b->replayed) != hipSuccess)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: waiting for the previous run failed");
        int rc = b->bp.autotune ? b->tune(d_noisy, work, s) : VSZ_OK;
        if (rc == VSZ_OK) rc = b->tune_schedule(d_noisy, d_out, work, s);
        if (rc) return rc;
        b->tuned = true;
        b->last_work = nullptr;  // the tables hold the last tried candidate's: upload again
    }
    const bool same = b->last_work == d_work && b->last_noisy.size() == n &&
                      std::equal(b->last_noisy.begin(), b->last_noisy.end(), d_noisy) &&
                      std::equal(b->last_dst.begin(), b->last_dst.end(), d_out);
    if (!same) {
        // an earlier replay, possibly on another stream, may still read the device tables
        if (b->launched && hipEventSynchronize(b->replayed) != hipSuccess)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: waiting for the previous run failed");
        const int rc = b->upload(d_noisy, d_out, work, s);
        if (rc) return rc;
        b->last_work = d_work;
        b->last_noisy.assign(d_noisy, d_noisy + n);
        b->last_dst.assign(d_out, d_out + n);
        if (b->exec) {  // the captured launches read the workspace base through the zero kernel
            (void)hipGraphExecDestroy(b->exec);
            (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
        }
    }
    if (!b->exec) {
        if (b->schedule == VSZ_BATCH_PERSISTENT) {
            const int rc = b->enqueue_persistent(nullptr, true);  // sizes the grid outside the capture
            if (rc) return rc;
        }
        if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipStreamBeginCapture failed");
        if (b->schedule == VSZ_BATCH_PERSISTENT) (void)b->enqueue_persistent(b->cap);
        else b->enqueue(work, b->cap);
        if (hipStreamEndCapture(b->cap, &b->graph) != hipSuccess || !b->graph)
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: graph capture failed");
        if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            b->exec = nullptr;
            return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipGraphInstantiate failed");
        }
    }
    if (hipGraphLaunch(b->exec, s) != hipSuccess) return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipGraphLaunch failed");
    if (hipEventRecord(b->replayed, s) != hipSuccess) return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_run: hipEventRecord failed");
    b->launched = true;
    return VSZ_OK;
}

int vsz_bm3d_batch_tile(const vsz_bm3d_batch *b, int plane, int pass, int *tbx, int *tby) {
    if (!b || plane < 0 || plane >= (int)b->planes.size() || pass < 0 || pass > 1 || !tbx || !tby)
        return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_tile: invalid arguments");
    *tbx = b->planes[plane].tile[pass].tbx;
    *tby = b->planes[plane].tile[pass].tby;
    return VSZ_OK;
}

int vsz_bm3d_batch_check(vsz_bm3d_batch *b, void *stream) {
    if (!b) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_check: invalid arguments");
    if (!b->d_q) return VSZ_OK;
    unsigned t = 0;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
        hipMemcpy(&t, &b->d_q->timeouts, sizeof(t), hipMemcpyDeviceToHost) != hipSuccess)
        return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_check: reading the queue failed");
    if (t) return vsz_fail(VSZ_EDEVICE, "vsz_bm3d_batch_check: %u dependency waits of the persistent launch timed out", t);
    return VSZ_OK;
}

int vsz_bm3d_batch_schedule(const vsz_bm3d_batch *b, int *schedule, int *shared_cols) {
    if (!b || !schedule || !shared_cols) return vsz_fail(VSZ_EINVAL, "vsz_bm3d_batch_schedule: invalid arguments");
    *schedule = b->schedule;
    *shared_cols = b->global_match ? 2 : b->shared_cols ? 1 : 0;
    return VSZ_OK;
}

void vsz_bm3d_batch_free(vsz_bm3d_batch *b) { delete b; }

}  // extern "C"
