// TEST INFRASTRUCTURE — the BM3D oracle.
//
// A scalar CPU restatement of the reference BM3D / BM3Dv2 / VAggregate device path, written
// from the reference's algorithm and used only as the checker: by tests/, by the smoke check
// and by the CPU-baseline leg of bench.py. The product never links or calls it.
//
// Reference files followed (paths relative to the reference repository):
//   src/bm3d.cu:32-280    8-point transforms (dct, haar, wht, bior1.5), forward and inverse
//   src/bm3d.cu:290-390   block distances ssd / sad / zssd / zsad / ssd_norm
//   src/bm3d.cu:416-546   hard thresholding and Wiener shrinkage, 3-D transform order
//   src/bm3d.cu:579-660   reference-block geometry and the spatial search
//   src/bm3d.cu:663-756   temporal predictive search
//   src/bm3d.cu:758-788   self-inclusion
//   src/bm3d.cu:790-895   group gather and aggregation into [plane][slot][num|den][H][S]
//   src/bm3d.cu:898-922   aggregate (num / den)
//   src/bm3d_vagg.cu:18-48 and src/bm3d.zig:46-48, 1284-1317   temporal aggregation
//
// Canonical FP op sequence (shared contract with the HIP kernels, both built without FP
// contraction): every FMA/FMS/FNMS of the reference codelets and every `acc += a*b`
// accumulation is one fused multiply-add; every other expression is evaluated as written,
// left to right, one rounding per operation; divisions and square roots are IEEE
// correctly rounded; denormals are kept. Reductions follow the reference's per-lane
// order (even/odd row partials, then e0+e1) and the 8-lane xor butterfly, which equals
// ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)) on every lane.
//
// Aggregation: with a non-zero extractor E = 2^e every contribution q(u) = (u+E)-E lies on
// the grid 2^(e-24), so the FP32 oracle sums them exactly in 64-bit integers and rounds
// once (order-free, hence deterministic). With E = 0 the FP32 oracle sums in reference-block
// raster order, then patch rank, then row, then column. The FP64 instance always sums in
// that raster order in double.
//
// Parity status of each part is listed in the test files that pin it.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

enum Metric { M_SSD = 0, M_SAD = 1, M_ZSSD = 2, M_ZSAD = 3, M_SSD_NORM = 4 };
enum Xform { X_DCT = 0, X_HAAR = 1, X_WHT = 2, X_BIOR15 = 3 };

// Fused forms of the reference codelet macros (src/bm3d.cu:3-5).
template <class T> inline T fma_(T a, T b, T c) { return std::fma(a, b, c); }
template <class T> inline T FMA_(T a, T b, T c) { return std::fma(a, b, c); }   // a*b + c
template <class T> inline T FMS_(T a, T b, T c) { return std::fma(a, b, -c); }  // a*b - c
template <class T> inline T FNMS_(T a, T b, T c) { return std::fma(-a, b, c); } // c - a*b

// Codelet constants. Each is a closed form (checked in tests/test_oracle_transforms.py);
// the decimal is rounded to double and then to T, as the reference's float brace-inits do.
template <class T> struct Kc {
    static constexpr T tan_pi_8 = T(0.414213562373095048801688724209698078569671875);        // tan(pi/8)
    static constexpr T two_cos_pi_8 = T(1.847759065022573512256366378793576573644833252);    // 2cos(pi/8)
    static constexpr T tan_pi_16 = T(0.198912367379658006911597622644676228597850501);       // tan(pi/16)
    static constexpr T two_cos_pi_16 = T(1.961570560806460898252364472268478073947867462);   // 2cos(pi/16)
    static constexpr T sqrt2 = T(1.414213562373095048801688724209698078569671875);           // sqrt(2)
    static constexpr T tan_3pi_16 = T(0.668178637919298919997757686523080761552472251);      // tan(3pi/16)
    static constexpr T two_cos_3pi_16 = T(1.662939224605090474157576755235811513477121624);  // 2cos(3pi/16)
    static constexpr T inv_sqrt2 = T(0.707106781186547524400844362104849039284835938);       // 1/sqrt(2)
    static constexpr T two = T(2.0);
    static constexpr T two_sqrt2 = T(2.828427124746190097603377448419396157139343751);       // 2sqrt(2)
    // bior1.5 analysis / synthesis gains
    static constexpr T b_877 = T(0.877670597010003062405456290501163901200537360);
    static constexpr T b_1797 = T(1.797135031972863413496886690073811797696338403);
    static constexpr T b_2277 = T(2.277437593371384746027611934034934695963515886);
    static constexpr T b_1609 = T(1.609389232649111887192845766718020518480884560);
    static constexpr T b_334 = T(0.334024180361136429417383083658457088741315663);
    static constexpr T b_1495 = T(1.495435764250674860795011090214036706658653686);
    static constexpr T b_2058 = T(2.058234225009388964222454285384072231477027482);
    static constexpr T b_486 = T(0.486135912065751423025580498947083714508324707);
};

// ---- 8-point transforms (src/bm3d.cu:32-280) -------------------------------------------------
// Forward DCT-II scaled by 4 (= 4 * orthonormal), FFTW-codelet op order (src/bm3d.cu:34-77).
template <class T> void dct_fwd(T *v) {
    using K = Kc<T>;
    const T a07 = v[0] - v[7], s07 = v[0] + v[7];
    const T a43 = v[4] - v[3], s43 = v[4] + v[3];
    const T a25 = v[2] - v[5];
    const T a16 = v[1] - v[6];
    const T p = a25 + a16, s16 = v[1] + v[6];
    const T m = a25 - a16, s25 = v[2] + v[5];
    const T tb = FNMS_(K::inv_sqrt2, p, a07);
    const T tg = FNMS_(K::inv_sqrt2, m, a43);
    T o[8];
    o[3] = K::two_cos_3pi_16 * FMA_(K::tan_3pi_16, tg, tb);
    o[5] = -(K::two_cos_3pi_16 * FNMS_(K::tan_3pi_16, tb, tg));
    const T tp = s07 + s43, tq = s25 + s16;
    o[4] = K::sqrt2 * (tp - tq);
    o[0] = K::sqrt2 * (tp + tq);
    const T th = FMA_(K::inv_sqrt2, p, a07);
    const T ti = FMA_(K::inv_sqrt2, m, a43);
    o[1] = K::two_cos_pi_16 * FNMS_(K::tan_pi_16, ti, th);
    o[7] = K::two_cos_pi_16 * FMA_(K::tan_pi_16, th, ti);
    const T tl = s07 - s43, to = s25 - s16;
    o[2] = K::two_cos_pi_8 * FNMS_(K::tan_pi_8, to, tl);
    o[6] = K::two_cos_pi_8 * FMA_(K::tan_pi_8, tl, to);
    for (int i = 0; i < 8; ++i) v[i] = o[i];
}

// Inverse: 4 * transpose(orthonormal DCT-II) (src/bm3d.cu:79-121).
template <class T> void dct_inv(T *v) {
    using K = Kc<T>;
    const T t1 = v[0] * K::sqrt2;
    const T t3 = FMA_(K::sqrt2, v[4], t1);
    const T tj = FNMS_(K::sqrt2, v[4], t1);
    const T t6 = FMA_(K::tan_pi_8, v[6], v[2]);
    const T tk = FMS_(K::tan_pi_8, v[2], v[6]);
    const T tb = v[5] + v[3];
    const T te = v[3] - v[5];
    const T tc = FMA_(K::inv_sqrt2, tb, v[1]);
    const T tn = FNMS_(K::inv_sqrt2, te, v[7]);
    const T tf = FMA_(K::inv_sqrt2, te, v[7]);
    const T tm = FNMS_(K::inv_sqrt2, tb, v[1]);
    T o[8];
    const T t7 = FMA_(K::two_cos_pi_8, t6, t3);
    const T tg = FMA_(K::tan_pi_16, tf, tc);
    o[7] = FNMS_(K::two_cos_pi_16, tg, t7);
    o[0] = FMA_(K::two_cos_pi_16, tg, t7);
    const T tp = FNMS_(K::two_cos_pi_8, tk, tj);
    const T tq = FMA_(K::tan_3pi_16, tm, tn);
    o[5] = FNMS_(K::two_cos_3pi_16, tq, tp);
    o[2] = FMA_(K::two_cos_3pi_16, tq, tp);
    const T th = FNMS_(K::two_cos_pi_8, t6, t3);
    const T ti = FNMS_(K::tan_pi_16, tc, tf);
    o[3] = FNMS_(K::two_cos_pi_16, ti, th);
    o[4] = FMA_(K::two_cos_pi_16, ti, th);
    const T tl = FMA_(K::two_cos_pi_8, tk, tj);
    const T to = FNMS_(K::tan_3pi_16, tn, tm);
    o[6] = FNMS_(K::two_cos_3pi_16, to, tl);
    o[1] = FMA_(K::two_cos_3pi_16, to, tl);
    for (int i = 0; i < 8; ++i) v[i] = o[i];
}

// Haar forward, 4 * orthonormal Haar (src/bm3d.cu:127-153).
template <class T> void haar_fwd(T *v) {
    using K = Kc<T>;
    const T s01 = v[0] + v[1], d01 = v[0] - v[1];
    const T s23 = v[2] + v[3], d23 = v[2] - v[3];
    const T s45 = v[4] + v[5], d45 = v[4] - v[5];
    const T s67 = v[6] + v[7], d67 = v[6] - v[7];
    const T a = s01 + s23;
    const T b = K::sqrt2 * (s01 - s23);
    const T c = s45 + s67;
    const T d = K::sqrt2 * (s45 - s67);
    const T sc = K::sqrt2;
    const T sc2 = sc * K::two;
    v[0] = sc * (a + c);
    v[1] = sc * (a - c);
    v[2] = sc * b;
    v[3] = sc * d;
    v[4] = sc2 * d01;
    v[5] = sc2 * d23;
    v[6] = sc2 * d45;
    v[7] = sc2 * d67;
}

// Haar "inverse" exactly as the reference writes it (src/bm3d.cu:154-178): it reads only
// v[0], v[1], v[2], v[4] and is therefore rank 4 -- reproduced literally for parity.
template <class T> void haar_inv(T *v) {
    using K = Kc<T>;
    const T s01 = v[0] + v[1], d01 = v[0] - v[1];
    const T p2 = K::sqrt2 * v[2];
    const T q4 = K::two * v[4];
    const T n2 = (-K::sqrt2) * v[2];
    const T pp = p2 + q4, pm = p2 - q4, np = n2 + q4, nm = n2 - q4;
    const T sc = K::sqrt2;
    v[0] = sc * (s01 + pp);
    v[1] = sc * (s01 + pm);
    v[2] = sc * (s01 + np);
    v[3] = sc * (s01 + nm);
    v[4] = sc * (d01 + pp);
    v[5] = sc * (d01 + pm);
    v[6] = sc * (d01 + np);
    v[7] = sc * (d01 + nm);
}

// Walsh-Hadamard (sequency-permuted, scaled by sqrt(2)); identical forward and inverse
// (src/bm3d.cu:182-212).
template <class T> void wht(T *v) {
    using K = Kc<T>;
    const T a = v[0] + v[1], b = v[0] - v[1];
    const T c = v[2] + v[3], d = v[2] - v[3];
    const T e = v[4] + v[5], f = v[4] - v[5];
    const T g = v[6] + v[7], h = v[6] - v[7];
    const T ac = a + c, a_c = a - c, bd = b + d, b_d = b - d;
    const T eg = e + g, e_g = e - g, fh = f + h, f_h = f - h;
    const T sc = K::sqrt2;
    v[0] = sc * (ac + eg);
    v[1] = sc * (ac - eg);
    v[2] = sc * (a_c - e_g);
    v[3] = sc * (a_c + e_g);
    v[4] = sc * (b_d + f_h);
    v[5] = sc * (b_d - f_h);
    v[6] = sc * (bd - fh);
    v[7] = sc * (bd + fh);
}

// Biorthogonal 1.5 analysis (src/bm3d.cu:216-245).
template <class T> void bior_fwd(T *v) {
    using K = Kc<T>;
    const T s03 = v[0] + v[3], d03 = v[0] - v[3];
    const T s12 = v[1] + v[2], d12 = v[1] - v[2];
    const T s47 = v[4] + v[7], d47 = v[4] - v[7];
    const T s56 = v[5] + v[6], d56 = v[5] - v[6];
    const T e01 = v[0] - v[1], e23 = v[2] - v[3], e45 = v[4] - v[5], e67 = v[6] - v[7];
    T o[8];
    o[0] = K::sqrt2 * (((s03 + s47) + s12) + s56);
    o[1] = (K::b_877 * (s03 - s47)) + (K::b_1797 * (s12 - s56));
    o[2] = ((K::b_2277 * d03) + (K::b_1609 * d12)) + (K::b_334 * (d56 - d47));
    o[3] = ((K::b_2277 * d47) + (K::b_1609 * d56)) + (K::b_334 * (d12 - d03));
    o[4] = K::two_sqrt2 * e01;
    o[5] = K::two_sqrt2 * e23;
    o[6] = K::two_sqrt2 * e45;
    o[7] = K::two_sqrt2 * e67;
    for (int i = 0; i < 8; ++i) v[i] = o[i];
}

// Biorthogonal 1.5 synthesis (src/bm3d.cu:247-279).
template <class T> void bior_inv(T *v) {
    using K = Kc<T>;
    const T u0 = K::sqrt2 * v[0];
    const T u1 = K::b_1495 * v[1];
    const T u2 = K::b_2058 * v[2];
    const T u3 = K::b_2058 * v[3];
    const T a4 = K::two_sqrt2 * v[4], b4 = K::b_486 * v[4];
    const T a5 = K::two_sqrt2 * v[5], b5 = K::b_486 * v[5];
    const T a6 = K::two_sqrt2 * v[6], b6 = K::b_486 * v[6];
    const T a7 = K::two_sqrt2 * v[7], b7 = K::b_486 * v[7];
    const T p = u0 + u1, q = u0 - u1;
    const T r = b5 - b7, s = b4 - b6;
    v[0] = (p + u2) + (a4 - r);
    v[1] = (p + u2) - (a4 + r);
    v[2] = (p - u2) + (s + a5);
    v[3] = (p - u2) + (s - a5);
    v[4] = (q + u3) + (r + a6);
    v[5] = (q + u3) + (r - a6);
    v[6] = (q - u3) - (s - a7);
    v[7] = (q - u3) - (s + a7);
}

template <class T> inline void xform(int kind, bool fwd, T *v) {
    switch (kind) {
    case X_DCT: fwd ? dct_fwd(v) : dct_inv(v); break;
    case X_HAAR: fwd ? haar_fwd(v) : haar_inv(v); break;
    case X_WHT: wht(v); break;
    default: fwd ? bior_fwd(v) : bior_inv(v); break;
    }
}

// ---- block distances (src/bm3d.cu:282-390) ----------------------------------------------------
// 8-lane butterfly result: ((c0+c1)+(c2+c3))+((c4+c5)+(c6+c7)).
template <class T> inline T tree8(const T *c) {
    return ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
}

// a: reference 8x8 patch, b: candidate, both with row stride `as` / `bs` elements.
template <class T> T block_distance(int metric, const T *a, long as, const T *b, long bs) {
    T col[8];
    if (metric == M_SSD || metric == M_SAD) {
        T e0[8], e1[8];
        for (int c = 0; c < 8; ++c) e0[c] = e1[c] = T(0);
        for (int r = 0; r < 8; ++r) {
            for (int c = 0; c < 8; ++c) {
                const T d = a[r * as + c] - b[r * bs + c];
                T &e = (r & 1) ? e1[c] : e0[c];
                if (metric == M_SSD) e = fma_(d, d, e);
                else e = e + std::fabs(d);
            }
        }
        for (int c = 0; c < 8; ++c) col[c] = e0[c] + e1[c];
        return tree8(col);
    }
    if (metric == M_ZSSD || metric == M_ZSAD) {
        T ca[8], cb[8];
        for (int c = 0; c < 8; ++c) {
            const T *pa = a + c;
            const T *pb = b + c;
            ca[c] = ((pa[0] + pa[as]) + (pa[2 * as] + pa[3 * as])) + ((pa[4 * as] + pa[5 * as]) + (pa[6 * as] + pa[7 * as]));
            cb[c] = ((pb[0] + pb[bs]) + (pb[2 * bs] + pb[3 * bs])) + ((pb[4 * bs] + pb[5 * bs]) + (pb[6 * bs] + pb[7 * bs]));
        }
        const T ma = tree8(ca) * T(1.0 / 64.0);
        const T mb = tree8(cb) * T(1.0 / 64.0);
        const T dm = ma - mb;
        T e0[8], e1[8];
        for (int c = 0; c < 8; ++c) e0[c] = e1[c] = T(0);
        for (int r = 0; r < 8; ++r) {
            for (int c = 0; c < 8; ++c) {
                const T d = (a[r * as + c] - b[r * bs + c]) - dm;
                T &e = (r & 1) ? e1[c] : e0[c];
                if (metric == M_ZSSD) e = fma_(d, d, e);
                else e = e + std::fabs(d);
            }
        }
        for (int c = 0; c < 8; ++c) col[c] = e0[c] + e1[c];
        return tree8(col);
    }
    // ssd/norm
    T qa0[8], qa1[8], qb0[8], qb1[8], x0[8], x1[8];
    for (int c = 0; c < 8; ++c) qa0[c] = qa1[c] = qb0[c] = qb1[c] = x0[c] = x1[c] = T(0);
    for (int r = 0; r < 8; ++r) {
        for (int c = 0; c < 8; ++c) {
            const T va = a[r * as + c], vb = b[r * bs + c];
            if (r & 1) {
                qa1[c] = fma_(va, va, qa1[c]);
                qb1[c] = fma_(vb, vb, qb1[c]);
                x1[c] = fma_(va, vb, x1[c]);
            } else {
                qa0[c] = fma_(va, va, qa0[c]);
                qb0[c] = fma_(vb, vb, qb0[c]);
                x0[c] = fma_(va, vb, x0[c]);
            }
        }
    }
    T ccol[8], ncol[8];
    for (int c = 0; c < 8; ++c) {
        ccol[c] = qa0[c] + qa1[c];
        ncol[c] = qb0[c] + qb1[c];
        col[c] = x0[c] + x1[c];
    }
    const T na = std::sqrt(tree8(ccol));
    const T nb = std::sqrt(tree8(ncol));
    const T dot = tree8(col);
    return T(2) - (T(2) * dot) / ((na * nb) + T(FLT_EPSILON));
}

// ---- matching ----------------------------------------------------------------------------------
template <class T> struct Cand {
    T err;
    int x, y, z;
};

// Strict-< sorted insertion (src/bm3d.cu:641-653): ties keep the earlier entry in front.
template <class T> inline void insert8(Cand<T> *l, T e, int x, int y, int z) {
    if (!(e < l[7].err)) return;
    int k = 7;
    while (k > 0 && e < l[k - 1].err) {
        l[k] = l[k - 1];
        --k;
    }
    l[k] = Cand<T>{e, x, y, z};
}

struct Geometry {
    int width, height, stride;
    int block_step, bm_range, radius, ps_num, ps_range;
    int nbx, nby;
};

inline int ref_origin(int b, int step, int extent) { return std::min(b * step, extent - 8); }

// Block matching for the reference block (bx, by) over `img` laid out [TW][H][S].
template <class T>
void match_block(const Geometry &g, int metric, const T *img, int bx, int by, Cand<T> *out) {
    const int W = g.width, H = g.height, S = g.stride, R = g.radius;
    const long TS = long(H) * S;
    const int x = ref_origin(bx, g.block_step, W);
    const int y = ref_origin(by, g.block_step, H);
    const T *centre = img + R * TS;
    const T *refp = centre + long(y) * S + x;

    Cand<T> l[8];
    for (int k = 0; k < 8; ++k) l[k] = Cand<T>{T(FLT_MAX), 0, 0, R};

    // spatial window, row-major (src/bm3d.cu:618-660)
    {
        const int left = std::max(x - g.bm_range, 0), right = std::min(x + g.bm_range, W - 8);
        const int top = std::max(y - g.bm_range, 0), bottom = std::min(y + g.bm_range, H - 8);
        for (int yy = top; yy <= bottom; ++yy)
            for (int xx = left; xx <= right; ++xx)
                insert8(l, block_distance<T>(metric, refp, S, centre + long(yy) * S + xx, S), xx, yy, R);
    }

    // temporal predictive search (src/bm3d.cu:663-756)
    if (R > 0) {
        Cand<T> centre_list[8];
        for (int k = 0; k < 8; ++k) centre_list[k] = l[k];
        for (int dir = -1; dir <= 1; dir += 2) {
            Cand<T> last[8];
            for (int k = 0; k < 8; ++k) last[k] = centre_list[k];
            for (int t = 1; t <= R; ++t) {
                const int tz = R + dir * t;
                const T *frame = img + long(tz) * TS;
                Cand<T> f[8];
                for (int k = 0; k < 8; ++k) f[k] = Cand<T>{T(FLT_MAX), 0, 0, tz};
                for (int i = 0; i < g.ps_num; ++i) {
                    const int cx = last[i].x, cy = last[i].y;
                    const int left = std::max(cx - g.ps_range, 0), right = std::min(cx + g.ps_range, W - 8);
                    const int top = std::max(cy - g.ps_range, 0), bottom = std::min(cy + g.ps_range, H - 8);
                    for (int yy = top; yy <= bottom; ++yy)
                        for (int xx = left; xx <= right; ++xx)
                            insert8(f, block_distance<T>(metric, refp, S, frame + long(yy) * S + xx, S), xx, yy, tz);
                }
                for (int i = 0; i < g.ps_num; ++i) insert8(l, f[i].err, f[i].x, f[i].y, tz);
                for (int k = 0; k < 8; ++k) last[k] = f[k];
            }
        }
    }

    // self-inclusion (src/bm3d.cu:758-788)
    bool found = false;
    for (int k = 0; k < 8; ++k)
        if (l[k].x == x && l[k].y == y && (R == 0 || l[k].z == R)) found = true;
    if (!found) {
        for (int k = 7; k > 0; --k) l[k] = l[k - 1];
        l[0] = Cand<T>{T(0), x, y, R};
    }
    for (int k = 0; k < 8; ++k) out[k] = l[k];
}

// ---- collaborative filtering (src/bm3d.cu:416-546) ---------------------------------------------
// grp[i][j][c]: i = match rank (group axis), j = row, c = column.
template <class T> void forward3d(int t2d, int t1d, T (*grp)[8][8]) {
    T v[8];
    for (int i = 0; i < 8; ++i) {
        for (int c = 0; c < 8; ++c) {  // vertical pass: along rows, per column
            for (int j = 0; j < 8; ++j) v[j] = grp[i][j][c];
            xform<T>(t2d, true, v);
            for (int j = 0; j < 8; ++j) grp[i][j][c] = v[j];
        }
        for (int j = 0; j < 8; ++j) xform<T>(t2d, true, grp[i][j]);  // horizontal pass
    }
    for (int j = 0; j < 8; ++j)
        for (int c = 0; c < 8; ++c) {  // group axis
            for (int i = 0; i < 8; ++i) v[i] = grp[i][j][c];
            xform<T>(t1d, true, v);
            for (int i = 0; i < 8; ++i) grp[i][j][c] = v[i];
        }
}

template <class T> void inverse3d(int t2d, int t1d, T (*grp)[8][8]) {
    T v[8];
    for (int i = 0; i < 8; ++i) {
        for (int c = 0; c < 8; ++c) {
            for (int j = 0; j < 8; ++j) v[j] = grp[i][j][c];
            xform<T>(t2d, false, v);
            for (int j = 0; j < 8; ++j) grp[i][j][c] = v[j];
        }
        for (int j = 0; j < 8; ++j) xform<T>(t2d, false, grp[i][j]);
    }
    for (int j = 0; j < 8; ++j)
        for (int c = 0; c < 8; ++c) {
            for (int i = 0; i < 8; ++i) v[i] = grp[i][j][c];
            xform<T>(t1d, false, v);
            for (int i = 0; i < 8; ++i) grp[i][j][c] = v[i];
        }
}

// Hard thresholding: returns the adaptive weight 1/k (src/bm3d.cu:416-455).
template <class T> T hard_threshold(T (*grp)[8][8], T sigma) {
    T kc[8];
    for (int c = 0; c < 8; ++c) {
        T k = T(0);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) {
                const T val = grp[i][j][c];
                const T thr = (i == 0 && j == 0 && c == 0) ? T(0) : sigma;
                const bool keep = std::fabs(val) >= thr;
                k = k + (keep ? T(1) : T(0));
                grp[i][j][c] = keep ? val * T(1.0 / 4096.0) : T(0);
            }
        kc[c] = k;
    }
    return T(1) / tree8(kc);
}

// Empirical Wiener shrinkage against the transformed basic estimate (src/bm3d.cu:480-515).
template <class T> T wiener_shrink(T (*grp)[8][8], T (*ref)[8][8], T sigma) {
    T kc[8];
    const T s2 = sigma * sigma;
    for (int c = 0; c < 8; ++c) {
        T k = T(0);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) {
                const T r = ref[i][j][c];
                const T r2 = r * r;
                T coeff = r2 / (r2 + s2);
                if (i == 0 && j == 0 && c == 0) coeff = T(1);
                const T val = grp[i][j][c] * coeff;
                k = fma_(coeff, coeff, k);
                grp[i][j][c] = val * T(1.0 / 4096.0);
            }
        kc[c] = k;
    }
    return T(1) / tree8(kc);
}

}  // namespace

// =================================================================================================
// C ABI for the Python test harness (ctypes).
// =================================================================================================
extern "C" {

struct OracleParams {
    int width, height, stride;
    int block_step, bm_range, radius, ps_num, ps_range;
    int chroma;      // 1: 3 planes share plane-0 matches (CBM3D)
    int final_pass;  // 1: Wiener pass, src holds [ref clip, noisy clip]
    float sigma[3];  // scaled thresholds (SIGMA_Y/U/V)
    float extractor; // E = 2^extractor_exp, or 0
    int extractor_exp;
    int metric;      // 0 ssd, 1 sad, 2 zssd, 3 zsad, 4 ssd/norm
    int t2d, t1d;    // 0 dct, 1 haar, 2 wht, 3 bior1.5
    int threads;     // 0 = all hardware threads
    int by_begin, by_end;  // block rows to process, [by_begin, by_end); by_end 0 = all rows
    int float_sums;        // 1: FP32 sums in reference-block raster order even with a non-zero extractor
};

}  // extern "C"

namespace {

int num_threads(int req) {
    if (req > 0) return req;
    unsigned h = std::thread::hardware_concurrency();
    return h ? int(h) : 1;
}

template <class Fn> void parallel_for(int n, int threads, Fn fn) {
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto &th : pool) th.join();
}

Geometry make_geometry(const OracleParams &p) {
    Geometry g{};
    g.width = p.width;
    g.height = p.height;
    g.stride = p.stride;
    g.block_step = p.block_step;
    g.bm_range = p.bm_range;
    g.radius = p.radius;
    g.ps_num = p.ps_num;
    g.ps_range = p.ps_range;
    // Reference blocks per axis: origins b*step while b*step < extent-8+step (src/bm3d.cu:588-592).
    g.nbx = (p.width - 8 + p.block_step - 1) / p.block_step + 1;
    g.nby = (p.height - 8 + p.block_step - 1) / p.block_step + 1;
    return g;
}

int check_params(const OracleParams &p) {
    if (p.width < 8 || p.height < 8 || p.stride < p.width) return -1;
    if (p.block_step < 1 || p.block_step > 8 || p.bm_range < 1 || p.radius < 0 || p.radius > 16) return -1;
    if (p.ps_num < 1 || p.ps_num > 8 || p.ps_range < 1) return -1;
    if (p.metric < 0 || p.metric > 4 || p.t2d < 0 || p.t2d > 3 || p.t1d < 0 || p.t1d > 3) return -1;
    return 0;
}

// Filtered output of one group on one plane.
template <class T> struct GroupOut {
    T weight;
    T vals[8][8][8];  // [rank][row][col], already multiplied by the weight and quantised
    bool active;
};

// Quantisation q(u) = (u + E) - E (src/bm3d.cu:884-885).
template <class T> inline T quantise(T u, T E) { return (u + E) - E; }

// Exact-sum scale: contributions are multiples of 2^(e-24) (see header).
inline int64_t to_fixed(float q, int e) {
    float s = std::ldexp(q, 24 - e);
    s = std::fmin(std::fmax(s, -4.611686018427387904e18f), 4.611686018427387904e18f);
    return int64_t(s);
}
inline float from_fixed(int64_t v, int e) { return std::ldexp(float(v), e - 24); }

inline bool exact_path(const OracleParams &p) {
    return p.extractor != 0.0f && p.extractor_exp > -40 && p.extractor_exp < 40;
}

template <class T>
int run_bm3d(const OracleParams &p, const T *src, T *res, int32_t *matches_out, T *match_err,
             const int32_t *given) {
    if (check_params(p)) return -1;
    const Geometry g = make_geometry(p);
    const int H = p.height, S = p.stride, R = p.radius, TW = 2 * R + 1;
    const int P = p.chroma ? 3 : 1;
    const long TS = long(H) * S;
    const long PLANE = long(TW) * TS;           // one plane of one clip: [TW][H][S]
    const long CLIP = long(P) * PLANE;          // one clip: [P][TW][H][S]
    const int threads = num_threads(p.threads);
    const T sig[3] = {T(p.sigma[0]), T(p.sigma[1]), T(p.sigma[2])};
    const bool exact = std::is_same<T, float>::value && exact_path(p) && !p.float_sums;
    const T E = T(p.extractor);

    const long res_elems = long(P) * TW * 2 * TS;
    std::vector<int64_t> acc_fixed(exact ? res_elems : 0, 0);
    for (long i = 0; i < res_elems; ++i) res[i] = T(0);

    const int CHUNK_ROWS = 4;
    const long blocks_per_chunk = long(CHUNK_ROWS) * g.nbx;
    std::vector<Cand<T>> lists(blocks_per_chunk * 8);
    std::vector<GroupOut<T>> outs(blocks_per_chunk * P);

    const int by_first = std::max(0, p.by_begin);
    const int by_last = p.by_end > 0 ? std::min(p.by_end, g.nby) : g.nby;
    for (int by0 = by_first; by0 < by_last; by0 += CHUNK_ROWS) {
        const int by1 = std::min(by0 + CHUNK_ROWS, by_last);
        const int nblk = (by1 - by0) * g.nbx;
        parallel_for(nblk, threads, [&](int bi) {
            const int by = by0 + bi / g.nbx, bx = bi % g.nbx;
            const long gid = long(by) * g.nbx + bx;
            Cand<T> *l = &lists[long(bi) * 8];
            if (given) {
                for (int k = 0; k < 8; ++k) {
                    const int32_t *m = given + (gid * 8 + k) * 3;
                    l[k] = Cand<T>{T(0), m[0], m[1], m[2]};
                }
            } else {
                // Matching reads plane 0 of clip 0 (the ref clip in the final pass).
                match_block<T>(g, p.metric, src, bx, by, l);
            }
            if (matches_out)
                for (int k = 0; k < 8; ++k) {
                    int32_t *m = matches_out + (gid * 8 + k) * 3;
                    m[0] = l[k].x;
                    m[1] = l[k].y;
                    m[2] = l[k].z;
                }
            if (match_err && !given)
                for (int k = 0; k < 8; ++k) match_err[gid * 8 + k] = l[k].err;

            for (int pl = 0; pl < P; ++pl) {
                GroupOut<T> &o = outs[long(bi) * P + pl];
                o.active = !(p.chroma && sig[pl] < T(FLT_EPSILON));  // src/bm3d.cu:804-810
                if (!o.active) continue;
                const T *plane_ref = src + pl * PLANE;          // clip 0
                const T *plane_noisy = plane_ref + CLIP;         // clip 1 (final pass only)
                T grp[8][8][8];
                T refg[8][8][8];
                for (int i = 0; i < 8; ++i) {
                    const long off = long(l[i].z) * TS + long(l[i].y) * S + l[i].x;
                    for (int j = 0; j < 8; ++j)
                        for (int c = 0; c < 8; ++c) {
                            if (p.final_pass) {
                                refg[i][j][c] = plane_ref[off + long(j) * S + c];
                                grp[i][j][c] = plane_noisy[off + long(j) * S + c];
                            } else {
                                grp[i][j][c] = plane_ref[off + long(j) * S + c];
                            }
                        }
                }
                T w;
                forward3d<T>(p.t2d, p.t1d, grp);
                if (p.final_pass) {
                    forward3d<T>(p.t2d, p.t1d, refg);
                    w = wiener_shrink<T>(grp, refg, sig[pl]);
                } else {
                    w = hard_threshold<T>(grp, sig[pl]);
                }
                inverse3d<T>(p.t2d, p.t1d, grp);
                o.weight = quantise<T>(w, E);
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j)
                        for (int c = 0; c < 8; ++c) o.vals[i][j][c] = quantise<T>(w * grp[i][j][c], E);
            }
        });

        if (exact) {
            // Integer sums do not depend on order: each thread owns a band of image rows and
            // adds every contribution that lands in it.
            const int bands = std::min(threads, H);
            parallel_for(bands, bands, [&](int t) {
                const int y0 = int(long(H) * t / bands), y1 = int(long(H) * (t + 1) / bands);
                for (int bi = 0; bi < nblk; ++bi) {
                    const Cand<T> *l = &lists[long(bi) * 8];
                    for (int pl = 0; pl < P; ++pl) {
                        const GroupOut<T> &o = outs[long(bi) * P + pl];
                        if (!o.active) continue;
                        const int64_t wq = to_fixed(float(o.weight), p.extractor_exp);
                        const long pbase = long(pl) * TW * 2 * TS;
                        for (int i = 0; i < 8; ++i) {
                            const long base = pbase + long(l[i].z) * 2 * TS + long(l[i].y) * S + l[i].x;
                            for (int j = 0; j < 8; ++j) {
                                const int y = l[i].y + j;
                                if (y < y0 || y >= y1) continue;
                                for (int c = 0; c < 8; ++c) {
                                    const long a = base + long(j) * S + c;
                                    acc_fixed[a] += to_fixed(float(o.vals[i][j][c]), p.extractor_exp);
                                    acc_fixed[a + TS] += wq;
                                }
                            }
                        }
                    }
                }
            });
            continue;
        }
        // Float sums: sequential, in reference-block raster order.
        for (int bi = 0; bi < nblk; ++bi) {
            const Cand<T> *l = &lists[long(bi) * 8];
            for (int pl = 0; pl < P; ++pl) {
                const GroupOut<T> &o = outs[long(bi) * P + pl];
                if (!o.active) continue;
                const long pbase = long(pl) * TW * 2 * TS;
                for (int i = 0; i < 8; ++i) {
                    const long base = pbase + long(l[i].z) * 2 * TS + long(l[i].y) * S + l[i].x;
                    for (int j = 0; j < 8; ++j)
                        for (int c = 0; c < 8; ++c) {
                            const long a = base + long(j) * S + c;
                            res[a] = res[a] + o.vals[i][j][c];
                            res[a + TS] = res[a + TS] + o.weight;
                        }
                }
            }
        }
    }
    if (exact)
        for (long i = 0; i < res_elems; ++i) res[i] = T(from_fixed(acc_fixed[i], p.extractor_exp));
    return 0;
}

}  // namespace

extern "C" {

int oracle_version() { return 1; }

int oracle_transform_f64(int kind, int forward, double *v) {
    if (kind < 0 || kind > 3) return -1;
    xform<double>(kind, forward != 0, v);
    return 0;
}

int oracle_transform_f32(int kind, int forward, float *v) {
    if (kind < 0 || kind > 3) return -1;
    xform<float>(kind, forward != 0, v);
    return 0;
}

float oracle_block_distance_f32(int metric, const float *a, long as, const float *b, long bs) {
    return block_distance<float>(metric, a, as, b, bs);
}

double oracle_block_distance_f64(int metric, const double *a, long as, const double *b, long bs) {
    return block_distance<double>(metric, a, as, b, bs);
}

// 3-D forward transform + shrinkage + inverse of one group (for unit tests).
// grp: [8][8][8] (rank, row, col); ref: same layout (Wiener only, may be null for hard).
float oracle_filter_group_f32(int t2d, int t1d, int wiener, float sigma, float *grp, float *ref) {
    auto *g = reinterpret_cast<float(*)[8][8]>(grp);
    auto *r = reinterpret_cast<float(*)[8][8]>(ref);
    forward3d<float>(t2d, t1d, g);
    float w;
    if (wiener) {
        forward3d<float>(t2d, t1d, r);
        w = wiener_shrink<float>(g, r, sigma);
    } else {
        w = hard_threshold<float>(g, sigma);
    }
    inverse3d<float>(t2d, t1d, g);
    return w;
}

int oracle_num_blocks(const OracleParams *p, int *nbx, int *nby) {
    if (check_params(*p)) return -1;
    const Geometry g = make_geometry(*p);
    *nbx = g.nbx;
    *nby = g.nby;
    return 0;
}

// src: [(final?2:1)][P][TW][H][S], res: [P][TW][2][H][S]; matches (optional): [nby][nbx][8][3]
// as (x, y, z); match_err (optional): [nby][nbx][8]; given (optional) replaces the matching.
int oracle_bm3d_f32(const OracleParams *p, const float *src, float *res, int32_t *matches, float *match_err,
                    const int32_t *given) {
    return run_bm3d<float>(*p, src, res, matches, match_err, given);
}

int oracle_bm3d_f64(const OracleParams *p, const double *src, double *res, int32_t *matches, double *match_err,
                    const int32_t *given) {
    return run_bm3d<double>(*p, src, res, matches, match_err, given);
}

// Matching only (for timing and index parity).
int oracle_match_f32(const OracleParams *p, const float *img, int32_t *matches, float *match_err) {
    if (check_params(*p)) return -1;
    const Geometry g = make_geometry(*p);
    parallel_for(g.nbx * g.nby, num_threads(p->threads), [&](int gid) {
        Cand<float> l[8];
        match_block<float>(g, p->metric, img, gid % g.nbx, gid / g.nbx, l);
        for (int k = 0; k < 8; ++k) {
            if (matches) {
                matches[(long(gid) * 8 + k) * 3 + 0] = l[k].x;
                matches[(long(gid) * 8 + k) * 3 + 1] = l[k].y;
                matches[(long(gid) * 8 + k) * 3 + 2] = l[k].z;
            }
            if (match_err) match_err[long(gid) * 8 + k] = l[k].err;
        }
    });
    return 0;
}

// Matching of selected blocks only (block ids are raster indices by * nbx + bx).
int oracle_match_blocks_f32(const OracleParams *p, const float *img, const int32_t *ids, int n, int32_t *matches,
                            float *match_err) {
    if (check_params(*p)) return -1;
    const Geometry g = make_geometry(*p);
    parallel_for(n, num_threads(p->threads), [&](int i) {
        const int gid = ids[i];
        Cand<float> l[8];
        match_block<float>(g, p->metric, img, gid % g.nbx, gid / g.nbx, l);
        for (int k = 0; k < 8; ++k) {
            matches[(long(i) * 8 + k) * 3 + 0] = l[k].x;
            matches[(long(i) * 8 + k) * 3 + 1] = l[k].y;
            matches[(long(i) * 8 + k) * 3 + 2] = l[k].z;
            if (match_err) match_err[long(i) * 8 + k] = l[k].err;
        }
    });
    return 0;
}

// dst[p] = num / den for planes in proc_mask (src/bm3d.cu:898-922). res: [P][2][H][S].
int oracle_aggregate_f32(int P, int width, int height, int stride, unsigned proc_mask, const float *res,
                         float *dst) {
    const long TS = long(height) * stride;
    for (int p = 0; p < P; ++p) {
        if (!((proc_mask >> p) & 1)) continue;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const long i = long(y) * stride + x;
                dst[p * TS + i] = res[p * 2 * TS + i] / res[p * 2 * TS + TS + i];
            }
    }
    return 0;
}

int oracle_aggregate_f64(int P, int width, int height, int stride, unsigned proc_mask, const double *res,
                         double *dst) {
    const long TS = long(height) * stride;
    for (int p = 0; p < P; ++p) {
        if (!((proc_mask >> p) & 1)) continue;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const long i = long(y) * stride + x;
                dst[p * TS + i] = res[p * 2 * TS + i] / res[p * 2 * TS + TS + i];
            }
    }
    return 0;
}

// Slot index of window position i for output frame n (src/bm3d.zig:46-48).
int oracle_agg_z(int i, int n, int nframes, int radius) {
    return std::min(std::max(2 * radius - i, n - nframes + 1 + radius), n + radius);
}

// Temporal aggregation of one plane (src/bm3d_vagg.cu:18-48, src/bm3d.zig:1284-1317):
// slots[i] points at the [TW][2][H][S] accumulator stack of centre frame clamp(n-r+i);
// the sum runs over i ascending, then one IEEE division.
int oracle_vaggregate_f32(int width, int height, int stride, int radius, int n, int nframes,
                          const float *const *slots, float *dst) {
    const int TW = 2 * radius + 1;
    const long TS = long(height) * stride;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const long i = long(y) * stride + x;
            float sv = 0.0f, sw = 0.0f;
            for (int t = 0; t < TW; ++t) {
                const int z = oracle_agg_z(t, n, nframes, radius);
                const float *base = slots[t] + long(z) * 2 * TS;
                sv = sv + base[i];
                sw = sw + base[TS + i];
            }
            dst[i] = sv / sw;
        }
    return 0;
}

}  // extern "C"
