// Learned basic pass: a residual CNN of six 3x3 convolutions (32 channels, ReLU between them,
// a noise-level map as the second input channel) that stands in for the hard-threshold pass and
// supplies the Wiener pass's reference. Training and inference kernels for gfx950: BF16 weights
// and activations, FP32 accumulation on v_mfma_f32_16x16x32_bf16.
//
// Activations are [N][H][W][32] bf16 (64 bytes per pixel). Layer 1's two input channels and
// layer 6's single output channel are zero-padded to 32 so every layer runs the same kernel.
// Weights are [9 taps][32 out][32 in] bf16 (tap = (dy+1)*3 + (dx+1)), biases fp32.
//
// MFMA orientation (lane layout checked on the GPU): D[out][pixel] += W[out][in] * X[in][pixel].
// A = weights: lane l holds out = l%16 (+16 for the second half), in = 8*(l/16)..+7.
// B = activations: lane l holds pixel l%16 of the row, in = 8*(l/16)..+7 (16 contiguous bytes).
// D: lane l holds out = 4*(l/16)..+3 for pixel l%16, so each lane stores 4 channels of a pixel.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <string.h>

#include "vszipcu_bm3d.h"
#include "vszipcu_cnn.h"
#include "internal.h"

namespace vszcnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CH = 32;
constexpr int TILE = 16;             // output tile: 16 x 16 pixels, 4 waves x 4 rows
constexpr int HALO = TILE + 2;
constexpr int PIX_BYTES = 80;        // 64 bytes of channels + 16 bytes of bank padding

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float bf2f(unsigned short u) { return __uint_as_float((unsigned)u << 16); }

// The (TILE+2)^2 input halo of image n (zeros outside the image), in registers: 16 bytes per
// thread and step, loads issued back to back so their latencies overlap.
constexpr int kHaloParts = HALO * HALO * 4, kHaloPer = (kHaloParts + 255) / 256;

__device__ __forceinline__ void halo_load(const unsigned short *x, int n, int H, int W, int x0, int y0,
                                          uint4 (&v)[kHaloPer]) {
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        const int pix = i >> 2, part = i & 3;
        const int hy = pix / HALO, hx = pix % HALO;
        const int gy = y0 + hy - 1, gx = x0 + hx - 1;
        v[k] = make_uint4(0, 0, 0, 0);
        if (i < kHaloParts && gy >= 0 && gy < H && gx >= 0 && gx < W)
            v[k] = *reinterpret_cast<const uint4 *>(x + (((long long)n * H + gy) * W + gx) * CH + part * 8);
    }
}

__device__ __forceinline__ void halo_store(char *lds, const uint4 (&v)[kHaloPer]) {
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i < kHaloParts) *reinterpret_cast<uint4 *>(lds + (i >> 2) * PIX_BYTES + (i & 3) * 16) = v[k];
    }
}

// Stages the halo of image n into LDS.
__device__ __forceinline__ void stage_halo(const unsigned short *x, int n, int H, int W, int x0, int y0, char *lds) {
    uint4 v[kHaloPer];
    halo_load(x, n, H, W, x0, y0, v);
    halo_store(lds, v);
}

// y = conv3x3(x) with the epilogue of `mode`:
//   0: relu(acc + bias) -> bf16 [N][H][W][32]
//   1: base[n][y][x] + acc[out 0] + bias[0] -> fp32 [N][H][ostride] (the residual output layer)
//   2: acc * (aux[n][y][x][out] > 0) -> bf16 (backward data through the previous ReLU)
// Persistent: workgroup g walks the 16x16 tiles g, g + G, ... (row-major over images) and loads
// the next tile's halo into registers while it computes the current one from LDS; the weights
// are loaded once. Wave w computes output channels hh*16 .. +15 (hh = w & 1) of rows
// (w >> 1)*8 .. +7 of a tile.
__global__ __launch_bounds__(256) void k_conv3x3(const unsigned short *__restrict__ x, const unsigned short *__restrict__ w,
                                                 const float *__restrict__ bias, int N, int H, int W, int tiles, int mode,
                                                 const unsigned short *__restrict__ aux, const float *__restrict__ base,
                                                 void *__restrict__ y, int ostride) {
    __shared__ __attribute__((aligned(16))) char lds[HALO * HALO * PIX_BYTES];
    const int tx_n = (W + TILE - 1) / TILE, ty_n = (H + TILE - 1) / TILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int hh = wave & 1, r0 = (wave >> 1) * 8;
    const bool active = !(mode == 1 && hh == 1);  // the residual output layer needs output channel 0 only
    bf16x8 wa[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) wa[t] = *reinterpret_cast<const bf16x8 *>(w + ((t * CH) + hh * 16 + j) * CH + g * 8);
    const int o = hh * 16 + g * 4;
    float bo[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (mode != 2)
#pragma unroll
        for (int q = 0; q < 4; ++q) bo[q] = bias[mode == 1 ? 0 : o + q];

    uint4 v[kHaloPer];
    int tile = blockIdx.x;
    if (tile < tiles) halo_load(x, tile / (tx_n * ty_n), H, W, (tile % tx_n) * TILE, ((tile / tx_n) % ty_n) * TILE, v);
    for (; tile < tiles; tile += gridDim.x) {
        const int n = tile / (tx_n * ty_n), x0 = (tile % tx_n) * TILE, y0 = ((tile / tx_n) % ty_n) * TILE;
        __syncthreads();  // every wave is done with the previous tile's halo
        halo_store(lds, v);
        __syncthreads();
        const int nt = tile + gridDim.x;
        if (nt < tiles) halo_load(x, nt / (tx_n * ty_n), H, W, (nt % tx_n) * TILE, ((nt / tx_n) % ty_n) * TILE, v);
        if (!active) continue;
#pragma unroll 2
        for (int rr = 0; rr < 8; ++rr) {
            const int r = r0 + rr;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int ky = t / 3, kx = t % 3;
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(lds + ((r + ky) * HALO + j + kx) * PIX_BYTES + g * 16);
                acc = mfma(wa[t], b, acc);
            }
            const int gy = y0 + r, gx = x0 + j;
            if (gy >= H || gx >= W) continue;
            const long long pix = ((long long)n * H + gy) * W + gx;
            if (mode == 1) {
                if (g == 0) static_cast<float *>(y)[((long long)n * H + gy) * ostride + gx] = base[pix] + (acc[0] + bo[0]);
                continue;
            }
            float vv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) vv[q] = acc[q];
            if (mode == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) vv[q] = fmaxf(vv[q] + bo[q], 0.0f);
            } else {
                const uint2 m = *reinterpret_cast<const uint2 *>(aux + pix * CH + o);
                const unsigned short mm[4] = {(unsigned short)(m.x & 0xffff), (unsigned short)(m.x >> 16),
                                              (unsigned short)(m.y & 0xffff), (unsigned short)(m.y >> 16)};
#pragma unroll
                for (int q = 0; q < 4; ++q) vv[q] = bf2f(mm[q]) > 0.0f ? vv[q] : 0.0f;
            }
            bf16x4 o4;
#pragma unroll
            for (int q = 0; q < 4; ++q) o4[q] = (__bf16)vv[q];
            *reinterpret_cast<bf16x4 *>(static_cast<unsigned short *>(y) + pix * CH + o) = o4;
        }
    }
}

// Weight gradient dW[t][out][in] += sum over pixels of dy[p][out] * x[p + t][in], FP32.
// A workgroup walks `tiles` consecutive 16x16 tiles (row-major over images) accumulating in
// registers, then adds its partial sums with device atomics. Wave w owns the (out half, in half)
// pair w for all 9 taps. LDS holds dy transposed ([out][16 rows][16 cols]) and x transposed in
// three horizontally shifted copies ([dx][in][18 rows][16 cols]) so every MFMA operand is 16
// aligned bytes.
constexpr int DY_CH_BYTES = TILE * TILE * 2 + 16;       // 528
constexpr int XS_CH_BYTES = HALO * TILE * 2 + 16;        // 592

__global__ __launch_bounds__(256) void k_conv3x3_dw(const unsigned short *__restrict__ x, const unsigned short *__restrict__ dy,
                                                    int N, int H, int W, int tiles, float *__restrict__ dw) {
    __shared__ __attribute__((aligned(16))) char sdy[CH * DY_CH_BYTES];
    __shared__ __attribute__((aligned(16))) char sx[3 * CH * XS_CH_BYTES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, g = lane >> 4;
    const int oh = wave >> 1, ih = wave & 1;
    const int tx_n = (W + TILE - 1) / TILE, ty_n = (H + TILE - 1) / TILE;
    const int total = N * tx_n * ty_n;

    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k = 0; k < tiles; ++k) {
        const int tile = blockIdx.x * tiles + k;
        if (tile >= total) break;
        const int n = tile / (tx_n * ty_n), rem = tile % (tx_n * ty_n);
        const int x0 = (rem % tx_n) * TILE, y0 = (rem / tx_n) * TILE;
        __syncthreads();  // previous tile's LDS reads are done
        // dy tile, transposed; pixels outside the image contribute zero
        for (int i = threadIdx.x; i < TILE * TILE * 4; i += 256) {
            const int pix = i >> 2, part = i & 3;
            const int ry = pix / TILE, rx = pix % TILE;
            const int gy = y0 + ry, gx = x0 + rx;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy < H && gx < W) v = *reinterpret_cast<const uint4 *>(dy + (((long long)n * H + gy) * W + gx) * CH + part * 8);
            const unsigned short *e = reinterpret_cast<const unsigned short *>(&v);
#pragma unroll
            for (int c = 0; c < 8; ++c)
                *reinterpret_cast<unsigned short *>(sdy + (part * 8 + c) * DY_CH_BYTES + (ry * TILE + rx) * 2) = e[c];
        }
        // x halo rows -1..16, copies for dx = -1, 0, +1
        for (int i = threadIdx.x; i < HALO * HALO * 4; i += 256) {
            const int pix = i >> 2, part = i & 3;
            const int hy = pix / HALO, hx = pix % HALO;  // halo coordinates: gx = x0 + hx - 1
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                v = *reinterpret_cast<const uint4 *>(x + (((long long)n * H + gy) * W + gx) * CH + part * 8);
            const unsigned short *e = reinterpret_cast<const unsigned short *>(&v);
#pragma unroll
            for (int s = 0; s < 3; ++s) {  // copy s holds column c = hx - s (dx = s - 1)
                const int c = hx - s;
                if (c < 0 || c >= TILE) continue;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    *reinterpret_cast<unsigned short *>(sx + (s * CH + part * 8 + q) * XS_CH_BYTES + (hy * TILE + c) * 2) = e[q];
            }
        }
        __syncthreads();
        // K = 256 pixels in 8 chunks of 32: chunk c = rows 2c, 2c+1; lane group g takes row
        // 2c + (g >> 1), columns 8*(g & 1)..+7
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ry = 2 * c + (g >> 1), cx = 8 * (g & 1);
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(sdy + (oh * 16 + j) * DY_CH_BYTES + (ry * TILE + cx) * 2);
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int ky = t / 3, kx = t % 3;  // x row = ry + ky - 1 -> halo row ry + ky
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(sx + (kx * CH + ih * 16 + j) * XS_CH_BYTES +
                                                                   ((ry + ky) * TILE + cx) * 2);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
    }
    // D lane l: out = oh*16 + 4*g + q, in = ih*16 + j
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) atomicAdd(dw + (t * CH + oh * 16 + g * 4 + q) * CH + ih * 16 + j, acc[t][q]);
}

// db[out] += sum over pixels of dy[p][out].
__global__ __launch_bounds__(256) void k_bias_grad(const unsigned short *__restrict__ dy, long long pixels, float *__restrict__ db) {
    __shared__ float part[8][CH];
    const int c = threadIdx.x & 31, r = threadIdx.x >> 5;
    float s = 0.0f;
    for (long long p = (long long)blockIdx.x * 8 + r; p < pixels; p += (long long)gridDim.x * 8) s += bf2f(dy[p * CH + c]);
    part[r][c] = s;
    __syncthreads();
    if (r == 0) {
        float t = 0.0f;
        for (int k = 0; k < 8; ++k) t += part[k][c];
        atomicAdd(db + c, t);
    }
}

// x0[n][y][x] = (noisy, sigma_map, 0 ... 0) as bf16; noisy is fp32 [N][H][stride].
__global__ __launch_bounds__(256) void k_pack_input(const float *__restrict__ noisy, int H, int W, int stride,
                                                    const float *__restrict__ sigma_map, unsigned short *__restrict__ x0,
                                                    float *__restrict__ base) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= (long long)H * W) return;
    const int yy = (int)(i / W), xx = (int)(i % W);
    const float v = noisy[((long long)n * H + yy) * stride + xx];
    const long long pix = (long long)n * H * W + i;
    __bf16 h[8] = {(__bf16)v, (__bf16)sigma_map[n], (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f,
                   (__bf16)0.0f, (__bf16)0.0f};
    uint4 *dst = reinterpret_cast<uint4 *>(x0 + pix * CH);
    dst[0] = *reinterpret_cast<const uint4 *>(h);
    dst[1] = dst[2] = dst[3] = make_uint4(0, 0, 0, 0);
    if (base) base[pix] = v;
}

// Mean squared error against `target`; writes the output-layer gradient (channel 0 of a
// [N][H][W][32] bf16 tensor, the rest zero) and accumulates the loss sum.
__global__ __launch_bounds__(256) void k_loss_grad(const float *__restrict__ out, const float *__restrict__ target,
                                                   long long pixels, float scale, unsigned short *__restrict__ g,
                                                   float *__restrict__ loss) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    float d = 0.0f;
    if (p < pixels) {
        d = out[p] - target[p];
        __bf16 h[8] = {(__bf16)(2.0f * scale * d), (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f,
                       (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
        uint4 *dst = reinterpret_cast<uint4 *>(g + p * CH);
        dst[0] = *reinterpret_cast<const uint4 *>(h);
        dst[1] = dst[2] = dst[3] = make_uint4(0, 0, 0, 0);
    }
    float s = d * d;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, s * scale);
}

// Adam on the fp32 master parameters, then the bf16 copies: forward layout w[t][o][i] and the
// backward-data layout wt[t'][i][o] = w[8 - t][o][i] (taps mirrored, channels swapped).
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ grad, float *__restrict__ m,
                                              float *__restrict__ v, int count, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float gr = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_pack_weights(const float *__restrict__ w32, int cin, int cout,
                                                      unsigned short *__restrict__ wf, unsigned short *__restrict__ wb) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // over [9][32][32]
    if (i >= 9 * CH * CH) return;
    const int t = i / (CH * CH), o = (i / CH) % CH, in = i % CH;
    const float v = (o < cout && in < cin) ? w32[(t * cout + o) * cin + in] : 0.0f;
    const __bf16 h = (__bf16)v;
    const unsigned short u = *reinterpret_cast<const unsigned short *>(&h);
    wf[i] = u;
    if (wb) wb[((8 - t) * CH + in) * CH + o] = u;
}

// ---- inference ends: first layer (2 -> 32) and output layer (32 -> 1) on FP32 VALU -----------
// Inputs and weights are the same bf16 values the MFMA layers would see (products of bf16
// values are exact in FP32); only the summation order differs. One thread per output pixel of a
// 16 x 16 tile; the weights are wave-uniform, so they stay in scalar registers.

// x1[n][y][x][o] = relu(b[o] + sum_t w[t][o][0] * bf16(noisy(p + t)) + w[t][o][1] * bf16(sigma) [p + t inside])
// w0: fp32 [9][2][32] holding bf16-rounded values, then [32] sums over t of the sigma channel's
// weights: a pixel whose 9 taps are all inside the image takes the sigma term as one product.
__global__ __launch_bounds__(256) void k_conv_first(const float *__restrict__ noisy, int H, int W, int stride,
                                                    const float *__restrict__ sigma_map, const float *__restrict__ w0,
                                                    const float *__restrict__ b0, unsigned short *__restrict__ y) {
    __shared__ float tile[HALO][HALO + 1];
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE, n = blockIdx.z;
    // outside the image: a NaN marker, so the sigma term can be dropped there as well. Both loads
    // of a thread are issued before its stores.
    constexpr int kPer = (HALO * HALO + 255) / 256;
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        const int hy = i / HALO, hx = i % HALO;
        const int gy = y0 + hy - 1, gx = x0 + hx - 1;
        v[k] = (i < HALO * HALO && gy >= 0 && gy < H && gx >= 0 && gx < W)
                   ? (float)(__bf16)noisy[((long long)n * H + gy) * stride + gx]
                   : __builtin_nanf("");
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i < HALO * HALO) tile[i / HALO][i % HALO] = v[k];
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int gx = x0 + tx, gy = y0 + ty;
    const float sg = (float)(__bf16)sigma_map[n];
    const float *wsum = w0 + 9 * 2 * CH;
    float acc[CH];
    if (gx >= 1 && gx + 1 < W && gy >= 1 && gy + 1 < H) {
#pragma unroll
        for (int o = 0; o < CH; ++o) acc[o] = fmaf(wsum[o], sg, b0[o]);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float v = tile[ty + t / 3][tx + t % 3];
#pragma unroll
            for (int o = 0; o < CH; ++o) acc[o] = fmaf(w0[(t * 2 + 0) * CH + o], v, acc[o]);
        }
    } else {
#pragma unroll
        for (int o = 0; o < CH; ++o) acc[o] = b0[o];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float v = tile[ty + t / 3][tx + t % 3];
            const bool in = v == v;
            const float xv = in ? v : 0.0f, sv = in ? sg : 0.0f;
#pragma unroll
            for (int o = 0; o < CH; ++o) {
                acc[o] = fmaf(w0[(t * 2 + 0) * CH + o], xv, acc[o]);
                acc[o] = fmaf(w0[(t * 2 + 1) * CH + o], sv, acc[o]);
            }
        }
    }
    // The tile's 256 x 64 output bytes go through LDS so that consecutive lanes store consecutive
    // 16-byte chunks of a tile row (one pixel per thread would store at a 64-byte stride).
    __shared__ uint4 outs[TILE * TILE * 4];
    __bf16 h[CH];
#pragma unroll
    for (int o = 0; o < CH; ++o) h[o] = (__bf16)fmaxf(acc[o], 0.0f);
#pragma unroll
    for (int k = 0; k < 4; ++k) outs[threadIdx.x * 4 + k] = reinterpret_cast<const uint4 *>(h)[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = (int)threadIdx.x + 256 * k, pix = idx >> 2;
        const int oy = y0 + pix / TILE, ox = x0 + pix % TILE;
        if (oy < H && ox < W)
            reinterpret_cast<uint4 *>(y + (((long long)n * H + oy) * W + ox) * CH)[idx & 3] = outs[idx];
    }
}

// out[n][y][x] (row stride ostride) = noisy[n][y][x] + b5[0] + sum_t sum_i w5[t][i] * x[p + t][i]
// w5: [9][16] bf16 pairs (channels 2k, 2k+1), the layout of two adjacent activation channels, so
// each pair is one v_dot2_f32_bf16; noisy: fp32 [n][H][stride].
__global__ __launch_bounds__(256) void k_conv_last(const unsigned short *__restrict__ x, int H, int W,
                                                   const unsigned *__restrict__ w5, const float *__restrict__ b5,
                                                   const float *__restrict__ noisy, int stride,
                                                   float *__restrict__ out, int ostride) {
    __shared__ __attribute__((aligned(16))) char lds[HALO * HALO * PIX_BYTES];
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE, n = blockIdx.z;
    stage_halo(x, n, H, W, x0, y0, lds);
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int gx = x0 + tx, gy = y0 + ty;
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const char *px = lds + ((ty + t / 3) * HALO + tx + t % 3) * PIX_BYTES;
#pragma unroll
        for (int part = 0; part < 4; ++part) {
            const uint4 v = *reinterpret_cast<const uint4 *>(px + part * 16);
            const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                acc = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w5[t * (CH / 2) + part * 4 + k]),
                                                      __builtin_bit_cast(bf16x2, u[k]), acc, false);
        }
    }
    if (gx >= W || gy >= H) return;
    out[((long long)n * H + gy) * ostride + gx] = noisy[((long long)n * H + gy) * stride + gx] + (acc + b5[0]);
}

// The end layers' weights from the packed forward layout: fp32 copies (bf16-rounded) of the first
// layer's, the sums over taps of its sigma channel, and the last layer's as bf16 pairs.
__global__ __launch_bounds__(256) void k_end_weights(const unsigned short *__restrict__ wf0,
                                                     const unsigned short *__restrict__ wf5, float *__restrict__ w0,
                                                     unsigned *__restrict__ w5) {
    const int i = threadIdx.x + blockIdx.x * 256;
    if (i < 9 * 2 * CH) {  // w0[t][i][o] = wf0[t][o][i], i in {0, 1}
        const int t = i / (2 * CH), ii = (i / CH) % 2, o = i % CH;
        w0[i] = bf2f(wf0[(t * CH + o) * CH + ii]);
    }
    if (i < CH) {          // wsum[o] = sum over t of wf0[t][o][1]
        float sum = 0.0f;
        for (int t = 0; t < 9; ++t) sum += bf2f(wf0[(t * CH + i) * CH + 1]);
        w0[9 * 2 * CH + i] = sum;
    }
    if (i < 9 * CH / 2) {  // w5[t][k] = (wf5[t][0][2k], wf5[t][0][2k+1])
        const int t = i / (CH / 2), k = i % (CH / 2);
        w5[i] = (unsigned)wf5[(t * CH + 0) * CH + 2 * k] | ((unsigned)wf5[(t * CH + 0) * CH + 2 * k + 1] << 16);
    }
}

}  // namespace vszcnn

using namespace vszcnn;

namespace {
int launch_ok(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return vsz_fail(VSZ_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    return VSZ_OK;
}
}  // namespace

extern "C" {

int vsz_cnn_conv3x3(const void *x, const void *w, const float *bias, int n, int h, int wd, int mode, const void *aux,
                    const float *base, void *y, int out_stride, void *stream) {
    if (!x || !w || !y || n < 1 || h < 1 || wd < 1 || mode < 0 || mode > 2 || (mode != 2 && !bias) ||
        (mode == 2 && !aux) || (mode == 1 && (!base || out_stride < wd)) || n > 65535 || (h + TILE - 1) / TILE > 65535)
        return vsz_fail(VSZ_EINVAL, "vsz_cnn_conv3x3: invalid arguments");
    const long long tiles = (long long)((wd + TILE - 1) / TILE) * ((h + TILE - 1) / TILE) * n;
    if (tiles > (1LL << 30)) return vsz_fail(VSZ_EINVAL, "vsz_cnn_conv3x3: image too large");
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return vsz_fail(VSZ_EDEVICE, "vsz_cnn_conv3x3: no device");
    // six workgroups per CU (their 26 KB halos fit the LDS together): measured best of 2..8
    const int grid = (int)std::min<long long>(tiles, 6LL * std::max(cus, 1));
    hipLaunchKernelGGL(k_conv3x3, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const unsigned short *)x,
                       (const unsigned short *)w, bias, n, h, wd, (int)tiles, mode, (const unsigned short *)aux, base, y,
                       out_stride);
    return launch_ok("vsz_cnn_conv3x3");
}

size_t vsz_cnn_workspace_size(int n, int h, int wd) {
    if (n < 1 || h < 1 || wd < 1) return 0;
    const size_t px = (size_t)n * h * wd;
    return 2 * px * CH * 2 + 4096;   // two activation buffers + the end layers' weight tables (3 KB)
}

int vsz_cnn_run(const vsz_cnn_model *m, const float *noisy, int n, int h, int wd, int stride, const float *sigma_map,
                float *out, void *work, void *stream) {
    if (!m || !noisy || !sigma_map || !out || !work || n < 1 || n > 65535 || h < 1 || wd < 1 || stride < wd ||
        (h + TILE - 1) / TILE > 65535)
        return vsz_fail(VSZ_EINVAL, "vsz_cnn_run: invalid arguments");
    for (int l = 0; l < 6; ++l)
        if (!m->w[l] || !m->b[l]) return vsz_fail(VSZ_EINVAL, "vsz_cnn_run: missing layer %d", l);
    hipStream_t s = (hipStream_t)stream;
    const size_t px = (size_t)n * h * wd;
    char *wk = static_cast<char *>(work);
    void *act[2] = {wk, wk + px * CH * 2};
    float *w0 = reinterpret_cast<float *>(wk + 2 * px * CH * 2);   // [9][2][32], then [32] sigma-tap sums
    unsigned *w5 = reinterpret_cast<unsigned *>(w0 + 9 * 2 * CH + CH);   // [9][16] bf16 pairs
    hipLaunchKernelGGL(k_end_weights, dim3(3), dim3(256), 0, s, (const unsigned short *)m->w[0],
                       (const unsigned short *)m->w[5], w0, w5);
    const dim3 grid((wd + TILE - 1) / TILE, (h + TILE - 1) / TILE, n);
    hipLaunchKernelGGL(k_conv_first, grid, dim3(256), 0, s, noisy, h, wd, stride, sigma_map, w0, m->b[0],
                       (unsigned short *)act[0]);
    int rc = launch_ok("vsz_cnn_run: first layer");
    for (int l = 1; l < 5 && rc == VSZ_OK; ++l)
        rc = vsz_cnn_conv3x3(act[(l - 1) & 1], m->w[l], m->b[l], n, h, wd, 0, nullptr, nullptr, act[l & 1], 0, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_conv_last, grid, dim3(256), 0, s, (const unsigned short *)act[0], h, wd, w5, m->b[5], noisy,
                       stride, out, stride);
    return launch_ok("vsz_cnn_run: output layer");
}

int vsz_cnn_conv3x3_dw(const void *x, const void *dy, int n, int h, int wd, float *dw, void *stream) {
    if (!x || !dy || !dw || n < 1 || h < 1 || wd < 1) return vsz_fail(VSZ_EINVAL, "vsz_cnn_conv3x3_dw: invalid arguments");
    const long long total = (long long)n * ((wd + TILE - 1) / TILE) * ((h + TILE - 1) / TILE);
    const int tiles = 8;
    const unsigned grid = (unsigned)((total + tiles - 1) / tiles);
    hipLaunchKernelGGL(k_conv3x3_dw, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const unsigned short *)x,
                       (const unsigned short *)dy, n, h, wd, tiles, dw);
    return launch_ok("vsz_cnn_conv3x3_dw");
}

int vsz_cnn_bias_grad(const void *dy, long long pixels, float *db, void *stream) {
    if (!dy || !db || pixels < 1) return vsz_fail(VSZ_EINVAL, "vsz_cnn_bias_grad: invalid arguments");
    const unsigned grid = (unsigned)std::min<long long>((pixels + 2047) / 2048, 1024);
    hipLaunchKernelGGL(k_bias_grad, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const unsigned short *)dy, pixels, db);
    return launch_ok("vsz_cnn_bias_grad");
}

int vsz_cnn_pack_input(const float *noisy, int n, int h, int wd, int stride, const float *sigma_map, void *x0, float *base,
                       void *stream) {
    if (!noisy || !sigma_map || !x0 || n < 1 || n > 65535 || h < 1 || wd < 1 || stride < wd)
        return vsz_fail(VSZ_EINVAL, "vsz_cnn_pack_input: invalid arguments");
    const dim3 grid((unsigned)(((long long)h * wd + 255) / 256), n, 1);
    hipLaunchKernelGGL(k_pack_input, grid, dim3(256), 0, (hipStream_t)stream, noisy, h, wd, stride, sigma_map,
                       (unsigned short *)x0, base);
    return launch_ok("vsz_cnn_pack_input");
}

int vsz_cnn_loss_grad(const float *out, const float *target, long long pixels, void *grad, float *loss, void *stream) {
    if (!out || !target || !grad || !loss || pixels < 1) return vsz_fail(VSZ_EINVAL, "vsz_cnn_loss_grad: invalid arguments");
    hipLaunchKernelGGL(k_loss_grad, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, (hipStream_t)stream, out, target,
                       pixels, 1.0f / (float)pixels, (unsigned short *)grad, loss);
    return launch_ok("vsz_cnn_loss_grad");
}

int vsz_cnn_adam(float *p, const float *grad, float *m, float *v, int count, float lr, float b1, float b2, float eps,
                 int step, void *stream) {
    if (!p || !grad || !m || !v || count < 1 || step < 1) return vsz_fail(VSZ_EINVAL, "vsz_cnn_adam: invalid arguments");
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, p, grad, m, v, count, lr, b1,
                       b2, eps, c1, c2);
    return launch_ok("vsz_cnn_adam");
}

int vsz_cnn_pack_weights(const float *w32, int cin, int cout, void *wf, void *wb, void *stream) {
    if (!w32 || !wf || cin < 1 || cin > CH || cout < 1 || cout > CH)
        return vsz_fail(VSZ_EINVAL, "vsz_cnn_pack_weights: invalid arguments");
    hipLaunchKernelGGL(k_pack_weights, dim3((9 * CH * CH + 255) / 256), dim3(256), 0, (hipStream_t)stream, w32, cin, cout,
                       (unsigned short *)wf, (unsigned short *)wb);
    return launch_ok("vsz_cnn_pack_weights");
}

}  // extern "C"
