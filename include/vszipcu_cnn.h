/* vszipcu_cnn.h — learned basic pass of libvszipcu_amd.so: a residual CNN of six 3x3
 * convolutions (32 channels, ReLU between them; input = noisy plane + noise-level map) whose
 * output replaces the hard-threshold basic estimate as the Wiener pass's reference.
 *
 * No reference counterpart: the reference only has the hard-threshold pass (src/bm3d.cu:416-478,
 * the `ref` argument of src/plugin.zig:84-90 is where a learned estimate enters). The kernels
 * are BF16 (weights and activations) with FP32 accumulation; the training kernels are here too.
 *
 * Layout: activations [n][h][w][32] bf16 (uint16 storage); weights [9][32 out][32 in] bf16 with
 * tap = (dy+1)*3 + (dx+1) (a correlation, like torch.nn.functional.conv2d); biases fp32 [32].
 * Layer 1's two inputs and layer 6's single output are zero-padded to 32 channels.
 * Every call is asynchronous on `stream` (a hipStream_t, NULL = null stream).
 */
#ifndef VSZIPCU_CNN_H
#define VSZIPCU_CNN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mode 0: y = relu(conv(x) + bias), bf16 activations;
 * mode 1: y[n][r][c] (fp32, row stride out_stride) = base[n][r][c] + conv(x)[out 0] + bias[0];
 * mode 2: y = conv(x) * (aux > 0), bf16 (the data gradient through the previous ReLU; w is then
 *         the mirrored, channel-swapped copy from vsz_cnn_pack_weights). */
int vsz_cnn_conv3x3(const void *x, const void *w, const float *bias, int n, int h, int w_, int mode, const void *aux,
                    const float *base, void *y, int out_stride, void *stream);
/* dw[t][o][i] += sum_p dy[p][o] * x[p + tap t][i] (fp32, accumulated with atomics). */
int vsz_cnn_conv3x3_dw(const void *x, const void *dy, int n, int h, int w_, float *dw, void *stream);
/* db[o] += sum_p dy[p][o]. */
int vsz_cnn_bias_grad(const void *dy, long long pixels, float *db, void *stream);
/* x0 = (noisy, sigma_map[image], 0...) as bf16 [n][h][w][32]; base (optional) = noisy, compact fp32. */
int vsz_cnn_pack_input(const float *noisy, int n, int h, int w_, int stride, const float *sigma_map, void *x0,
                       float *base, void *stream);
/* Mean squared error: grad channel 0 = 2 (out - target) / pixels, *loss += sum (out - target)^2 / pixels. */
int vsz_cnn_loss_grad(const float *out, const float *target, long long pixels, void *grad, float *loss, void *stream);
/* One Adam step on fp32 parameters (bias-corrected with `step` >= 1). */
int vsz_cnn_adam(float *p, const float *grad, float *m, float *v, int count, float lr, float b1, float b2, float eps,
                 int step, void *stream);
/* fp32 [9][cout][cin] -> bf16 forward layout wf [9][32][32] and (optional) backward-data layout wb. */
int vsz_cnn_pack_weights(const float *w32, int cin, int cout, void *wf, void *wb, void *stream);

/* Inference. */
typedef struct vsz_cnn_model {
    const void *w[6];   /* packed bf16 [9][32][32] per layer (device) */
    const float *b[6];  /* fp32 [32] per layer (device) */
} vsz_cnn_model;

size_t vsz_cnn_workspace_size(int n, int h, int w_);
/* out[n][r][c] (row stride `stride`) = learned basic estimate of noisy[n] (same layout);
 * sigma_map: n device floats (the noise level sigma/255 of each image). */
int vsz_cnn_run(const vsz_cnn_model *m, const float *noisy, int n, int h, int w_, int stride, const float *sigma_map,
                float *out, void *work, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* VSZIPCU_CNN_H */
