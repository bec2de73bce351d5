// Kernels of the learned preconditioner (a 2-D U-Net): bfloat16 activations
// and weights, float32 accumulation on the matrix cores (unet.hip).
//
// Tensors are NHWC.  A "view" is (pointer to channel 0 of pixel 0, channel
// stride): a convolution reads or writes a slice of channels of a wider
// buffer, which is how the decoder's concatenations are formed without a
// copy.  bf16 values are stored as uint16_t bit patterns (round to nearest,
// ties to even).
#pragma once

#include "core.h"

namespace swp {
namespace nn {

// out[p, co] = act(bias[co] + sum_{tap, ci} w[co, tap, ci] * in[p + off(tap), ci]),
// same-size (zero padding ks / 2), ks in {1, 3}; out bf16 or (out_f32) float.
// cin, in_stride, out_stride multiples of 8; pointers 16-byte aligned.
void conv_fwd(const uint16_t *in, int N, int H, int W, int cin, int in_stride, const uint16_t *w, int cout,
              int ks, const float *bias, bool relu, void *out, int out_stride, bool out_f32);

// dw[co, tap, ci] = sum_p dout[p, co] * in[p + off(tap), ci] (float32,
// overwritten); db[co] = sum_p dout[p, co] if db.  partials: scratch of
// conv_wgrad_partials(...) floats.
int64_t conv_wgrad_partials(int N, int H, int W, int cin, int cout, int ks);
void conv_wgrad(const uint16_t *in, int N, int H, int W, int cin, int in_stride, const uint16_t *dout, int cout,
                int dout_stride, int ks, float *dw, float *db, float *partials);

// wt[ci, tap, co] = w[co, ks*ks - 1 - tap, ci]: the weights of the data
// gradient, which is then conv_fwd(dout, wt) with cin and cout swapped.
void conv_flip(const uint16_t *w, int cout, int cin, int ks, uint16_t *wt);

// d[p, c] = act[p, c] > 0 ? d[p, c] : 0 over P pixels, C channels.
void relu_bwd(uint16_t *d, int d_stride, const uint16_t *act, int act_stride, int64_t P, int C);
// 2 x 2 average pool (H, W even) and its adjoint (accumulated into din).
void pool2_fwd(const uint16_t *in, int in_stride, int N, int H, int W, int C, uint16_t *out, int out_stride);
void pool2_bwd_acc(const uint16_t *dout, int dout_stride, int N, int H, int W, int C, uint16_t *din,
                   int din_stride);
// nearest 2x upsample of an (N, H, W, C) view and its adjoint (overwrites din).
void up2_fwd(const uint16_t *in, int in_stride, int N, int H, int W, int C, uint16_t *out, int out_stride);
void up2_bwd(const uint16_t *dout, int dout_stride, int N, int H, int W, int C, uint16_t *din, int din_stride);

// x[p, c] = bf16(scale[c] * src[c][p]) for c < nsrc, 0 up to stride: the
// network input from nsrc float planes of P = N*H*W cells.
void pack_input(const float *const *src, const float *scale, int nsrc, int64_t P, uint16_t *x, int stride);
// Mean squared error of out[p, 0] (float, channel stride out_stride) against
// target[p] (* mask[p] if mask): loss[0] = sum (o - t)^2 m / count (float64,
// fixed order), dout[p, 0] = bf16(2 (o - t) m / count), other channels 0.
void mse(const float *out, int out_stride, const float *target, const float *mask, int64_t P, double count,
         uint16_t *dout, int dout_stride, double *loss, double *partials);
// Adam (Kingma & Ba 2015) on float32 master weights, bf16 copy refreshed.
void adam(float *w, const float *g, float *m, float *v, uint16_t *wbf, int64_t n, float lr, float b1, float b2,
          float eps, int step);
void to_bf16(const float *src, uint16_t *dst, int64_t n);

}  // namespace nn
}  // namespace swp
