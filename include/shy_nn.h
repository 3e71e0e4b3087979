/*
 * shy_nn.h -- BF16 MFMA kernels of the discharge/momentum surrogate.
 *
 * No reference counterpart: the reference converges discharge and momentum
 * only by running the particle method (world.h:80-86).  These kernels train
 * and run a small convolutional network that predicts the converged fields
 * from a terrain, to warm-start the particle method.
 *
 * Tensors are device pointers the caller owns; activations are NHWC bf16;
 * weights are bf16 Wt[Cout][taps*Cin] (tap-major, tap = 3*(dy+1)+(dx+1));
 * biases and gradients FP32 (a convolution's bias 16-byte aligned: the
 * epilogue reads 4 at once).  All products use v_mfma_f32_16x16x32_bf16 with
 * FP32 accumulation.  Every call is asynchronous on `stream` (a hipStream_t;
 * NULL = the default stream) and returns 0, or SHY_NN_ERR with
 * shy_nn_last_error() describing the problem.
 */
#ifndef SHY_NN_H
#define SHY_NN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHY_NN_ERR (-10)

const char* shy_nn_last_error(void);

/* y = act(conv(x, w) + bias [+ res]); taps 9 (3x3, pad 1) or 1; Cin a
   multiple of 8; exactly one of y_bf16 / y_f32. */
int shy_nn_conv_fwd(const void* x, const void* w, const float* bias, const void* res,
                    void* y_bf16, float* y_f32, int N, int H, int W, int Cin, int Cout,
                    int taps, int relu, void* stream);
/* The same 3x3 convolution of the channel concatenation [up2(x_half), skip]
   -- a nearest x2 upsample of x_half [N][H/2][W/2][C1] in channels [0, C1),
   skip [N][H][W][C2] in [C1, C1+C2) -- read straight from the two tensors,
   with the same bits as shy_nn_up2 + shy_nn_channels + shy_nn_conv_fwd.
   Weights w[Cout][9*(C1+C2)]; H, W even; C1, C2, Cout multiples of 32. */
int shy_nn_conv_fwd_upcat(const void* x_half, const void* skip, const void* w, const float* bias,
                          const void* res, void* y_bf16, float* y_f32, int N, int H, int W,
                          int C1, int C2, int Cout, int relu, void* stream);
/* shy_nn_conv_fwd (3x3, Cout = 32, BF16 output y, which may be null: not
   stored) with a 1x1 convolution applied to the rounded output in the same
   pass: head_y [N][H][W][head_c] (FP32) = head_b + head_w [head_c][32] . y,
   head_c in {4, 8, 12, 16}, Cin 32 or 64, head_w and head_b 16-byte
   aligned.  A network's last convolution and its output head without the
   32-channel tensor round trip. */
int shy_nn_conv_fwd_head(const void* x, const void* w, const float* bias, const void* res,
                         void* y_bf16, const void* head_w, const float* head_b, float* head_y,
                         int N, int H, int W, int Cin, int head_c, int relu, void* stream);
/* shy_nn_conv_fwd (3x3) that also writes the 2x2 average pool of its
   output, pool_y [N][H/2][W/2][Cout] (8-byte aligned), with the bits of
   shy_nn_pool2 on y: an encoder stage's last convolution without reading its
   output back for the pool.  H, W even; Cin, Cout multiples of 32. */
int shy_nn_conv_fwd_pool(const void* x, const void* w, const float* bias, const void* res,
                         void* y_bf16, void* pool_y, int N, int H, int W, int Cin, int Cout,
                         int relu, void* stream);
/* The data gradient dx [N][H][W][Cin] (BF16) of a convolution from g
   [N][H][W][Cout] with the flipped weights wf of shy_nn_flip_weights, with
   the backward pass's next two element-wise steps in its epilogue: gres
   (optional, [N][H][W][Cin]) added to the rounded result, then the result
   zeroed where mask (optional, the forward input [N][H][W][Cin]) is not
   positive -- the same bits as shy_nn_conv_fwd, shy_nn_add and
   shy_nn_relu_bwd in sequence.  Cin, Cout multiples of 32. */
int shy_nn_conv_dgrad(const void* g, const void* wf, const void* gres, const void* mask,
                      void* dx_bf16, int N, int H, int W, int Cin, int Cout, int taps,
                      void* stream);
/* weights for the data gradient: wf[Cin][taps*Cout] = w[Cout][(taps-1-t)*Cin] */
int shy_nn_flip_weights(const void* w, void* wf, int Cin, int Cout, int taps, void* stream);
/* dw[Cout][taps*Cin] (FP32) = sum over pixels of g[p][co] x[p+tap][ci];
   part holds part_splits*Cout*taps*Cin floats of scratch */
int shy_nn_conv_wgrad(const void* x, const void* g, float* dw, float* part, int part_splits,
                      int N, int H, int W, int Cin, int Cout, int taps, void* stream);
/* The same weight gradient with no workspace argument and no allocation.
   No counterpart in the reference (nor in shy_nn_conv_wgrad's signature):
   for callers that cannot allocate device memory, such as code under HIP
   graph capture or a serving loop.  The split-K partials live in a
   file-scope device array of SHY_NN_NOALLOC_FLOATS floats (16 MiB) that
   every device the library runs on holds until the process ends; the call
   itself never calls hipMalloc or hipFree.  Equals shy_nn_conv_wgrad with
   part_splits = min(SHY_NN_NOALLOC_FLOATS / (Cout*taps*Cin),
   SHY_NN_NOALLOC_MAX_SPLITS) bit for bit.  Fails (SHY_NN_ERR) when
   Cout*taps*Cin exceeds SHY_NN_NOALLOC_FLOATS.  All calls on one device share
   the array: calls that overlap on one GPU (other streams or threads) race,
   so serialise them. */
#define SHY_NN_NOALLOC_FLOATS 4194304
#define SHY_NN_NOALLOC_MAX_SPLITS 1024
int shy_nn_conv_wgrad_noalloc(const void* x, const void* g, float* dw, int N, int H, int W,
                              int Cin, int Cout, int taps, void* stream);
/* db[c] = sum over pixels of g[p][c]; part holds part_rows*C floats */
int shy_nn_bias_grad(const void* g, float* db, float* part, int part_rows, int P, int C,
                     void* stream);
int shy_nn_relu_bwd(const void* y, const void* g, void* out, size_t n, void* stream);
/* 2x2 average pool (backward: x = dy, y = dx at full size) */
int shy_nn_pool2(const void* x, void* y, int N, int H, int W, int C, int backward, void* stream);
/* nearest x2 upsample of x [N][H][W][C] into channels [c_off, c_off+C) of
   y [N][2H][2W][Cy]; backward sums the 4 children back */
int shy_nn_up2(const void* x, void* y, int N, int H, int W, int C, int Cy, int c_off,
               int backward, void* stream);
/* dir 0: copy src [P][C] into channels [c_off, c_off+C) of dst [P][Cy];
   dir 1: the reverse */
int shy_nn_channels(const void* src, void* dst, size_t P, int C, int Cy, int c_off, int dir,
                    void* stream);
int shy_nn_add(const void* a, const void* b, void* y, size_t n, void* stream);
/* masked squared error: g = 2 (y-t) m inv_count (bf16), part[block] = sum of
   ((y-t) m)^2 over 256-element blocks (ceil(n/256) floats) */
int shy_nn_mse(const float* y, const float* t, const float* mask, void* g_bf16, float* part,
               size_t n, float inv_count, void* stream);
/* Adam on FP32 master weights, optional bf16 copy for the MFMA kernels */
int shy_nn_adam(float* w, const float* g, float* m, float* v, void* w_bf16, size_t n, float lr,
                float b1, float b2, float eps, int step, float weight_decay, void* stream);
/* shy_nn_adam on a convolution's weights w[Cout][taps*Cin] that also writes
   the rounded weights flipped for the data gradient (the layout of
   shy_nn_flip_weights) into wf_bf16, so no separate flip runs per step */
int shy_nn_adam_conv(float* w, const float* g, float* m, float* v, void* w_bf16, void* wf_bf16,
                     int Cin, int Cout, int taps, float lr, float b1, float b2, float eps, int step,
                     float weight_decay, void* stream);
int shy_nn_cast(const void* src, void* dst, size_t n, int to_bf16, void* stream);

#ifdef __cplusplus
}
#endif
#endif
