/*
 * mpd_nn.h -- C ABI of the learned warm start's gfx950 kernels
 * (mpd-plasma-dynamics_amd/libmpd_nn.so): a BF16 convolutional
 * encoder-decoder with FP32 accumulation on the matrix cores, and the kernels
 * that train it on the GPU.  No counterpart in the reference, which always
 * starts from its expression or VTK initial condition
 * (src/main.cpp:117-148); the network's prediction is an initial state that
 * the LF time loop then relaxes to its own steady state.
 *
 * Conventions: every pointer is a device pointer; tensors are NHWC (batch,
 * z rows, r columns, channels), contiguous; bf16 values are passed as
 * uint16_t bit patterns; `stream` is a hipStream_t (NULL: the null stream).
 * Calls enqueue work and return 0, or a negative MPD_NN_E* code with a
 * message in mpd_nn_last_error().  Conv weights are [Cout][3][3][Cin].
 */
#ifndef MPD_NN_H
#define MPD_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPD_NN_OK 0
#define MPD_NN_E_DEVICE -2
#define MPD_NN_E_ARG -3

const char* mpd_nn_last_error(void);

/* Device memory of the caller's tensors (hipMalloc / hipFree / synchronous
 * copies and memset on this library's HIP runtime). */
int mpd_nn_malloc(void** p, long long bytes);
int mpd_nn_free(void* p);
int mpd_nn_memcpy_h2d(void* dst, const void* src, long long bytes);
int mpd_nn_memcpy_d2h(void* dst, const void* src, long long bytes);
int mpd_nn_memset(void* p, int value, long long bytes);
int mpd_nn_synchronize(void);

/* 3x3 convolution, stride 1, zero padding 1: z = conv(x, w) + bias (bias may
 * be NULL).  x [B][H][W][Cin] bf16, z [B][H][W][Cout] FP32. */
int mpd_nn_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const float* bias, float* z, int B, int H, int W,
                       int Cin, int Cout, void* stream);
/* The same, also writing a = bf16(silu(z)) [B][H][W][Cout] (as
 * mpd_nn_silu_fwd on z, bit for bit) from the GEMM's registers. */
int mpd_nn_conv3x3_fwd_silu(const uint16_t* x, const uint16_t* w, const float* bias, float* z, uint16_t* a, int B,
                            int H, int W, int Cin, int Cout, void* stream);
/* dx = dL/dx [B][H][W][Cin] FP32 from dz = dL/dz [B][H][W][Cout] bf16. */
int mpd_nn_conv3x3_bwd_data(const uint16_t* dz, const uint16_t* w, float* dx, int B, int H, int W, int Cin, int Cout,
                            void* stream);
/* The data gradient with the layer below's SiLU backward fused (channel
 * counts multiples of 8, Cin not an odd multiple of 96): for dx as above and
 * z_in that layer's pre-activation [B][H][W][Cin], writes d = dx silu'(z_in)
 * as bf16 (dz_in_bf16; bit for bit mpd_nn_silu_bwd_bias's) and db_in[c] =
 * the sum of d over the pixels (FP32, per pixel tile then over the tiles in
 * a fixed order: deterministic, though not mpd_nn_silu_bwd_bias's order); no
 * FP32 dx is written.  `work` holds mpd_nn_bwd_data_silu_work_floats() floats. */
long long mpd_nn_bwd_data_silu_work_floats(int B, int H, int W, int Cin, int Cout);
int mpd_nn_conv3x3_bwd_data_silu(const uint16_t* dz, const uint16_t* w, const float* z_in, uint16_t* dz_in_bf16,
                                 float* db_in, float* work, int B, int H, int W, int Cin, int Cout, void* stream);
/* dw = dL/dw [Cout][3][3][Cin] FP32; `work` holds mpd_nn_wgrad_work_floats()
 * floats of split partial sums (reduced in a fixed order). */
long long mpd_nn_wgrad_work_floats(int B, int H, int W, int Cin, int Cout);
int mpd_nn_conv3x3_bwd_weight(const uint16_t* x, const uint16_t* dz, float* dw, float* work, int B, int H, int W,
                              int Cin, int Cout, void* stream);
/* db[c] = sum over the P pixels of dz[p][c] (FP32), summed in a fixed order
 * over pixel slices; `work` holds mpd_nn_bias_work_floats(P, C) floats. */
long long mpd_nn_bias_work_floats(int P, int C);
int mpd_nn_bias_grad(const float* dz, float* db, float* work, int P, int C, void* stream);

/* a = bf16(silu(z)), silu(z) = z / (1 + exp(-z)); backward: dz = da silu'(z)
 * written as bf16 and (if dz_f32 != NULL) FP32. */
int mpd_nn_silu_fwd(const float* z, uint16_t* a, long long n, void* stream);
int mpd_nn_silu_bwd(const float* z, const float* da, uint16_t* dz_bf16, float* dz_f32, long long n, void* stream);
int mpd_nn_to_bf16(const float* x, uint16_t* y, long long n, void* stream);
/* dz = da silu'(z) ([P][C]; dz = da when z is NULL, a layer without
 * activation) written as bf16, with db = the bias gradient, as
 * mpd_nn_bias_grad on the FP32 dz bit for bit, in one pass over z and da;
 * `work` holds mpd_nn_bias_work_floats(P, C) floats. */
int mpd_nn_silu_bwd_bias(const float* z, const float* da, uint16_t* dz_bf16, float* db, float* work, int P, int C,
                         void* stream);

/* 2x2 average pooling (H, W even) and nearest-neighbour x2 upsampling
 * (H, W: the input sizes); backward passes on FP32 gradients. */
int mpd_nn_avgpool2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream);
int mpd_nn_avgpool2_bwd(const float* dy, float* dx, int B, int H, int W, int C, void* stream);
int mpd_nn_upsample2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream);
int mpd_nn_upsample2_bwd(const float* dy, float* dx, int B, int H, int W, int C, void* stream);

/* Channel concatenation y = [a, b] over P pixels; backward splits dy into da
 * and db, adding to them where acc_a / acc_b is nonzero. */
int mpd_nn_concat_fwd(const uint16_t* a, int Ca, const uint16_t* b, int Cb, uint16_t* y, long long P, void* stream);
int mpd_nn_concat_bwd(const float* dy, float* da, int Ca, int acc_a, float* db, int Cb, int acc_b, long long P,
                      void* stream);
int mpd_nn_add(float* y, const float* x, long long n, void* stream);

/* The decoder's skip joins in one pass each (H, W: the coarse sizes; channel
 * counts multiples of 8 forward, of 4 backward; bit for bit the separate
 * passes): y [B][2H][2W][Cu + Cs] = [upsample2(x) (Cu), skip (Cs)];
 * backward dx [B][H][W][Cu] = upsample2 backward of dy[..., :Cu] and dskip
 * [B][2H][2W][Cs] = dy[..., Cu:] (overwritten). */
int mpd_nn_upcat_fwd(const uint16_t* x, int Cu, const uint16_t* skip, int Cs, uint16_t* y, int B, int H, int W,
                     void* stream);
int mpd_nn_upcat_bwd(const float* dy, float* dx, int Cu, float* dskip, int Cs, int B, int H, int W, void* stream);
/* dx = avgpool2 backward of dy + skip (H, W: the fine sizes, even; C a
 * multiple of 4), as mpd_nn_avgpool2_bwd then mpd_nn_add. */
int mpd_nn_pool_bwd_add(const float* dy, const float* skip, float* dx, int B, int H, int W, int C, void* stream);
/* mpd_nn_pool_bwd_add followed by mpd_nn_silu_bwd_bias (P = B H W, z the
 * pre-activation [B][H][W][C]) in one pass, with no FP32 dx written: dz_bf16
 * and db bit for bit those of the two calls.  H, W the fine sizes, even;
 * B H W C below 2^31; `work` holds mpd_nn_bias_work_floats(B H W, C) floats. */
int mpd_nn_pool_silu_bwd_bias(const float* dy, const float* skip, const float* z, uint16_t* dz_bf16, float* db,
                              float* work, int B, int H, int W, int C, void* stream);

/* Mean squared error of y against t over n values: dy = 2 (y - t) / n (dy
 * may be NULL) and mpd_nn_mse_blocks(n) partial sums of (y - t)^2 in part. */
int mpd_nn_mse_blocks(long long n);
int mpd_nn_mse(const float* y, const float* t, float* dy, float* part, long long n, void* stream);

/* Adam step (bias-corrected, step >= 1) on FP32 master weights w with
 * moments m, v, using the gradient g scaled by gscale (global-norm clipping:
 * 1 leaves it unchanged); w_bf16 (may be NULL) receives the bf16 copy. */
int mpd_nn_adam(float* w, uint16_t* w_bf16, const float* g, float* m, float* v, long long n, float lr, float b1,
                float b2, float eps, int step, float gscale, void* stream);
/* The same step over `count` tensors (host arrays of device pointers and
 * sizes; w_bf16 or its entries may be NULL) in as few launches as possible,
 * bit for bit as mpd_nn_adam on each. */
int mpd_nn_adam_multi(int count, float* const* w, uint16_t* const* w_bf16, const float* const* g, float* const* m,
                      float* const* v, const long long* n, float lr, float b1, float b2, float eps, int step,
                      float gscale, void* stream);

/* mpd_nn_sumsq_blocks(n) partial sums of x[i]^2 in part (the caller adds
 * them in order: a deterministic squared norm). */
int mpd_nn_sumsq_blocks(long long n);
int mpd_nn_sumsq(const float* x, float* part, long long n, void* stream);

/* dst[j] = src[idx[j]] for rows of row_len floats (mini-batch gather). */
int mpd_nn_gather_rows(const float* src, const int* idx, float* dst, int nrows, long long row_len, void* stream);

#ifdef __cplusplus
}
#endif
#endif
