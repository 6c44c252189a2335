/*
 * glide_unet.h - C ABI of the warm-start network kernels (MI355X, gfx950).
 *
 * The reference has no learned component: these entry points serve a small
 * convolutional U-Net whose predicted (u, v) seeds the FASCD V-cycles of a
 * forward step (the solve of glide/physics.py:141-209), with plain FASCD as
 * the fallback and the arbiter of the residual.
 *
 * Tensors: activations BF16 (raw uint16 bits, round-to-nearest-even), NHWC,
 * every channel count a positive multiple of 32; 3x3 weights [9][Cout][Cin]
 * with tap = 3*(dy+1) + (dx+1); FP32 biases, gradients and master weights.
 * Conventions as in glide_hip.h: device pointers, caller's stream (void*),
 * nothing allocated (scratch is caller-provided), 0 / GLIDE_E_* / hipError_t
 * return codes, glide_last_error() for the message.
 */
#ifndef GLIDE_UNET_H
#define GLIDE_UNET_H

#include <stdint.h>

#include "glide_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out[p][co] = act(bias[co] + sum_{t,ci} W[t][co][ci] x[p+off_t][ci]), zero
 * padding, x = concat(xa (CA channels), xb (CB channels, may be 0 / NULL));
 * act = ReLU when relu != 0.  BF16 in and out, FP32 accumulation (MFMA). */
int glide_unet_conv3x3(const uint16_t *xa, int CA, const uint16_t *xb, int CB, int N, int H, int W,
                       const uint16_t *weights, int Cout, const float *bias, int relu, uint16_t *out,
                       void *stream);

/* Data gradient of glide_unet_conv3x3: from dz (BF16, Cout channels) and the
 * flipped weights of glide_unet_prep_weights, writes FP32 gradients of the
 * two inputs: ga (CA channels) and gb (CB channels); acc_* != 0 adds. */
int glide_unet_conv3x3_dgrad(const uint16_t *dz, int Cout, int N, int H, int W, const uint16_t *wflip,
                             int CA, float *ga, int acc_a, int CB, float *gb, int acc_b, void *stream);

/* Weight gradient dw[t][co][ci] = sum_p dz[p][co] x[p+off_t][ci] (FP32),
 * split over pixel chunks into partial (glide_unet_wgrad_chunks(N, H, W) x 9
 * x Cout x (CA+CB) floats) and reduced in chunk order (deterministic). */
long glide_unet_wgrad_chunks(int N, int H, int W);
int glide_unet_conv3x3_wgrad(const uint16_t *dz, int Cout, const uint16_t *xa, int CA, const uint16_t *xb,
                             int CB, int N, int H, int W, float *partial, float *dw, void *stream);

/* C[m][n] = sum_k A[m][k] B[n][k] (BF16 in, FP32 out; M, N multiples of 16,
 * K of 32): the MFMA fragment layout the convolutions use, exposed for tests. */
int glide_unet_gemm_bf16(const uint16_t *A, const uint16_t *B, float *C, int M, int N, int K, void *stream);

/* BF16 copy of FP32 master weights [9][Cout][Cin] and (wflip != NULL) the
 * data-gradient weights wflip[8-t][ci][co] = W[t][co][ci]. */
int glide_unet_prep_weights(const float *w_master, int Cout, int Cin, uint16_t *wb, uint16_t *wflip,
                            void *stream);

/* dz = bf16(dy * (y > 0)): ReLU backward from the post-activation output. */
int glide_unet_relu_bwd(const float *dy, const uint16_t *y, long n, uint16_t *dz, void *stream);

/* 2x2 average pooling (H, W even) and its backward (acc != 0 adds). */
int glide_unet_pool2(const uint16_t *x, int N, int H, int W, int C, uint16_t *y, void *stream);
int glide_unet_pool2_bwd(const float *dy, int N, int H, int W, int C, float *dx, int accumulate,
                         void *stream);

/* Nearest x2 upsampling of an N x H x W x C tensor and its backward. */
int glide_unet_up2(const uint16_t *x, int N, int H, int W, int C, uint16_t *y, void *stream);
int glide_unet_up2_bwd(const float *dy, int N, int H, int W, int C, float *dx, void *stream);

/* db[c] = sum_p dz[p][c] via glide_unet_bias_chunks(P) x C partials. */
long glide_unet_bias_chunks(long P);
int glide_unet_bias_grad(const uint16_t *dz, long P, int C, float *partial, float *db, void *stream);

/* 1x1 output head with NO (1..4) outputs, FP32:
 * out[p][o] = bh[o] + sum_c wh[o][c] x[p][c]; backward: dx (FP32, P x C) and
 * [dwh (NO x C) | dbh (NO)] through glide_unet_bias_chunks(P) x (NO C + NO)
 * partials. */
int glide_unet_head(const uint16_t *x, long P, int C, int NO, const float *wh, const float *bh, float *out,
                    void *stream);
int glide_unet_head_bwd(const float *dout, const uint16_t *x, long P, int C, int NO, const float *wh,
                        float *dx, float *partial, float *dwh_dbh, void *stream);

/* Masked MSE: grad = 2 m (pred - target) inv_count (mask per pixel of 2
 * values, may be NULL); per-block loss partials (glide_unet_mse_blocks(n)). */
long glide_unet_mse_blocks(long n);
int glide_unet_mse(const float *pred, const float *target, const float *mask, long n, float inv_count,
                   float *grad, float *partial, void *stream);

/* Adam step on FP32 master weights (step >= 1; decoupled weight decay added
 * to the gradient). */
int glide_unet_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
                    float eps, int step, float weight_decay, void *stream);

/* Network input of one domain (ny x nx cells, FP64 fields) as BF16 NHWC with
 * 32 channels: H/1000, tanh((bed + 0.917 H)/200), 100 |grad s|, log(beta) + 2,
 * B/24 - 1, 100 ds/dx, 100 ds/dy, then zeros; s = max(bed, -0.917 H) + H
 * (glide/physics.py:223-229). */
int glide_unet_features(const double *H, const double *bed, const double *beta, const double *B, int ny,
                        int nx, double dx, uint16_t *x, void *stream);

/* Training targets, 3 FP32 values per cell: cell-centred u and v times
 * scale, and the step's thickness change (H - H_prev) times hscale. */
int glide_unet_targets(const double *u, const double *v, const double *H, const double *H_prev, int ny,
                       int nx, double scale, double hscale, float *tgt, void *stream);

/* Solver state from a prediction (3 per cell): MAC faces from the
 * cell-centred velocities / scale (interior faces average their two cells,
 * boundary faces copy their one cell) and H = max(H_prev + dH / hscale,
 * gamma). */
int glide_unet_apply_prediction(const float *pred, int ny, int nx, double scale, double hscale, double *u,
                                double *v, double *H, const double *H_prev, const double *gamma,
                                void *stream);

#ifdef __cplusplus
}
#endif

#endif
