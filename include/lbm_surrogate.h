/*
 * lbm_surrogate.h -- learned warm start for steady lattice-Boltzmann
 * channel flows (part of liblbm_hip.so).
 *
 * No reference counterpart: SFCMM/LBM always starts from rest
 * (src/lbm/solver.cpp:267-295).  A convolutional encoder-decoder (U-Net,
 * 23 3x3 convolutions on 6 resolution levels, about 1.9M parameters) maps
 * per-cell geometry features to the steady (u_x, u_y, rho) field of a 2D
 * channel; its prediction initialises the solver (lbm_init_fields), which
 * then converges as usual.  Every kernel is hand-written HIP for gfx950:
 * the convolutions run on bf16 MFMA (32x32x16) with FP32 accumulation,
 * training keeps FP32 master weights and Adam moments.
 *
 * Layouts (host arrays): features x[n][H][W][in_channels] FP32, targets
 * t[n][H][W][3] FP32, fluid masks m[n][H][W] (1 = cell enters the loss).
 * H and W must be multiples of 32, in_channels <= 16.  Loss: mean over
 * fluid cells and the 3 outputs of the squared error.  Status codes as in
 * lbm_hip.h.
 */
#ifndef LBM_SURROGATE_H
#define LBM_SURROGATE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_sur lbm_sur;

/* batch: samples per training step and per forward call (device buffers
 * are sized for it); seed: weight initialisation (He normal). */
int lbm_sur_create(int32_t batch, int32_t height, int32_t width, int32_t in_channels, uint64_t seed, lbm_sur **out);

/* number of FP32 parameters (weights [co][3][3][ci] then bias [co], per
 * layer in network order) */
int lbm_sur_param_count(const lbm_sur *s, int64_t *n);
int lbm_sur_get_params(lbm_sur *s, float *params);
int lbm_sur_set_params(lbm_sur *s, const float *params);

/* prediction for n <= batch samples: y[n][H][W][3] */
int lbm_sur_forward(lbm_sur *s, const float *x, int32_t n, float *y);

/* loss of n <= batch samples and, when grad is not NULL, its gradient with
 * respect to every parameter (same order as lbm_sur_get_params). */
int lbm_sur_loss_grad(lbm_sur *s, const float *x, const float *t, const uint8_t *mask, int32_t n, double *loss,
                      float *grad);

/* device-resident training set of n samples */
int lbm_sur_set_dataset(lbm_sur *s, const float *x, const float *t, const uint8_t *mask, int64_t n);

/* nsteps Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr;
 * step k trains on samples idx[k*batch .. k*batch+batch-1] of the dataset.
 * losses[nsteps] (may be NULL) receives the loss of every step. */
int lbm_sur_train(lbm_sur *s, const int64_t *idx, int32_t nsteps, float lr, double *losses);

/* prediction for dataset samples i0 .. i0+n-1 (n <= batch): y[n][H][W][3] */
int lbm_sur_predict_dataset(lbm_sur *s, int64_t i0, int32_t n, float *y);

const char *lbm_sur_last_error(const lbm_sur *s);
int lbm_sur_destroy(lbm_sur *s);

#ifdef __cplusplus
}
#endif
#endif
