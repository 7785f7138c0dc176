/* rewl_mlp.h -- learned warm start for the REWL engine: a small MLP that
 * predicts ln g(E)/N from (E/N, 1/L, K, disorder sigma), trained and evaluated
 * on the GPU by hand-written gfx950 kernels (BF16 weights and activations on
 * the MFMA matrix cores, FP32 accumulation, FP32 master weights and Adam).
 *
 * The reference has no counterpart: every REWL_Walker starts from a flat
 * ln g = 0 and ln f = 1 (wang_landau.hpp:35-40, rewl_walker.hpp:118-139).
 * A caller predicts ln g for each window's bins with rewl_mlp_forward, writes
 * it with rewl_set_logdos and starts at a small ln f with rewl_set_window_lnf;
 * plain flat-initialised WL stays the fallback.
 *
 * Network: x (4) -> 128 -> 128 -> 128 -> 1, GELU (erf form) after each hidden
 * layer.  Parameters, in the order of the FP32 vector that get/set exchange:
 *   W1[128][4], b1[128], W2[128][128], b2[128], W3[128][128], b3[128],
 *   w4[128], b4                                     (REWL_MLP_NUM_PARAMS)
 * W[i][j] maps input j to output i.  Forward pass, per point:
 *   z1 = W1 x + b1 (FP32 on bf16(W1)), a1 = bf16(gelu(z1))
 *   z2 = bf16(W2) a1 + b2 (MFMA),      a2 = bf16(gelu(z2))
 *   z3 = bf16(W3) a2 + b3 (MFMA),      a3 = gelu(z3) (FP32)
 *   y  = bf16(w4) . a3 + b4
 * bf16() rounds to nearest even.  The loss is sum_i w_i (y_i - t_i)^2 /
 * sum_i w_i; the backward pass rounds dL/dz3 and dL/dz2 to bf16 for the MFMA
 * products and keeps every sum in FP32.
 *
 * Status codes are those of rewl_c.h.
 */
#ifndef REWL_MLP_H
#define REWL_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REWL_MLP_INPUTS 4
#define REWL_MLP_WIDTH 128
#define REWL_MLP_NUM_PARAMS 33793

typedef struct rewl_mlp rewl_mlp;

typedef struct rewl_mlp_train_params {
    int64_t steps;        /* Adam steps */
    int64_t batch;        /* points per step, drawn with replacement (Philox) */
    double lr;            /* peak learning rate */
    double lr_final;      /* cosine decay from lr to lr_final over the steps */
    double beta1, beta2, eps;
    uint64_t seed;        /* minibatch draw stream */
    int64_t warmup;       /* linear warm-up steps */
    int64_t reserved[3];
} rewl_mlp_train_params;

/* He-uniform initialisation from seed (splitmix64 stream), biases zero. */
int rewl_mlp_create(uint64_t seed, rewl_mlp** out);
int rewl_mlp_destroy(rewl_mlp* m);
int rewl_mlp_get_params(const rewl_mlp* m, float* theta, int64_t n);
/* Sets the FP32 master weights (and their bf16 copies); clears Adam state. */
int rewl_mlp_set_params(rewl_mlp* m, const float* theta, int64_t n);
/* y[i] = MLP(x[4 i .. 4 i + 3]) for i < n; host buffers. */
int rewl_mlp_forward(rewl_mlp* m, const float* x, int64_t n, float* y);
/* Full-batch gradient of the weighted MSE at the current weights (FP32,
 * REWL_MLP_NUM_PARAMS values) and the loss; w may be NULL (all ones). */
int rewl_mlp_gradient(rewl_mlp* m, const float* x, const float* t, const float* w, int64_t n, float* grad,
                      double* loss);
/* Adam on minibatches of (x, t, w); loss_history[s] = minibatch loss of step
 * s (may be NULL).  The data is copied to the device once per call. */
int rewl_mlp_train(rewl_mlp* m, const float* x, const float* t, const float* w, int64_t n,
                   const rewl_mlp_train_params* p, float* loss_history);
/* Device time of the last train/forward/gradient call, milliseconds. */
int rewl_mlp_last_ms(const rewl_mlp* m, double* ms);

#ifdef __cplusplus
}
#endif

#endif
