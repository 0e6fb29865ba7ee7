/*
 * amgt_policy.h -- C-ABI of the per-level precision / strength-threshold
 * policy network (part of libamgt_mbsr.so).
 *
 * The reference chooses a level's SpMV/SpGEMM precision with a fixed rule,
 * get_calculationPrecison (AmgT_HYPRE/src/seq_mv/csr_spgemm_device.c:1847-1862:
 * level 0 FP64, level 1 FP32, deeper levels FP16), and BoomerAMG uses one
 * strength threshold (par_amg.c, strong_threshold) on every level.  These
 * entry points evaluate and train a small network that predicts both per
 * level from the level's matrix statistics; the host decides how to use the
 * prediction.  They have no counterpart in the reference.
 *
 * Network: 12 features -> 64 tanh -> 64 tanh -> 4 outputs (logits of FP64,
 * FP32, FP16 and a strength threshold), FP32 throughout.  Parameter vector
 * (AMGT_POLICY_PARAMS floats, device memory, row-major W[out][in]):
 *   W1[64][12] b1[64] W2[64][64] b2[64] W3[4][64] b3[4].
 * Training loss, mean over the n samples of
 *   -log softmax(o[0..2])[cls] + lambda * (o[3] - target)^2.
 * All arrays are device pointers; `work` is caller-provided device memory of
 * at least amgt_policy_work_bytes(n) bytes (the entry points allocate
 * nothing).  Results are deterministic (fixed summation order, no atomics).
 */
#ifndef AMGT_POLICY_H
#define AMGT_POLICY_H

#include <stddef.h>
#include <stdint.h>

#define AMGT_POLICY_FEATURES 12
#define AMGT_POLICY_HIDDEN 64
#define AMGT_POLICY_OUTPUTS 4
#define AMGT_POLICY_PARAMS 5252

#ifdef __cplusplus
extern "C" {
#endif

/* Bytes of device work memory the calls below need for n samples. */
size_t amgt_policy_work_bytes(int64_t nsamples);
/* out[n][4] = network(features[n][12]); replaces the level -> precision rule
 * of get_calculationPrecison, csr_spgemm_device.c:1847-1862. */
int amgt_policy_forward(const float *params, const float *features, int64_t n, float *out,
                        void *work, size_t work_bytes, void *stream);
/* *loss = mean training loss, grad[AMGT_POLICY_PARAMS] = its gradient, for
 * labels cls[n] in {0 FP64, 1 FP32, 2 FP16} and threshold targets target[n].
 * A label outside 0..2 makes *loss NaN (that sample adds no gradient). */
int amgt_policy_loss_grad(const float *params, const float *features, const int32_t *cls,
                          const float *target, int64_t n, float lambda, float *grad, float *loss,
                          void *work, size_t work_bytes, void *stream);
/* Statistics of one level's operator that the policy's features are built
 * from (CSR on the device, int32 indices, diagonal found by column); stats is a
 * HOST array of 8 doubles: sum over rows of min(|a_ii| / sum_(j != i) |a_ij|,
 * 10) (10 for rows without off-diagonals), the number of rows where that ratio
 * is below 1, the number of positive off-diagonal entries, max |a_ij|, min
 * |a_ii| over rows with a_ii != 0 (+inf if none), the sum and the sum of
 * squares of log10 |a_ii| over those rows, and their number.  Synchronises
 * the stream. */
size_t amgt_level_stats_work_bytes(int32_t nrows);
int amgt_level_stats(int32_t n, const int32_t *A_i, const int32_t *A_j, const double *A_data,
                     double *stats, void *work, size_t work_bytes, void *stream);
/* One Adam step (bias-corrected, step >= 1) on params with moments m, v. */
int amgt_policy_adam(float *params, const float *grad, float *m, float *v, int64_t nparams,
                     int64_t step, float lr, float beta1, float beta2, float eps, void *stream);

#ifdef __cplusplus
}
#endif

#endif
