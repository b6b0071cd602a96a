/*
 * fastem_setnet.h -- the learned initializer's set network (libfastem.so).
 *
 * No counterpart in the fastem reference: the reference initializes EM with
 * the Jain / k-means / k-means++ / file initializers of
 * EM2/EMInitializer.cpp (see fem_kmeans_init in fastem_abi.h for the device
 * k-means family).  This network maps a set of normalised points to K
 * starting means and log-variances; its forward, hand-written backward and
 * Adam update run as HIP kernels in FP32 with fixed-order reductions.
 *
 * Parameters are one flat FP32 vector, in this order:
 *   W1[d][64] b1[64] W2[64][64] b2[64] S[K][64] Wg[64][64]
 *   Wm[64][d] bm[d] Wv[64][d] bv[d] Wp[64][64] Wq[64][64]
 * Training arithmetic is FP32 (fem_setnet_forward, fem_setnet_step); the
 * inference forward fem_setnet_forward_bf16 runs the per-point layers, the
 * point head's projections, scores and weighted sum on BF16 MFMA
 * (v_mfma_f32_16x16x32_bf16, operands rounded to bf16, FP32 accumulation).
 * Network, per set of n points x_i (width 64):
 *   h1 = gelu(x W1 + b1), h2 = gelu(h1 W2 + b2), g = mean_i h2_i,
 *   q_k = S_k + g Wg, a_ik = softmax_i(q_k . h2_i / 8), z_k = sum_i a_ik h2_i,
 *   mu_k = z_k Wm + bm, lv_k = z_k Wv + bv,
 * and the point head, a learned mean-shift step over the set:
 *   m_i = sum_j softmax_j((h2_i Wp) . (h2_j Wq) / 8) x_j.
 * Training loss per set against a fitted mixture (M_j, V_j, weights p_j):
 *   sum_j p_j |mu_k(j) - M_j|^2 + (1/K) sum_k min_j |mu_k - M_j|^2
 *   + lam sum_j p_j |lv_k(j) - V_j|^2 / d,   k(j) = argmin_k |mu_k - M_j|^2
 *   + gam (1/n) sum_i |m_i - T_i|^2,          T_i the mean of point i's cluster.
 * Supported: d in {8, 16, 32, 64}, 1 <= K <= 64, targets J <= 4K.
 */
#ifndef FASTEM_SETNET_H
#define FASTEM_SETNET_H
#include <stdint.h>

#include "fastem_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_setnet fem_setnet;

/* Creates a network with Gaussian-initialised weights (seeded mt19937_64),
 * zero biases and zero Adam state on HIP device `device`. */
int fem_setnet_create(fem_setnet **net, int d, int K, uint64_t seed, int device);
int fem_setnet_destroy(fem_setnet *net);
const char *fem_setnet_last_error(const fem_setnet *net);
int fem_setnet_num_params(const fem_setnet *net);
int fem_setnet_shape(const fem_setnet *net, int *d, int *K);
int fem_setnet_get_params(fem_setnet *net, float *theta);
/* Loads parameters and resets the Adam state. */
int fem_setnet_set_params(fem_setnet *net, const float *theta);

/* Forward pass over B sets of n normalised points (X: B*n*d FP32, host);
 * writes means and log-variances (B*K*d each, host) and, if m is not NULL,
 * the point head's outputs (B*n*d). */
int fem_setnet_forward(fem_setnet *net, int B, int n, const float *X, float *mu, float *lv, float *m);

/* The same forward for inference with the products on BF16 MFMA (FP32
 * accumulation; pooling, seed-query attention, heads and the softmax in
 * FP32).  Outputs as fem_setnet_forward. */
int fem_setnet_forward_bf16(fem_setnet *net, int B, int n, const float *X, float *mu, float *lv, float *m);

/* One training step over B sets: forward, loss against the targets
 * (M, V: B*J*d; p: B*J; point targets T: B*n*d, or NULL to leave the point
 * head out of the loss), backward.  grad (host, may be NULL) receives the
 * gradient of the loss summed over the batch; *loss receives that sum.  With
 * lr > 0 one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) is applied with the
 * gradient of the batch-mean loss. */
int fem_setnet_step(fem_setnet *net, int B, int n, const float *X, int J, const float *M, const float *V,
                    const float *p, const float *T, float lam, float gam, float lr, float *grad, float *loss);

/* ---- Weight files ----------------------------------------------------------
 * Raw little-endian file: 8 bytes "FEMSNET1", int32 d, int32 K, int32 count,
 * then count FP32 parameters in the order above.  fem_setnet_load creates the
 * network on `device` (errors through fem_last_error(NULL)). */
int fem_setnet_save(fem_setnet *net, const char *path);
int fem_setnet_load(fem_setnet **net, const char *path, int device);

/* ---- The learned initializer -------------------------------------------------
 * For the dataset held by a fem_ctx of the network's d and K.  One Rand()
 * stream, rnd(rnd_state) (or std::mt19937_64(seed) through generate_canonical,
 * Util/Rand.cpp:7-14), drives every draw, in this order:
 *  1. n_sub (>= K) row indices (long)(Rand() * N), with replacement;
 *  2. the rows are normalised by their column means c and one scale s (the
 *     root mean square deviation over all coordinates), the BF16 forward runs
 *     on them, and the point head's outputs m_i (each point moved towards its
 *     cluster's mean) are seeded by the k-means++ rule of
 *     Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84 (first row
 *     uniform, then D^2 draws);
 *  3. the proposal C0 = c + s * seeds (K x d) goes through the device Lloyd
 *     sweeps and cluster mixture of fem_kmeans_init_seeded_rng.
 * fem_setnet_propose_means[_rng] stop after step 2 and write C0.
 * fem_setnet_propose[_rng] runs all three; if the proposal leaves a cluster
 * with fewer than two points, or, when the context has no covariance
 * regularisation (fem_set_regularization g = 0), fewer than d + 1 (a singular
 * covariance, from which unregularised EM cannot start), k-means++
 * (fem_kmeans_init_rng) on the same stream is the fallback.  *used = 1 for the network's proposal, 0 for the
 * fallback.  The result becomes the context's parameters.
 * fem_setnet_judge[_rng] judges the proposal by the final EM log-likelihood:
 * it takes the learned initialisation and then the k-means++ one (same
 * stream), runs EM from each on the context to the reference's stopping rule
 * (regularisation g before the loop and after each M-step, stop at
 * |dLL/LL| < eps, or an LL decrease, or em_sweeps E-steps;
 * EM2/EMAlgorithm.cpp:139-191), and keeps the converged mixture with the
 * higher final log-likelihood (the learned one on a tie): *used = 1 learned,
 * 0 learned path's k-means++ fallback, 2 the k-means++ start; ll[2] (may be
 * NULL) = final LLs from the learned and the k-means++ starts (-inf where EM
 * failed).  Errors through fem_last_error(ctx). */
int fem_setnet_propose_means_rng(fem_setnet *net, fem_ctx *ctx, int n_sub, double (*rnd)(void *),
                                 void *rnd_state, double *C0);
int fem_setnet_propose_means(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, double *C0);
int fem_setnet_propose_rng(fem_setnet *net, fem_ctx *ctx, int n_sub, double (*rnd)(void *), void *rnd_state,
                           int max_sweeps, double *w, double *mu, double *cov, int *used);
int fem_setnet_propose(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, int max_sweeps, double *w,
                       double *mu, double *cov, int *used);
int fem_setnet_judge_rng(fem_setnet *net, fem_ctx *ctx, int n_sub, double (*rnd)(void *), void *rnd_state,
                         int max_sweeps, int em_sweeps, double eps, double g, double *w, double *mu, double *cov,
                         int *used, double *ll);
int fem_setnet_judge(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, int max_sweeps, int em_sweeps,
                     double eps, double g, double *w, double *mu, double *cov, int *used, double *ll);

#ifdef __cplusplus
}
#endif
#endif
