/*
 * fastem_tools.h -- host-side tools of libfastem that sit beside the EM path:
 * the seeded synthetic-mixture generator (the workloads' data source) and the
 * k-means++ mixture initializer.  Plain C types; FP64 row-major parameters as
 * in fastem_abi.h.
 */
#ifndef FASTEM_TOOLS_H
#define FASTEM_TOOLS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Seeded Gaussian mixture ("gengmm").  With a master std::mt19937_64(seed):
 * weights ~ Dirichlet(1) (normalised -log(1-U)), means ~ U[mean_lo,mean_hi]^d,
 * covariances A A^T / d + diag I with A_ij ~ N(0,1) (Box-Muller).  Rows come in
 * blocks of 8192; block b draws from mt19937_64(seed + 0x9E3779B97F4A7C15*(b+1)):
 * label by inverse CDF of the weights, x = mu + L z, stored FP32.  Any output
 * pointer except X may be NULL.  Deterministic for any thread count. */
int fem_gengmm(int64_t n, int d, int K, double mean_lo, double mean_hi, double diag,
               uint64_t seed, float *X, int *labels, double *w, double *mu, double *cov);

/* k-means++ mixture initializer on host data: copies X to a temporary FP64
 * device context and runs fem_kmeans_init (include/fastem_abi.h) with
 * k-means++ seeding, std::mt19937_64(seed) as the Rand() stream and at most
 * lloyd_iters Lloyd sweeps (the reference's KMeansEMInitializer uses MaxIter
 * 30, i.e. 29 sweeps; 0 = seeding only).  Needs a GPU.  Returns 0, or a
 * negative FEM_ERR_* code (FEM_ERR_ARG if a cluster ends with fewer than two
 * points). */
int fem_kmeanspp_init(int64_t n, int d, int K, const float *X, uint64_t seed, int lloyd_iters,
                      double *w, double *mu, double *cov);

#ifdef __cplusplus
}
#endif
#endif
