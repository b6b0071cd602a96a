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

/* Condition number |lambda_max / lambda_min| of the lower-symmetrised d x d
 * matrix C, through NormalDensity::Init + ComputeConditionNumber
 * (EM2/NormalDensity.cpp:43-71), the check em2 runs on every component
 * unless -S is given.  Host only.  FEM_ERR_NOT_PD if C is not positive
 * definite, FEM_ERR_ARG if the eigenvalue iteration does not converge. */
int fem_condition_number(int d, const double *C, double *cond);

/* Bhattacharyya distance between N(mu1, C1) and N(mu2, C2) through
 * BhattDist::ComputeDist (EM2/NormalDensity.cpp:128-138).  Host only.
 * FEM_ERR_NOT_PD if a covariance (or their mean) is not positive definite. */
int fem_bhatt_dist(int d, const double *mu1, const double *C1, const double *mu2, const double *C2,
                   double *dist);

#ifdef __cplusplus
}
#endif
#endif
