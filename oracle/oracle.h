/*
 * oracle.h -- FP64 CPU restatement of fastem's single-node Gaussian-mixture
 * EM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product (libfastem) never links, loads or calls it, and has no CPU
 * fallback.
 *
 * Every function names the reference file:line whose semantics it follows
 * (paths relative to the fastem source tree).  All arithmetic is IEEE double,
 * compiled without -ffast-math, with a fixed "virtual thread" partition so the
 * result does not depend on how many OpenMP threads actually run.
 *
 * Layouts: data X is N x d row-major FP32; weights K; means K x d row-major;
 * covariances K x d x d row-major; posteriors N x K row-major FP64;
 * U_k = (L_k^-1)^T stored as K x d x d row-major (upper triangular).
 */
#ifndef FASTEM_ORACLE_H
#define FASTEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Util/Rand.cpp: std::mt19937_64 + uniform_real_distribution<double> ---- */
typedef struct { uint64_t mt[312]; int mti; } orc_mt64;
void     orc_mt64_seed(orc_mt64 *s, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *s);
double   orc_mt64_canonical(orc_mt64 *s);

/* ---- Util/OpenMP.h:37-46 FindOpenMPItems ---- */
void orc_find_items(int64_t n, int size, int id, int64_t *offset, int64_t *count);
/* ---- EM2/BlockManager.cpp:10-28; returns 0, or -1 for a bad spec ---- */
int  orc_parse_block_params(const char *spec, int64_t n, int d, int K, int *ndens, int *nmstep);

/* ---- EM2/NormalDensity.cpp:43-65: 0 on success, -1 if not positive definite ---- */
int  orc_factor(int d, const double *C, double *U, double *logdet);
/* ---- EM2/NormalDensity.cpp:67-71 (eigenvalues only; ascending into ev) ---- */
int  orc_sym_eigvals(int d, const double *C, double *ev);
double orc_cond(int d, const double *C);
/* ---- EM2/EMAlgorithm.cpp:237-244 ---- */
void orc_regularize(int K, int d, double g, double *covs);

/* ---- EM2/MatrixEM.cpp:39-87, 275-307 (SimpleMatrixEM E-step) ----
 * nvt = virtual threads, nblk = sub-blocks per thread.  Returns LL. */
double orc_estep(int64_t n, int d, int K, const float *X, const double *w,
                 const double *mu, const double *U, const double *logdet,
                 double *R, int nvt, int nblk);
/* ---- EM2/MatrixEM.cpp:309-378 (SimpleMatrixEM M-step); log2 != 0 selects the
 * tree reducer of Util/OMPReducer.h:69-80, else the serial one (:51-56). ---- */
void orc_mstep(int64_t n, int d, int K, const float *X, const double *R,
               int64_t ntotal, double *w, double *mu, double *cov,
               int nvt, int nblk, int log2red);

/* ---- EM2/EMAlgorithm.cpp:139-191 TrainNoCopy loop.  Parameters are in/out.
 * ll_trace (may be NULL) receives the LL of every sweep (up to maxsweeps).
 * Returns the final IterCounter (reference semantics) or a negative code:
 * -1 Cholesky failure, -2 condition-number abort. *out_ll gets the returned LL. */
int orc_train(int64_t n, int d, int K, const float *X, double *w, double *mu,
              double *cov, int maxsweeps, double eps, double athr, double g,
              int svd, int nvt, int nblk, int log2red, double *ll_trace,
              double *out_ll);

/* ---- Clust/CentroidVector.cpp:56-73 ClassifyDataset (ties -> lowest index) ---- */
void orc_classify(int64_t n, int d, int K, const float *X, const double *C, int *cls);
/* ---- Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84 ---- */
void orc_kmeanspp(int64_t n, int d, int K, const float *X, orc_mt64 *rng, double *C);
/* ---- Clust/KMAlgorithm.cpp:44-87,205-255 NaiveKMA::RunKMeans (random repair
 * from Clust/CentroidRepair.cpp:21-28).  Returns the number of sweeps. ---- */
int  orc_lloyd(int64_t n, int d, int K, const float *X, double *C, double minrel,
               int maxiter, orc_mt64 *rng, double *bestfit);
/* ---- EM2/EMInitializer.cpp:103-134 tail: mixture from a clustering ---- */
void orc_gmm_from_clusters(int64_t n, int d, int K, const float *X, const double *C,
                           const int *cls, double *w, double *mu, double *cov);

/* ---- Clust/KMeansInitializer.cpp:33-49 ForgyInitializer::Init ---- */
void orc_forgy(int64_t n, int d, int K, const float *X, orc_mt64 *rng, double *C);
/* ---- Util/MultithreadedDataset.cpp:179-283 TraceCov (sequential row order) ---- */
double orc_data_trace(int64_t n, int d, const float *X);
/* ---- EM2/EMInitializer.cpp:43-80 JainEMInitializer::Init ---- */
void orc_jain(int64_t n, int d, int K, const float *X, double trace, orc_mt64 *rng,
              double *w, double *mu, double *cov);
/* ---- Clust/KMeansOrOrInitializer.cpp:143-174 + weighted k-means++ (one-thread
 * order); returns the number of candidate centers collected ---- */
int64_t orc_kmeansoror(int64_t n, int d, int K, const float *X, double os, orc_mt64 *rng,
                       double *C);

#ifdef __cplusplus
}
#endif
#endif
