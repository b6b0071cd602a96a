/*
 * fastem_abi.h -- C ABI of the MI355X Gaussian-mixture EM engine (libfastem.so).
 *
 * Plain C types only.  One context owns the device copy of one dataset, the
 * mixture parameters, every scratch buffer and one HIP stream.  Host arrays
 * passed in are borrowed for the duration of the call.  Every call is
 * synchronous at return unless stated otherwise.  Calls on one context are not
 * thread-safe.  Return value: FEM_OK (0) or a negative FEM_ERR_* code, with a
 * message from fem_last_error().
 *
 * Layouts (row-major, FP64 unless stated): weights[K]; means[K][d];
 * covariances[K][d][d]; data X[n][ld] FP32; posteriors R[n][ld] FP64.
 *
 * Each entry point names the reference hook it replaces (paths relative to
 * the fastem source tree).
 */
#ifndef FASTEM_ABI_H
#define FASTEM_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fem_ctx fem_ctx;

enum fem_status {
    FEM_OK = 0,
    FEM_ERR_ARG = -1,     /* bad argument or unsupported shape */
    FEM_ERR_HIP = -2,     /* HIP runtime error */
    FEM_ERR_NOT_PD = -3,  /* a covariance is not positive definite (see info[]) */
    FEM_ERR_STATE = -4    /* call out of order (e.g. E-step before factor) */
};

/* Arithmetic of the E-step and M-step contractions.  Cholesky, log-det and
 * the final reduction of the sufficient statistics are FP64 in every mode. */
enum fem_precision {
    FEM_PREC_FP64 = 0,    /* FP64 MFMA everywhere: the parity form */
    FEM_PREC_FP32 = 1,    /* FP32 MFMA contractions, FP32 posteriors */
    FEM_PREC_BF16X3 = 2   /* BF16x3 split products, FP32 accumulation */
};

/* Allocates a context for n points of dimension d and K components on HIP
 * device `device`.  Replaces the constructors EMAlgorithm(ncols,nrows,ncl)
 * (EM2/EMAlgorithm.cpp:115-149) and MatrixEMBase/MatrixEM/SimpleMatrixEM
 * (EM2/MatrixEM.cpp:10-30,161-174,361-374): Posteriors, Dataset and the
 * per-thread reducer buffers become device buffers.  1 <= d <= 128, K >= 1. */
int fem_create(fem_ctx **ctx, int64_t n, int d, int K, int precision, int device);
int fem_destroy(fem_ctx *ctx);

/* Copies the FP32 dataset (n rows, row stride ld floats) from host memory to
 * HBM.  Replaces EMAlgorithm::CopyDataset / MatrixEM::CopyDataset
 * (EM2/EMAlgorithm.cpp:225-235, EM2/MatrixEM.cpp:102-121); the FP32->FP64
 * widening happens in registers inside the kernels. */
int fem_upload_data(fem_ctx *ctx, const float *X, int64_t ld);
/* Same, from a device pointer already in HBM (no PCIe transfer). */
int fem_upload_data_device(fem_ctx *ctx, const float *dX, int64_t ld);

/* Switches the arithmetic of later E-/M-steps (allocating the posterior
 * buffer of the new form on first use; the dataset and parameters stay in
 * HBM).  A fem_factor must follow before the next fem_estep.  Used to finish
 * a mixed-precision fit in FP64. */
int fem_set_precision(fem_ctx *ctx, int precision);

/* Number of rows the weights are normalised by (EMAlgorithm::TotalRowCount,
 * EM2/EMAlgorithm.cpp:134; MatrixEM.cpp:375).  Defaults to n. */
int fem_set_total_rows(fem_ctx *ctx, int64_t ntotal);

/* Covariance regularisation coefficient g in [0,1] applied by every M-step:
 * C <- C(1-g) + g tr(C)/d I  (EMAlgorithm::RegularizeCovariances,
 * EM2/EMAlgorithm.cpp:237-244, called after MStep at :158). */
int fem_set_regularization(fem_ctx *ctx, double g);

/* Uploads mixture parameters (EMAlgorithm::DecodeMixture,
 * EM2/EMAlgorithm.cpp:125-131).  Covariances are read from their lower
 * triangle when factorised, as NormalDensity::Init does. */
int fem_set_params(fem_ctx *ctx, const double *w, const double *mu, const double *cov);
/* Downloads the current parameters; any pointer may be NULL. */
int fem_get_params(fem_ctx *ctx, double *w, double *mu, double *cov);
/* fem_get_params in two halves, for a host that works on the parameters
 * while the device runs on: _start queues the copies into pinned staging the
 * context owns (no synchronisation), _finish waits for them and copies out
 * (any pointer may be NULL).  _finish may run on another host thread than
 * the one queueing further work on the context; a pair must not overlap
 * another pair or fem_destroy.  No reference counterpart (the reference's
 * parameters never leave host memory); it lets `em2` overlap
 * EMAlgorithm::FindLargestConditionNumber (EM2/EMAlgorithm.cpp:159-163) with
 * the next E-step without a host-side wait between the two. */
int fem_get_params_start(fem_ctx *ctx);
int fem_get_params_finish(fem_ctx *ctx, double *w, double *mu, double *cov);

/* Stage (d): batched FP64 Cholesky, log-determinant and U_k = (L_k^-1)^T for
 * every component.  Replaces EMAlgorithm::InitDensities ->
 * NormalDensity::Init (EM2/EMAlgorithm.cpp:75-87, EM2/NormalDensity.cpp:43-65)
 * and MatrixEMBase::CalcLogCoeffs (EM2/MatrixEM.cpp:83-87).  logdet[K] and
 * info[K] may be NULL; info[k] = 0, or j+1 if the pivot of column j was not
 * positive.  Returns FEM_ERR_NOT_PD if any component failed. */
int fem_factor(fem_ctx *ctx, double *logdet, int *info);

/* Stage (a): fused E-step -- log-densities through U_k on MFMA, log-sum-exp
 * over K, posteriors -- and the log-likelihood of the current parameters.
 * Replaces MatrixEM/SimpleMatrixEM::PrecomputeDensityMatrix +
 * ComputeLogLike (EM2/MatrixEM.cpp:39-81,126-157,275-307; MatrixEM.h:88). */
int fem_estep(fem_ctx *ctx, double *loglik);

/* Stages (b)+(c): N_k and means (R^T X), then covariances about the new
 * means (weighted SYRK), weights N_k/ntotal, then regularisation.  Replaces
 * MatrixEM/SimpleMatrixEM::MStep/doMStep (EM2/MatrixEM.cpp:176-249,309-378)
 * including both OpenMP reducers (EM2/OpenMPEMReducer.cpp, Util/OMPReducer.h).
 * The outputs may be NULL (parameters stay on the device). */
int fem_mstep(fem_ctx *ctx, double *w, double *mu, double *cov);

/* Copies the posteriors of the last E-step to host memory (n rows, stride ld
 * doubles).  The reference's EMAlgorithm::Posteriors matrix
 * (EM2/EMAlgorithm.h:41) read by RetrieveClasses (EM2/EMAlgorithm.cpp:95-110). */
int fem_get_posteriors(fem_ctx *ctx, double *R, int64_t ld);

/* Seeding of the k-means initialisation (the reference's KMeansInitializer
 * subclasses). */
enum fem_seeding {
    FEM_SEED_FORGY = 0,      /* K rows at (long)(Rand()*N), Clust/KMeansInitializer.cpp:33-49 */
    FEM_SEED_KMEANSPP = 1,   /* Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84 */
    FEM_SEED_KMEANSOROR = 2, /* k-means||, oversampling 1.4, 5 rounds, then weighted k-means++
                                over the candidates: Clust/KMeansOrOrInitializer.cpp:15-174,
                                EM2/EMInitializer.cpp:158-161 (one-thread draw order) */
    FEM_SEED_KMEANSPP_GREEDY = 3 /* no counterpart in the reference: greedy k-means++ (Arthur and
                                Vassilvitskii 2007, section 6): per seed, L = 2 + floor(ln K)
                                candidates drawn by the k-means++ rule, one variate each, and
                                the one of least potential sum_i min(D_i^2, |x_i - c|^2) kept
                                (ties to the earlier draw); the batched workload's restarts */
};

/* Class of each of the n rows (cls[n], host): the component of largest
 * posterior from the last E-step, ties to the lower index; computed on the
 * device.  Replaces EMAlgorithm::RetrieveClasses (EM2/EMAlgorithm.cpp:107-122). */
int fem_get_classes(fem_ctx *ctx, int *cls);

/* k-means initialisation on the context's data, the step before the EM path:
 * a seeding (enum fem_seeding), up to max_sweeps Lloyd sweeps with MinRel
 * 1e-4 (Clust/KMAlgorithm.cpp:44-87,205-255; the reference's MaxIter =
 * max_sweeps + 1; empty clusters re-seeded by Clust/CentroidRepair.cpp:21-28),
 * then the mixture of the final clustering (EM2/EMInitializer.cpp:103-134:
 * covariances about the centroids divided by n_k - 1, weights n_k / n).
 * Replaces KMeansEMInitializer::Init and KMeansOrOrEMInitializer::Init.
 * Distances are FP64 in ascending feature order without multiply-add
 * contraction (k-means||: float differences and squares summed in FP64, as
 * Clust/CentroidVector.h:79-87), ties go to the lower index, prefix sums,
 * costs and draws are sequential FP64 on the host.  The random stream is
 * std::mt19937_64(seed) through generate_canonical (Util/Rand.cpp:7-14), or
 * rnd(rnd_state) for the _rng variant.  The result becomes the context's
 * parameters and is copied to w[K], mu[K][d], cov[K][d][d].  Returns
 * FEM_ERR_ARG (after writing the result) if a cluster has fewer than two
 * points.  Overwrites the posteriors. */
int fem_kmeans_init(fem_ctx *ctx, uint64_t seed, int seeding, int max_sweeps, double *w, double *mu,
                    double *cov);
int fem_kmeans_init_rng(fem_ctx *ctx, double (*rnd)(void *), void *rnd_state, int seeding,
                        int max_sweeps, double *w, double *mu, double *cov);
/* The same Lloyd sweeps and cluster mixture from caller-supplied starting
 * centroids C0[K][d] instead of a seeding pass (the reference's
 * FileEMInitializer path reads a finished mixture, EM2/EMInitializer.cpp:148-151;
 * this variant is used by the learned initializer, whose proposals replace the
 * seeding).  Empty clusters are re-seeded from rows drawn with
 * std::mt19937_64(seed), or rnd(rnd_state) for the _rng variant.  Same result
 * and error convention as fem_kmeans_init. */
int fem_kmeans_init_seeded_rng(fem_ctx *ctx, const double *C0, double (*rnd)(void *), void *rnd_state,
                               int max_sweeps, double *w, double *mu, double *cov);
int fem_kmeans_init_seeded(fem_ctx *ctx, const double *C0, uint64_t seed, int max_sweeps, double *w,
                           double *mu, double *cov);

/* ---- Many independent fits at once ------------------------------------------
 * A batch holds nfits datasets of uneven (n, d <= 64, K) in pooled HBM buffers.
 * Every EM phase is one persistent launch over a work list covering all
 * active fits; one iteration of all fits is captured in a hipGraph and
 * replayed until every fit has stopped; each fit follows the stopping rule of
 * EMAlgorithm::TrainNoCopy (EM2/EMAlgorithm.cpp:152-183) on the device.
 * Per-fit arithmetic is the FP64 form's.  No reference counterpart: it serves
 * random restarts and model-order (K) sweeps scored by BIC. */
typedef struct fem_batch fem_batch;
int fem_batch_create(fem_batch **batch, int nfits, const int64_t *n, const int *d, const int *K,
                     int device);
int fem_batch_destroy(fem_batch *batch);
int fem_batch_upload(fem_batch *batch, int fit, const float *X, int64_t ld);
int fem_batch_set_params(fem_batch *batch, int fit, const double *w, const double *mu,
                         const double *cov);
/* Autotuner: per class of fits with the same (16-feature blocks, padded K)
 * it times that class's work under each candidate tile shape -- E-step point
 * groups per wave (2, 4), covariance chunk rows (1, 2, 4 times the base
 * length) -- and keeps the fastest, then times the persistent grids of the
 * combined work lists.  The current parameters are restored (report: text,
 * may be NULL). */
int fem_batch_tune(fem_batch *batch, char *report, int report_len);
/* EM on all fits until each converges (|dLL/LL| < eps), fails (Cholesky) or
 * reaches maxsweeps E-steps; g = the regularisation coefficient, applied to the
 * initial covariances (g > 0) and after every M-step, as TrainNoCopy does. */
int fem_batch_run(fem_batch *batch, int maxsweeps, double eps, double g, int use_graph);
/* status: 1 converged, 2 sweep cap, -1 Cholesky failure; ll = LL of the last
 * E-step; iter_counter = E-steps + 1 (the reference's IterCounter at exit). */
int fem_batch_get_result(fem_batch *batch, int fit, double *w, double *mu, double *cov, double *ll,
                         int *iter_counter, int *status);
const char *fem_batch_last_error(const fem_batch *batch);

/* ---- No allocation ----------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fem_factor + fem_estep of the FP64 form over
 * caller-owned device buffers, with no context.  It allocates nothing: its
 * working memory (factor products, a zero-padded tile of rows, the tile's
 * posteriors, log-likelihood partials) is one static 64 MiB device array per
 * GPU, held until the process ends.  Limits: 1 <= d <= 128, 1 <= K <= 128
 * (FEM_ERR_ARG beyond).  Every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU race: serialise them.  Runs on the current
 * device, on `stream` (a hipStream_t; NULL = the null stream); synchronous at
 * return; errors through fem_last_error(NULL).
 *   X        device, n rows of ldx floats (ldx >= d)
 *   w,mu,cov device FP64: K; K x d; K x d x d (covariances read from their
 *            lower triangle, as fem_factor does)
 *   R        device FP64, n rows of ldr (ldr >= K): the posteriors, equal bit
 *            for bit to fem_estep's (fem_get_posteriors layout)
 *   loglik   host, may be NULL: the log-likelihood (fem_estep's up to
 *            summation order: tiles of at most 65,536 rows are summed in order)
 *   info     host, K ints, may be NULL: as fem_factor; FEM_ERR_NOT_PD if any
 *            component is not positive definite (R is then not written). */
int fem_estep_noalloc(const float *X, int64_t n, int64_t ldx, int d, int K, const double *w,
                      const double *mu, const double *cov, double *R, int64_t ldr, double *loglik,
                      int *info, void *stream);

/* Device time in milliseconds of the last fem_estep / fem_mstep (means and
 * covariance passes) / fem_factor, from HIP events on the context's stream.
 * Any pointer may be NULL. */
int fem_last_times(fem_ctx *ctx, float *estep_ms, float *means_ms, float *cov_ms,
                   float *factor_ms);

/* Which E-step kernel the last fem_estep ran (no counterpart in the
 * reference; for tests and profiling).  In the BF16x3 form the first E-step
 * after the data, the means (fem_set_params, an initializer) or the precision
 * were set centres every (point, component) pair on mu_k; later ones, for
 * d <= 64 and the component counts whose stage fits two workgroups per CU
 * (K <= 75 at d = 64, 118 at d = 48), centre each point once on the component of its
 * largest log-density in the previous E-step and add U_k^T (mu_a - mu_k)
 * from a table (FASTEM_ESTEPB_REF=0 at fem_create / fem_set_precision keeps
 * the first kind throughout). */
enum fem_estep_form {
    FEM_ESTEP_FP64 = 0,
    FEM_ESTEP_FP32 = 1,
    FEM_ESTEP_BF16X3 = 2,      /* centred on each component's mean */
    FEM_ESTEP_BF16X3_REF = 3   /* centred on a reference component per point */
};
int fem_last_estep_form(fem_ctx *ctx, int *form);

/* No counterpart in the reference (which evaluates every component for every
 * point, EM2/MatrixEM.cpp:69-87).  The FP64 E-step of a context with K <= 64
 * and d > 16 keeps the points, inside the context, sorted by the component
 * of their largest log-density in the last E-step, takes each workgroup's
 * components first and skips a component for a wave of 32 points when the
 * first 16 rows of U_k^T (x - mu_k) already put every point's log-density
 * more than 760 below its running maximum: that log-density's posterior is
 * exactly zero, and the E-step's outputs are bit for bit those of the same
 * order without the skip (posteriors and classes leave the context in point
 * order).  *frac = the share of (wave, component) evaluations the last FP64
 * E-step skipped (0 for other forms).  FASTEM_ESTEP_DENSE=1 at fem_create
 * evaluates every component in point order. */
int fem_last_estep_skipped(fem_ctx *ctx, double *frac);

/* Message of the last failed call on ctx (or of the last failed fem_create
 * when ctx is NULL). */
const char *fem_last_error(const fem_ctx *ctx);

/* Library version string. */
const char *fem_version(void);

#ifdef __cplusplus
}
#endif
#endif
