// fem_internal.h -- device context and kernel launchers of libfastem.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/fastem_abi.h"

namespace fem {

// Points are padded to a multiple of this; padded rows of X are zero and
// their posteriors are written as zero, so they never contribute.
constexpr int64_t kRowAlign = 4096;
// Largest supported dimension (the factor kernel keeps L in LDS).
constexpr int kMaxDim = 128;
// Candidates per seed of the greedy k-means++ seeding: 2 + floor(ln K) <= 6
// for K <= 128.
constexpr int kMaxTrials = 8;

// Number of (jb, lb) fragment pairs the E-step consumes for NB 16-feature
// blocks: lb runs over 4-feature blocks, jb >= lb/4 (U is upper triangular).
inline int estep_pairs(int NB) {
    int p = 0;
    for (int lb = 0; lb < 4 * NB; lb++) p += NB - lb / 4;
    return p;
}
inline int upper_blocks(int NB) { return NB * (NB + 1) / 2; }
// FP32 E-step stage block per component: fragments, mu_hi, -U^T mu_lo.
inline int stage32_words(int NB) { return estep_pairs(NB) * 64 + 2 * 16 * NB; }
// FP64 E-step stage block per component (fem_tiles.h stage64_words_of, checked
// equal there): 2 NB (NB - 1) 16x16x4 fragments of 64 doubles, 10 NB 4x4x4
// quads of 16 doubles, -U^T mu.
constexpr int stage64_words(int NB) { return 2 * NB * (NB - 1) * 64 + 10 * NB * 16 + 16 * NB; }

struct Launch {
    hipStream_t stream;
};

}  // namespace fem

struct fem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0, npad = 0, ntotal = 0;
    int d = 0, dpad = 0, NB = 0, K = 0, KP = 0, prec = 0;
    double g = 0.0;
    bool have_params = false, factored = false, have_post = false;

    // dataset and posteriors
    float *X = nullptr;        // npad x dpad, FP32, zero padded
    double *Rt = nullptr;      // KP x npad FP64 posteriors (component-major)
    float *Rt32 = nullptr;     // KP x npad FP32 posteriors (mixed precision)

    // parameters (dense d)
    double *w = nullptr;       // K
    double *mu = nullptr;      // K x d
    double *cov = nullptr;     // K x d x d
    double *nk = nullptr;      // K   posterior sums of the last M-step

    // factorisation products
    double *logdet = nullptr;  // K
    int *info = nullptr;       // K
    double *coef = nullptr;    // K   log w_k - d/2 log 2pi - logdet_k/2
    int *korder = nullptr;     // K   components by descending weight: the FP64 E-step's order
    double *U = nullptr;       // K x dpad x dpad  U_k = L_k^-T (zero padded)
    double *stage64 = nullptr; // K x (pairs x 64 + dpad): FP64 E-step fragments, -U_k^T mu_k
    float *stage32 = nullptr;  // K x stage32_words: FP32 fragments, mu_hi, -U_k^T mu_lo
    unsigned int *stageb = nullptr;  // K x stage words: BF16 hi/lo E-step fragments, mu_hi, U_k^T mu_lo
    // reference-centred BF16x3 E-step (allocated where it fits, d <= 64)
    unsigned int *stager = nullptr;  // K stage blocks (fragments, seed table), then the mu_hi table
    int *refc = nullptr;             // npad: reference component of each point
    bool ref_ok = false;   // refc was written by an E-step since the data or means were last set
    bool stage_ref = false;  // the last BF16x3 prep packed stager (else stageb)
    int last_form = 0;       // fem_estep_form of the last fem_estep

    // reductions
    int est_blocks = 0;         // capacity of ll_part (one per E-step workgroup)
    int estep_groups = 2;       // 16-point groups per wave in the FP64 E-step (2 or 4; 4 for d <= 32)
    int estep_groups32 = 2;     // same for the FP32 E-step
    int estep_groupsb = 4;      // same for the BF16x3 E-step (4 measured faster at d <= 64)
    double *ll_part = nullptr;  // per E-step workgroup
    double *ll_dev = nullptr;   // 1
    int m_chunks = 0, c_chunks = 0;
    int64_t c_chunk_rows = 0;
    // covariances over the nonzero posteriors only (FASTEM_COV_DENSE=1: every point)
    bool cov_sparse = true;

    // FP64 E-step pruning (K <= 64, d > 16; FASTEM_ESTEP_DENSE=1 turns it
    // off): the EM kernels run over the points in slot order, Xs[s] = X[perm[s]]
    // once the points are sorted by the component of their largest
    // log-density (ref, per slot), so a workgroup's points share few
    // components; the E-step takes those first and skips, per wave, every
    // other component whose first 16 whitened coordinates already put all of
    // the wave's points beyond exp's underflow range.  Posteriors and classes
    // leave the context in point order.
    bool prune = true;
    bool sorted = false;        // Xs and perm hold a sorted order
    float *Xs = nullptr;        // npad x dpad, X in slot order
    int *perm = nullptr;        // npad: point of each slot
    int *perm2 = nullptr;       // npad: next permutation (sort output)
    int *ref = nullptr;         // npad: component of largest log-density per slot (K: none)
    int *ref2 = nullptr;        // npad: sorted keys
    int *slots = nullptr;       // npad: 0..npad-1 (sort values in)
    int *rel = nullptr;         // npad: old slot of each new slot (sort values out)
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    unsigned long long *nskip = nullptr;  // wave-component evaluations skipped by the last FP64 E-step
    int esteps_since_sort = 0, sort_every = 1;
    bool refs_valid = false;    // ref was written by a pruned E-step on the current order
    const float *Xe() const { return sorted ? Xs : X; }
    double *mpart = nullptr;    // m_chunks x KP x (dpad+1)
    double *cpart = nullptr;    // c_chunks x KP x ublocks x 256

    // k-means scratch (allocated on first use)
    double *kC = nullptr;       // K x d centroids
    double *dmin = nullptr;     // npad  k-means++ closest squared distances
    int *cls = nullptr;         // npad  cluster of each point
    int64_t *kpick = nullptr;   // 2     k-means++ draw decided on the device (row, decided)
    double *kfit = nullptr;     // per-workgroup fit partials
    double *kcand = nullptr;    // kMaxTrials x d greedy k-means++ candidates
    double *kpot = nullptr;     // (npad / 256) x kMaxTrials candidate potential partials
    bool owns_rt64 = false;     // Rt allocated for k-means in a non-FP64 context

    hipEvent_t ev[8] = {};
    double *params_host = nullptr;     // pinned staging of fem_get_params_start: w, mu, cov
    hipEvent_t ev_params = nullptr;    // recorded after the staging copies
    bool params_pending = false;       // a start whose finish has not run
    float t_estep = 0, t_means = 0, t_cov = 0, t_factor = 0;
    std::string err;
};

namespace fem {
// Records msg as the error fem_last_error(NULL) reports; returns code.
int fail_global(int code, const std::string &msg);
// Rows idx[0..m) of the context's data to host memory (m x d floats).
int gather_rows_host(fem_ctx *c, const int64_t *idx, int m, float *out);
hipError_t launch_factor(fem_ctx *c);
hipError_t launch_estep(fem_ctx *c);
hipError_t launch_means(fem_ctx *c);
hipError_t launch_cov(fem_ctx *c);
hipError_t launch_zero(double *p, int64_t n, hipStream_t s);
hipError_t launch_transpose_post(const fem_ctx *c, double *out_dev, int64_t ld);
hipError_t launch_classes(const fem_ctx *c, int *cls_dev);
// the pruned FP64 E-step's slot order (fem_sort.hip)
hipError_t sort_points(fem_ctx *c);
void reset_order(fem_ctx *c);
hipError_t launch_unpermute_int(const fem_ctx *c, const int *in, int *out);
// Row copies between pageable host memory and pitched device memory (bytes),
// staged through two pinned buffers so the DMA engine moves large contiguous
// chunks while the host packs or unpacks rows; device pad bytes beyond
// `width` are written as zero.  Synchronous at return.
hipError_t copy_rows_h2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, int64_t rows,
                         hipStream_t s);
hipError_t copy_rows_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, int64_t rows,
                         hipStream_t s);
hipError_t launch_means64_partials(fem_ctx *c);
hipError_t launch_means_finalize(fem_ctx *c, double *mu_out, double *nk_out, int raw);
hipError_t launch_cov64_partials(fem_ctx *c);
hipError_t launch_cov64_narrow(fem_ctx *c);  // NB <= 3 (fem_cov64_narrow.hip)
hipError_t launch_cov_finalize(fem_ctx *c, double g);
// FP32 form (fem_kernels32.hip)
hipError_t launch_estep32(fem_ctx *c);
hipError_t launch_means32_partials(fem_ctx *c);
hipError_t launch_cov32_partials(fem_ctx *c);
hipError_t launch_transpose_post32(const fem_ctx *c, double *out_dev, int64_t ld);
// BF16x3 form (fem_kernels_bf16.hip); posteriors in Rt32
int estepb_stage(int NB);
hipError_t launch_prep_bf16(fem_ctx *c);
hipError_t launch_estepb(fem_ctx *c);
// the reference-centred BF16x3 E-step handles (NB, K); words of its stage
bool estepr_fits(int NB, int K);
size_t estepr_words(int NB, int K);
hipError_t launch_meansb_partials(fem_ctx *c);
hipError_t launch_covb_partials(fem_ctx *c);
}  // namespace fem
