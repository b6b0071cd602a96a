/*
 * oracle.c -- FP64 CPU restatement of fastem's single-node Gaussian-mixture EM
 * path (see oracle.h for the layouts).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Loaded only by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark; the product
 * library never links or calls it.
 *
 * The reference's arithmetic lives in Eigen 3.4 and Intel MKL (LLT, TRMM,
 * GEMM, SYRK, eigen-solver); neither is present, so the dense linear algebra
 * here is written from the textbook definitions.  Summation orders that the
 * reference leaves to BLAS are fixed to ascending index order.
 *
 * Pinning (tests/test_oracle_*.py): mt19937_64 against the C++ standard's
 * known answer; factorisation against closed forms and numpy; E-step against
 * scipy logpdf/logsumexp; M-step against numpy weighted moments and the K=1
 * closed form; whole trajectories against scikit-learn's GaussianMixture
 * stepped one iteration at a time; k-means against brute force.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ------------------------------------------------------------------------ */
/* Util/Rand.cpp:7-14,33-36 -- std::mt19937_64 (C++11 [rand.predef]) and
 * libstdc++ generate_canonical<double,53>, which for a 64-bit engine is
 * (double)x * 2^-64, clamped below 1.                                       */
#define MT_N 312
#define MT_M 156
#define MT_UPPER 0xFFFFFFFF80000000ULL
#define MT_LOWER 0x000000007FFFFFFFULL

void orc_mt64_seed(orc_mt64 *s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_N; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = MT_N;
}

static void mt64_twist(orc_mt64 *s) {
    for (int i = 0; i < MT_N; i++) {
        uint64_t y = (s->mt[i] & MT_UPPER) | (s->mt[(i + 1) % MT_N] & MT_LOWER);
        uint64_t v = s->mt[(i + MT_M) % MT_N] ^ (y >> 1);
        if (y & 1ULL) v ^= 0xB5026F5AA96619E9ULL;
        s->mt[i] = v;
    }
    s->mti = 0;
}

uint64_t orc_mt64_next(orc_mt64 *s) {
    if (s->mti >= MT_N) mt64_twist(s);
    uint64_t x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

double orc_mt64_canonical(orc_mt64 *s) {
    double r = (double)orc_mt64_next(s) / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Util/OpenMP.h:37-46 */
void orc_find_items(int64_t n, int size, int id, int64_t *offset, int64_t *count) {
    int64_t c = n / size;
    int64_t rem = n % size;
    if (id < rem) { c++; *offset = c * id; }
    else *offset = c * id + rem;
    *count = c;
}

/* EM2/BlockManager.cpp:10-28 */
int orc_parse_block_params(const char *spec, int64_t n, int d, int K, int *ndens, int *nmstep) {
    double l3;
    if (spec == NULL) { *ndens = *nmstep = 1; return 0; }
    if (sscanf(spec, "automax:%lf", &l3) == 1) {
        l3 *= (1024.0 * 1024.0);
        int nb = (int)((double)n * (16.0 * (double)d + 16.0 * (double)K) / l3 * 1.2);
        if (nb < 1) nb = 1;
        *ndens = *nmstep = nb;
        return 0;
    }
    if (sscanf(spec, "%d:%d", ndens, nmstep) != 2) return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* EM2/NormalDensity.cpp:43-65.  C is symmetrised from its lower triangle,
 * L = chol(C) (column-by-column Cholesky-Crout), logdet = 2 sum log L_ii,
 * U = (L^-1)^T by forward substitution on the identity.                    */
int orc_factor(int d, const double *C, double *U, double *logdet) {
    double *L = (double *)calloc((size_t)d * d, sizeof(double));
    double *Li = (double *)calloc((size_t)d * d, sizeof(double));
    int ok = 1;
    for (int j = 0; j < d && ok; j++) {
        double s = C[(size_t)j * d + j];
        for (int k = 0; k < j; k++) s -= L[(size_t)j * d + k] * L[(size_t)j * d + k];
        if (!(s > 0.0)) { ok = 0; break; }
        double ljj = sqrt(s);
        L[(size_t)j * d + j] = ljj;
        for (int i = j + 1; i < d; i++) {
            double t = C[(size_t)i * d + j];  /* lower triangle */
            for (int k = 0; k < j; k++) t -= L[(size_t)i * d + k] * L[(size_t)j * d + k];
            L[(size_t)i * d + j] = t / ljj;
        }
    }
    if (!ok) { free(L); free(Li); return -1; }
    double ld = 0.0;
    for (int i = 0; i < d; i++) ld += log(L[(size_t)i * d + i]);
    *logdet = 2.0 * ld;
    /* L * Li = I, column by column (forward substitution). */
    for (int c = 0; c < d; c++) {
        for (int i = c; i < d; i++) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int k = c; k < i; k++) t -= L[(size_t)i * d + k] * Li[(size_t)k * d + c];
            Li[(size_t)i * d + c] = t / L[(size_t)i * d + i];
        }
    }
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) U[(size_t)i * d + j] = Li[(size_t)j * d + i];
    free(L); free(Li);
    return 0;
}

/* Eigenvalues of the lower-symmetrised C: Householder reduction to tridiagonal
 * form, then the implicit QL iteration with Wilkinson-style shifts (Golub &
 * Van Loan, "Matrix Computations", 8.3; EISPACK tql1).  Ascending order.   */
int orc_sym_eigvals(int d, const double *C, double *ev) {
    double *A = (double *)malloc((size_t)d * d * sizeof(double));
    double *e = (double *)calloc((size_t)d + 1, sizeof(double));
    double *v = (double *)malloc((size_t)d * sizeof(double));
    double *p = (double *)malloc((size_t)d * sizeof(double));
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++)
            A[(size_t)i * d + j] = (i >= j) ? C[(size_t)i * d + j] : C[(size_t)j * d + i];
    for (int k = 0; k + 2 < d; k++) {
        int m = d - k - 1;                 /* length of x = A[k+1:, k] */
        double nrm = 0.0;
        for (int i = 0; i < m; i++) nrm += A[(size_t)(k + 1 + i) * d + k] * A[(size_t)(k + 1 + i) * d + k];
        nrm = sqrt(nrm);
        double x0 = A[(size_t)(k + 1) * d + k];
        if (nrm == 0.0) { e[k] = 0.0; continue; }
        double alpha = (x0 > 0.0) ? -nrm : nrm;
        for (int i = 0; i < m; i++) v[i] = A[(size_t)(k + 1 + i) * d + k];
        v[0] -= alpha;
        double vtv = 0.0;
        for (int i = 0; i < m; i++) vtv += v[i] * v[i];
        e[k] = alpha;
        if (vtv == 0.0) continue;
        double beta = 2.0 / vtv;
        /* p = beta * A22 v ; K = beta/2 v.p ; q = p - K v ; A22 -= v q^T + q v^T */
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int j = 0; j < m; j++) s += A[(size_t)(k + 1 + i) * d + (k + 1 + j)] * v[j];
            p[i] = beta * s;
        }
        double kk = 0.0;
        for (int i = 0; i < m; i++) kk += v[i] * p[i];
        kk *= 0.5 * beta;
        for (int i = 0; i < m; i++) p[i] -= kk * v[i];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++)
                A[(size_t)(k + 1 + i) * d + (k + 1 + j)] -= v[i] * p[j] + p[i] * v[j];
    }
    for (int i = 0; i < d; i++) ev[i] = A[(size_t)i * d + i];
    if (d >= 2) e[d - 2] = A[(size_t)(d - 1) * d + (d - 2)];
    e[d - 1] = 0.0;
    /* implicit QL on (ev, e): e[i] couples i and i+1 */
    int fail = 0;
    for (int l = 0; l < d; l++) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < d - 1; m++) {
                double dd = fabs(ev[m]) + fabs(ev[m + 1]);
                if (fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) { fail = 1; break; }
            double g = (ev[l + 1] - ev[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = ev[m] - ev[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i;
            int early = 0;
            for (i = m - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { ev[i + 1] -= pp; e[m] = 0.0; early = 1; break; }
                s = f / r; c = g / r;
                g = ev[i + 1] - pp;
                r = (ev[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                ev[i + 1] = g + pp;
                g = c * r - b;
            }
            if (early) continue;
            ev[l] -= pp; e[l] = g; e[m] = 0.0;
        }
        if (fail) break;
    }
    /* ascending insertion sort */
    for (int i = 1; i < d; i++) {
        double t = ev[i]; int j = i - 1;
        while (j >= 0 && ev[j] > t) { ev[j + 1] = ev[j]; j--; }
        ev[j + 1] = t;
    }
    free(A); free(e); free(v); free(p);
    return fail ? -1 : 0;
}

/* EM2/NormalDensity.cpp:67-71 */
double orc_cond(int d, const double *C) {
    double *ev = (double *)malloc((size_t)d * sizeof(double));
    orc_sym_eigvals(d, C, ev);
    double c = fabs(ev[d - 1] / ev[0]);
    free(ev);
    return c;
}

/* EM2/EMAlgorithm.cpp:237-244: C <- C (1-g) + (g tr(C)/d) I */
void orc_regularize(int K, int d, double g, double *covs) {
    for (int k = 0; k < K; k++) {
        double *C = covs + (size_t)k * d * d;
        double tr = 0.0;
        for (int i = 0; i < d; i++) tr += C[(size_t)i * d + i];
        double diag = g * tr / (double)d;
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++)
                C[(size_t)i * d + j] = C[(size_t)i * d + j] * (1.0 - g) + (i == j ? diag : 0.0);
    }
}

/* ------------------------------------------------------------------------ */
/* E-step, EM2/MatrixEM.cpp:69-87 (BlockDensities, CalcLogCoeffs) and :39-51
 * (PostProcessDensityMatrix), partitioned as SimpleMatrixEM (:275-307):
 * virtual thread t owns FindOpenMPItems(N, nvt, t) and walks nblk sub-blocks;
 * its LL partial is summed sequentially; partials are added in thread order.
 * Rows are independent, so evaluating row-by-row instead of
 * component-by-component gives the same values.                             */
double orc_estep(int64_t n, int d, int K, const float *X, const double *w,
                 const double *mu, const double *U, const double *logdet,
                 double *R, int nvt, int nblk) {
    double *coef = (double *)malloc((size_t)K * sizeof(double));
    const double lc = -0.5 * (double)d * log(2.0 * M_PI);
    for (int k = 0; k < K; k++) coef[k] = log(w[k]) + lc - 0.5 * logdet[k];
    double *part = (double *)calloc((size_t)nvt, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nvt; t++) {
        double *tv = (double *)malloc((size_t)d * sizeof(double));
        double *y = (double *)malloc((size_t)d * sizeof(double));
        double *s = (double *)malloc((size_t)K * sizeof(double));
        int64_t pst, pcnt;
        orc_find_items(n, nvt, t, &pst, &pcnt);
        double ll = 0.0;
        for (int b = 0; b < nblk; b++) {
            int64_t off, cnt;
            orc_find_items(pcnt, nblk, b, &off, &cnt);
            for (int64_t i = pst + off; i < pst + off + cnt; i++) {
                const float *x = X + (size_t)i * d;
                for (int k = 0; k < K; k++) {
                    const double *m = mu + (size_t)k * d;
                    const double *Uk = U + (size_t)k * d * d;
                    for (int j = 0; j < d; j++) { tv[j] = (double)x[j] - m[j]; y[j] = 0.0; }
                    /* TRMM row: y_j = sum_{l<=j} t_l U[l][j], l ascending */
                    for (int l = 0; l < d; l++) {
                        const double tl = tv[l];
                        const double *Ur = Uk + (size_t)l * d;
                        for (int j = l; j < d; j++) y[j] += tl * Ur[j];
                    }
                    double sq = 0.0;
                    for (int j = 0; j < d; j++) sq += y[j] * y[j];
                    s[k] = -0.5 * sq + coef[k];
                }
                double mx = s[0];
                for (int k = 1; k < K; k++) if (s[k] > mx) mx = s[k];
                double se = 0.0;
                for (int k = 0; k < K; k++) se += exp(s[k] - mx);
                double lse = log(se) + mx;
                ll += lse;
                double *r = R + (size_t)i * K;
                for (int k = 0; k < K; k++) r[k] = exp(s[k] - lse);
            }
        }
        part[t] = ll;
        free(tv); free(y); free(s);
    }
    double LL = 0.0;
    for (int t = 0; t < nvt; t++) LL += part[t];
    free(part); free(coef);
    return LL;
}

/* Util/OMPReducer.h:51-56 (serial) and :69-80 (log2 tree) over nvt buffers
 * of len doubles each, result in buffer 0.                                  */
static void reduce_parts(double *buf, size_t len, int nvt, int log2red) {
    if (!log2red) {
        for (int t = 1; t < nvt; t++)
            for (size_t q = 0; q < len; q++) buf[q] += buf[(size_t)t * len + q];
        return;
    }
    for (int s = 1; s < nvt; s *= 2)
        for (int t = 0; t < nvt; t += 2 * s)
            if (t + s < nvt)
                for (size_t q = 0; q < len; q++)
                    buf[(size_t)t * len + q] += buf[(size_t)(t + s) * len + q];
}

/* M-step, EM2/MatrixEM.cpp:309-378 (SimpleMatrixEM::doMStep) with
 * ComputeWeightedTempDataset (:53-67) and BlockCovariances (:90-99).       */
void orc_mstep(int64_t n, int d, int K, const float *X, const double *R,
               int64_t ntotal, double *w, double *mu, double *cov,
               int nvt, int nblk, int log2red) {
    const size_t s1 = (size_t)K * (d + 1);
    double *p1 = (double *)calloc((size_t)nvt * s1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nvt; t++) {
        double *ps = p1 + (size_t)t * s1;          /* [K] posterior sums  */
        double *ms = ps + K;                       /* [K][d] mean sums    */
        double *bps = (double *)malloc((size_t)K * sizeof(double));
        double *bms = (double *)malloc((size_t)K * d * sizeof(double));
        int64_t pst, pcnt;
        orc_find_items(n, nvt, t, &pst, &pcnt);
        for (int b = 0; b < nblk; b++) {
            int64_t off, cnt;
            orc_find_items(pcnt, nblk, b, &off, &cnt);
            memset(bps, 0, (size_t)K * sizeof(double));
            memset(bms, 0, (size_t)K * d * sizeof(double));
            for (int64_t i = pst + off; i < pst + off + cnt; i++) {
                const double *r = R + (size_t)i * K;
                const float *x = X + (size_t)i * d;
                for (int k = 0; k < K; k++) {
                    bps[k] += r[k];
                    double *m = bms + (size_t)k * d;
                    for (int j = 0; j < d; j++) m[j] += r[k] * (double)x[j];
                }
            }
            for (int k = 0; k < K; k++) ps[k] += bps[k];
            for (size_t q = 0; q < (size_t)K * d; q++) ms[q] += bms[q];
        }
        free(bps); free(bms);
    }
    reduce_parts(p1, s1, nvt, log2red);
    double *psum = p1;
    for (int k = 0; k < K; k++)
        for (int j = 0; j < d; j++)
            mu[(size_t)k * d + j] = p1[K + (size_t)k * d + j] / psum[k];

    const size_t s2 = (size_t)K * d * d;
    double *p2 = (double *)calloc((size_t)nvt * s2, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nvt; t++) {
        double *cs = p2 + (size_t)t * s2;
        double *tv = (double *)malloc((size_t)d * sizeof(double));
        int64_t pst, pcnt;
        orc_find_items(n, nvt, t, &pst, &pcnt);
        for (int b = 0; b < nblk; b++) {
            int64_t off, cnt;
            orc_find_items(pcnt, nblk, b, &off, &cnt);
            for (int k = 0; k < K; k++) {
                const double inv = 1.0 / sqrt(psum[k]);
                const double *m = mu + (size_t)k * d;
                double *C = cs + (size_t)k * d * d;
                for (int64_t i = pst + off; i < pst + off + cnt; i++) {
                    const float *x = X + (size_t)i * d;
                    const double sp = sqrt(R[(size_t)i * K + k]) * inv;
                    for (int j = 0; j < d; j++) tv[j] = ((double)x[j] - m[j]) * sp;
                    for (int a = 0; a < d; a++) {
                        const double ta = tv[a];
                        double *Cr = C + (size_t)a * d;
                        for (int c = a; c < d; c++) Cr[c] += ta * tv[c];
                    }
                }
            }
        }
        free(tv);
    }
    reduce_parts(p2, s2, nvt, log2red);
    for (int k = 0; k < K; k++) {
        const double *S = p2 + (size_t)k * d * d;
        double *C = cov + (size_t)k * d * d;
        for (int a = 0; a < d; a++)
            for (int c = 0; c < d; c++)
                C[(size_t)a * d + c] = (a <= c) ? S[(size_t)a * d + c] : S[(size_t)c * d + a];
        w[k] = psum[k] / (double)ntotal;
    }
    free(p1); free(p2);
}

/* Factorise every component (EMAlgorithm::InitDensities, :75-87).  Returns
 * -1 if any Cholesky fails.                                                 */
static int factor_all(int K, int d, const double *cov, double *U, double *ld) {
    int err = 0;
#pragma omp parallel for reduction(+:err)
    for (int k = 0; k < K; k++)
        if (orc_factor(d, cov + (size_t)k * d * d, U + (size_t)k * d * d, ld + k) != 0) err++;
    return err ? -1 : 0;
}

static double max_cond(int K, int d, const double *cov) {
    double mc = 0.0;
    for (int k = 0; k < K; k++) {
        double c = orc_cond(d, cov + (size_t)k * d * d);
        if (c > mc) mc = c;
    }
    return mc;
}

/* EM2/EMAlgorithm.cpp:133-191 (Train -> TrainNoCopy). */
int orc_train(int64_t n, int d, int K, const float *X, double *w, double *mu,
              double *cov, int maxsweeps, double eps, double athr, double g,
              int svd, int nvt, int nblk, int log2red, double *ll_trace,
              double *out_ll) {
    double *U = (double *)malloc((size_t)K * d * d * sizeof(double));
    double *ld = (double *)malloc((size_t)K * sizeof(double));
    double *R = (double *)malloc((size_t)n * K * sizeof(double));
    double prev = -DBL_MAX, LL = prev;
    int iter, rc = 0;
    orc_regularize(K, d, g, cov);
    if (factor_all(K, d, cov, U, ld) != 0) { rc = -1; iter = 0; goto done; }
    if (svd && max_cond(K, d, cov) > athr) { rc = -2; iter = 0; goto done; }
    for (iter = 1; iter <= maxsweeps; iter++) {
        LL = orc_estep(n, d, K, X, w, mu, U, ld, R, nvt, nblk);
        if (ll_trace) ll_trace[iter - 1] = LL;
        orc_mstep(n, d, K, X, R, n, w, mu, cov, nvt, nblk, log2red);
        orc_regularize(K, d, g, cov);
        if (factor_all(K, d, cov, U, ld) != 0) { rc = -1; goto done; }
        if (svd && max_cond(K, d, cov) > athr) { rc = -2; goto done; }
        double e = fabs((LL - prev) / LL);
        if (LL < prev) e = 0.0;
        prev = LL;
        if (e < eps) { iter++; break; }
    }
done:
    if (out_ll) *out_ll = LL;
    free(U); free(ld); free(R);
    return rc < 0 ? rc : iter;
}

/* ------------------------------------------------------------------------ */
/* Clust/CentroidVector.h:105-116: squared distance, double centroid minus a
 * float row promoted to double, ascending feature order.                    */
static inline double sqdist(int d, const double *c, const float *x) {
    double s = 0.0;
    for (int j = 0; j < d; j++) { double t = c[j] - (double)x[j]; s += t * t; }
    return s;
}

void orc_classify(int64_t n, int d, int K, const float *X, const double *C, int *cls) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double best = DBL_MAX; int bj = -1;
        for (int k = 0; k < K; k++) {
            double s = sqdist(d, C + (size_t)k * d, X + (size_t)i * d);
            if (s < best) { best = s; bj = k; }
        }
        cls[i] = bj;
    }
}

void orc_kmeanspp(int64_t n, int d, int K, const float *X, orc_mt64 *rng, double *C) {
    double *cd = (double *)malloc((size_t)n * sizeof(double));
    double *agg = (double *)malloc((size_t)n * sizeof(double));
    int64_t r0 = (int64_t)(orc_mt64_canonical(rng) * (double)n);
    for (int j = 0; j < d; j++) C[j] = (double)X[(size_t)r0 * d + j];
    for (int64_t i = 0; i < n; i++) cd[i] = INFINITY;
    for (int c = 1; c < K; c++) {
        const double *prevc = C + (size_t)(c - 1) * d;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            double s = sqdist(d, prevc, X + (size_t)i * d);
            if (cd[i] > s) cd[i] = s;
        }
        double sum = 0.0;
        for (int64_t i = 0; i < n; i++) { sum += cd[i] * 1.0; agg[i] = sum; }
        for (int64_t i = 0; i < n; i++) agg[i] /= sum;
        double r = orc_mt64_canonical(rng);
        for (int64_t i = 0; i < n; i++) {
            if (r <= agg[i]) {
                for (int j = 0; j < d; j++) C[(size_t)c * d + j] = (double)X[(size_t)i * d + j];
                break;
            }
        }
    }
    free(cd); free(agg);
}

/* One Lloyd sweep (NaiveKMA::ComputeMSEAndCorrect).  Fixed 16-way virtual
 * partition, partial sums combined in order.                                */
static double lloyd_sweep(int64_t n, int d, int K, const float *X, double *C, orc_mt64 *rng) {
    enum { NV = 16 };
    double *cen = (double *)calloc((size_t)NV * K * d, sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)NV * K, sizeof(int64_t));
    double fitp[NV];
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < NV; t++) {
        int64_t off, c;
        orc_find_items(n, NV, t, &off, &c);
        double f = 0.0;
        for (int64_t i = off; i < off + c; i++) {
            const float *x = X + (size_t)i * d;
            double best = DBL_MAX; int bj = -1;
            for (int k = 0; k < K; k++) {
                double s = sqdist(d, C + (size_t)k * d, x);
                if (s < best) { best = s; bj = k; }
            }
            f += best;
            cnt[(size_t)t * K + bj]++;
            double *ce = cen + ((size_t)t * K + bj) * d;
            for (int j = 0; j < d; j++) ce[j] += (double)x[j];
        }
        fitp[t] = f;
    }
    double fit = 0.0;
    for (int t = 0; t < NV; t++) fit += fitp[t];
    for (int t = 1; t < NV; t++) {
        for (size_t q = 0; q < (size_t)K * d; q++) cen[q] += cen[(size_t)t * K * d + q];
        for (int k = 0; k < K; k++) cnt[k] += cnt[(size_t)t * K + k];
    }
    for (int k = 0; k < K; k++) {
        if (cnt[k] > 0) {
            double f = 1.0 / (double)cnt[k];
            for (int j = 0; j < d; j++) C[(size_t)k * d + j] = cen[(size_t)k * d + j] * f;
        } else {
            int64_t src = (int64_t)(orc_mt64_canonical(rng) * (double)n);
            for (int j = 0; j < d; j++) C[(size_t)k * d + j] = (double)X[(size_t)src * d + j];
        }
    }
    free(cen); free(cnt);
    return fit / (double)n;
}

int orc_lloyd(int64_t n, int d, int K, const float *X, double *C, double minrel,
              int maxiter, orc_mt64 *rng, double *bestfit) {
    double best = DBL_MAX;
    int i;
    if (maxiter == 0) { if (bestfit) *bestfit = 0.0; return 0; }
    for (i = 1;; i++) {
        double fit = lloyd_sweep(n, d, K, X, C, rng);
        double rel = (best - fit) / best;
        if (fit < nextafter(fit, best)) best = fit;
        else break;
        if (minrel > 0.0 && minrel > rel) break;
        if (maxiter > 0 && i == maxiter - 1) break;
    }
    if (bestfit) *bestfit = best;
    return i;
}

void orc_gmm_from_clusters(int64_t n, int d, int K, const float *X, const double *C,
                           const int *cls, double *w, double *mu, double *cov) {
    int64_t *cnt = (int64_t *)calloc((size_t)K, sizeof(int64_t));
    double *row = (double *)malloc((size_t)d * sizeof(double));
    memcpy(mu, C, (size_t)K * d * sizeof(double));
    memset(cov, 0, (size_t)K * d * d * sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        int c = cls[i];
        cnt[c]++;
        for (int j = 0; j < d; j++) row[j] = (double)X[(size_t)i * d + j] - mu[(size_t)c * d + j];
        double *S = cov + (size_t)c * d * d;
        for (int a = 0; a < d; a++)
            for (int b = 0; b < d; b++) S[(size_t)a * d + b] += row[a] * row[b];
    }
    for (int k = 0; k < K; k++) {
        double f = 1.0 / ((double)cnt[k] - 1.0);
        for (size_t q = 0; q < (size_t)d * d; q++) cov[(size_t)k * d * d + q] *= f;
        w[k] = (double)cnt[k] / (double)n;
    }
    free(cnt); free(row);
}

/* ------------------------------------------------------------------------ */
/* Clust/KMeansInitializer.cpp:33-49 ForgyInitializer::Init: K rows drawn at
 * (long)(Rand()*N), in component order (repeats allowed).                   */
void orc_forgy(int64_t n, int d, int K, const float *X, orc_mt64 *rng, double *C) {
    for (int k = 0; k < K; k++) {
        int64_t r = (int64_t)(orc_mt64_canonical(rng) * (double)n);
        for (int j = 0; j < d; j++) C[(size_t)k * d + j] = (double)X[(size_t)r * d + j];
    }
}

/* Util/MultithreadedDataset.cpp:179-283 ComputeStatistics -> TraceCov: column
 * means, then sum of (x - mean)^2 / (N - 1), summed over columns.  Sequential
 * ascending row order (the reference sums per thread, then over threads).   */
double orc_data_trace(int64_t n, int d, const float *X) {
    double *m = (double *)calloc((size_t)d, sizeof(double));
    double *s = (double *)calloc((size_t)d, sizeof(double));
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < d; j++) m[j] += (double)X[(size_t)i * d + j];
    for (int j = 0; j < d; j++) m[j] /= (double)n;
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < d; j++) {
            double t = (double)X[(size_t)i * d + j] - m[j];
            s[j] += t * t;
        }
    double tr = 0.0;
    for (int j = 0; j < d; j++) tr += s[j] / ((double)n - 1.0);
    free(m); free(s);
    return tr;
}

/* EM2/EMInitializer.cpp:43-80 JainEMInitializer::Init.  For component i one
 * index is drawn and kept; the duplicate check draws further indices into a
 * shadowed variable (:57) until one differs from the kept indices of earlier
 * components, and discards them.  Weights 1/K, means = kept rows,
 * covariances (trace / d / 10) I.                                           */
void orc_jain(int64_t n, int d, int K, const float *X, double trace, orc_mt64 *rng,
              double *w, double *mu, double *cov) {
    int64_t *obj = (int64_t *)malloc((size_t)K * sizeof(int64_t));
    const double wt = 1.0 / (double)K;
    const double cv = trace / (double)d / 10.0;
    memset(cov, 0, (size_t)K * d * d * sizeof(double));
    for (int i = 0; i < K; i++) {
        int64_t idx = (int64_t)(orc_mt64_canonical(rng) * (double)n);
        int found;
        do {
            int64_t idx2 = (int64_t)(orc_mt64_canonical(rng) * (double)n);
            found = 0;
            for (int j = 0; j < i; j++)
                if (obj[j] == idx2) { found = 1; break; }
        } while (found);
        obj[i] = idx;
        w[i] = wt;
        for (int j = 0; j < d; j++) mu[(size_t)i * d + j] = (double)X[(size_t)idx * d + j];
        for (int j = 0; j < d; j++) cov[(size_t)i * d * d + (size_t)j * d + j] = cv;
    }
    free(obj);
}

/* Clust/CentroidVector.h:79-87: squared distance of two float rows; each
 * difference and square in float, accumulated in double, ascending order.   */
static inline double sqdist_ff(int d, const float *a, const float *b) {
    double s = 0.0;
    for (int j = 0; j < d; j++) {
        float t = b[j] - a[j];
        float t2 = t * t;
        s += (double)t2;
    }
    return s;
}

/* Clust/KMeansOrOrInitializer.cpp:15-174 (k-means||, oversampling factor os,
 * 5 rounds for os > 0.1 else 10) followed by the weighted k-means++ of
 * Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:23-84 over the collected
 * centers.  Restated in the order of a one-thread run: one Rand() per row per
 * round in row order (the reference draws from one shared stream inside an
 * OpenMP loop, so its multi-thread result is not reproducible), new centers
 * in row order, ties keep the earlier center.  Returns the number of
 * collected centers (its candidate set), or -1 on allocation failure.       */
int64_t orc_kmeansoror(int64_t n, int d, int K, const float *X, double os, orc_mt64 *rng,
                       double *C) {
    const int rounds = os <= 0.1 ? 10 : 5;
    double *cd = (double *)malloc((size_t)n * sizeof(double));
    int *cc = (int *)malloc((size_t)n * sizeof(int));
    int64_t cap = 1024, m = 0;
    int64_t *rows = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    if (!cd || !cc || !rows) { free(cd); free(cc); free(rows); return -1; }
    for (int64_t i = 0; i < n; i++) { cd[i] = INFINITY; cc[i] = -1; }
    /* first center and InsertCentroid (:38-51) */
    rows[m++] = (int64_t)(orc_mt64_canonical(rng) * (double)n);
    for (int64_t i = 0; i < n; i++) {
        double s = sqdist_ff(d, X + (size_t)rows[0] * d, X + (size_t)i * d);
        if (s < cd[i]) { cd[i] = s; cc[i] = 0; }
    }
    double cost = 0.0;
    for (int64_t i = 0; i < n; i++) cost += cd[i];
    for (int r = 0; r < rounds; r++) {
        /* ScanDataset (:83-103) */
        int64_t base = m;
        for (int64_t i = 0; i < n; i++) {
            double prob = os * (double)K * cd[i] / cost;
            if (orc_mt64_canonical(rng) < prob) {
                if (m == cap) {
                    cap *= 2;
                    int64_t *nr = (int64_t *)realloc(rows, (size_t)cap * sizeof(int64_t));
                    if (!nr) { free(cd); free(cc); free(rows); return -1; }
                    rows = nr;
                }
                rows[m++] = i;
            }
        }
        /* InsertCentroids (:53-70) */
        for (int64_t i = 0; i < n; i++) {
            const float *x = X + (size_t)i * d;
            for (int64_t j = base; j < m; j++) {
                double s = sqdist_ff(d, X + (size_t)rows[j] * d, x);
                if (s < cd[i]) { cd[i] = s; cc[i] = (int)j; }
            }
        }
        cost = 0.0;
        for (int64_t i = 0; i < n; i++) cost += cd[i];
    }
    /* ComputeWeights (:114-126) */
    double *wt = (double *)calloc((size_t)m, sizeof(double));
    for (int64_t i = 0; i < n; i++) wt[cc[i]] += 1.0;
    /* weighted k-means++ over the m centers (PlusPlusInitializer.cpp:33-84) */
    double *cd2 = (double *)malloc((size_t)m * sizeof(double));
    double *agg = (double *)malloc((size_t)m * sizeof(double));
    int64_t r0 = (int64_t)(orc_mt64_canonical(rng) * (double)m);
    for (int j = 0; j < d; j++) C[j] = (double)X[(size_t)rows[r0] * d + j];
    for (int64_t j = 0; j < m; j++) cd2[j] = INFINITY;
    for (int c = 1; c < K; c++) {
        const double *prevc = C + (size_t)(c - 1) * d;
        for (int64_t j = 0; j < m; j++) {
            double s = sqdist(d, prevc, X + (size_t)rows[j] * d);
            if (cd2[j] > s) cd2[j] = s;
        }
        double sum = 0.0;
        for (int64_t j = 0; j < m; j++) { sum += cd2[j] * wt[j]; agg[j] = sum; }
        for (int64_t j = 0; j < m; j++) agg[j] /= sum;
        double rr = orc_mt64_canonical(rng);
        for (int64_t j = 0; j < m; j++) {
            if (rr <= agg[j]) {
                for (int q = 0; q < d; q++) C[(size_t)c * d + q] = (double)X[(size_t)rows[j] * d + q];
                break;
            }
        }
    }
    free(cd); free(cc); free(rows); free(wt); free(cd2); free(agg);
    return m;
}
