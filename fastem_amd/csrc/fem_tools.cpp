// fem_tools.cpp -- synthetic mixture generator and the k-means++ initializer
// entry on host data (include/fastem_tools.h).  Host code, OpenMP.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "../../include/fastem/EMException.h"
#include "../../include/fastem/NormalDensity.h"
#include "../../include/fastem_abi.h"
#include "../../include/fastem_tools.h"

namespace {

// libstdc++ generate_canonical<double,53> over a 64-bit engine, which is what
// the reference's Rand() (Util/Rand.cpp:7-14) evaluates to.
inline double canonical(std::mt19937_64 &g) {
    double r = (double)g() / 18446744073709551616.0;
    return r < 1.0 ? r : std::nextafter(1.0, 0.0);
}

struct BoxMuller {
    std::mt19937_64 &g;
    bool have = false;
    double spare = 0.0;
    explicit BoxMuller(std::mt19937_64 &e) : g(e) {}
    double operator()() {
        if (have) { have = false; return spare; }
        const double u1 = canonical(g), u2 = canonical(g);
        const double r = std::sqrt(-2.0 * std::log1p(-u1));
        const double t = 2.0 * M_PI * u2;
        spare = r * std::sin(t);
        have = true;
        return r * std::cos(t);
    }
};

bool chol_lower(int d, const double *C, double *L) {
    std::memset(L, 0, sizeof(double) * d * d);
    for (int j = 0; j < d; j++) {
        double s = C[j * d + j];
        for (int k = 0; k < j; k++) s -= L[j * d + k] * L[j * d + k];
        if (!(s > 0.0)) return false;
        L[j * d + j] = std::sqrt(s);
        for (int i = j + 1; i < d; i++) {
            double t = C[i * d + j];
            for (int k = 0; k < j; k++) t -= L[i * d + k] * L[j * d + k];
            L[i * d + j] = t / L[j * d + j];
        }
    }
    return true;
}

}  // namespace

extern "C" int fem_gengmm(int64_t n, int d, int K, double mean_lo, double mean_hi, double diag,
                          uint64_t seed, float *X, int *labels, double *w_out, double *mu_out,
                          double *cov_out) {
    if (n < 1 || d < 1 || K < 1 || !X) return -1;
    std::mt19937_64 master(seed);
    BoxMuller nrm(master);
    std::vector<double> w(K), mu((size_t)K * d), cov((size_t)K * d * d), L((size_t)K * d * d);
    double ws = 0.0;
    for (int k = 0; k < K; k++) {
        w[k] = -std::log1p(-canonical(master));
        ws += w[k];
    }
    for (int k = 0; k < K; k++) w[k] /= ws;
    for (size_t q = 0; q < mu.size(); q++) mu[q] = mean_lo + (mean_hi - mean_lo) * canonical(master);
    std::vector<double> A((size_t)d * d);
    for (int k = 0; k < K; k++) {
        for (auto &a : A) a = nrm();
        double *C = cov.data() + (size_t)k * d * d;
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) {
                double s = 0.0;
                for (int q = 0; q < d; q++) s += A[(size_t)i * d + q] * A[(size_t)j * d + q];
                C[(size_t)i * d + j] = s / d + (i == j ? diag : 0.0);
            }
        if (!chol_lower(d, C, L.data() + (size_t)k * d * d)) return -2;
    }
    std::vector<double> cdf(K);
    double acc = 0.0;
    for (int k = 0; k < K; k++) cdf[k] = (acc += w[k]);
    cdf[K - 1] = 1.0;
    const int64_t kBlock = 8192;
    const int64_t nblocks = (n + kBlock - 1) / kBlock;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nblocks; b++) {
        std::mt19937_64 g(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(b + 1));
        BoxMuller bm(g);
        std::vector<double> z(d);
        const int64_t i1 = (b + 1) * kBlock < n ? (b + 1) * kBlock : n;
        for (int64_t i = b * kBlock; i < i1; i++) {
            const double u = canonical(g);
            int lab = 0;
            while (lab < K - 1 && !(u < cdf[lab])) lab++;
            bm.have = false;
            for (int j = 0; j < d; j++) z[j] = bm();
            const double *Lk = L.data() + (size_t)lab * d * d;
            const double *m = mu.data() + (size_t)lab * d;
            for (int j = 0; j < d; j++) {
                double s = m[j];
                for (int q = 0; q <= j; q++) s += Lk[(size_t)j * d + q] * z[q];
                X[(size_t)i * d + j] = (float)s;
            }
            if (labels) labels[i] = lab;
        }
    }
    if (w_out) std::memcpy(w_out, w.data(), sizeof(double) * K);
    if (mu_out) std::memcpy(mu_out, mu.data(), sizeof(double) * K * d);
    if (cov_out) std::memcpy(cov_out, cov.data(), sizeof(double) * K * d * d);
    return 0;
}

// k-means++ initializer on host data: a temporary FP64 context runs the
// device k-means (fem_kmeans_init) on a copy of X.
extern "C" int fem_kmeanspp_init(int64_t n, int d, int K, const float *X, uint64_t seed,
                                 int lloyd_iters, double *w, double *mu, double *cov) {
    if (n < 1 || d < 1 || K < 1 || !X || !w || !mu || !cov || lloyd_iters < 0) return FEM_ERR_ARG;
    fem_ctx *c = nullptr;
    int rc = fem_create(&c, n, d, K, FEM_PREC_FP64, 0);
    if (rc != FEM_OK) return rc;
    rc = fem_upload_data(c, X, d);
    if (rc == FEM_OK) rc = fem_kmeans_init(c, seed, 1, lloyd_iters, w, mu, cov);
    fem_destroy(c);
    return rc;
}

namespace {

void load_density(NormalDensity &D, int d, const double *mu, const double *C) {
    EigenMatrix M(d, d);
    std::memcpy(M.data(), C, sizeof(double) * d * d);
    D.Init(M, EigenRowVector(mu, d));
}

}  // namespace

extern "C" int fem_condition_number(int d, const double *C, double *cond) {
    if (d < 1 || !C || !cond) return FEM_ERR_ARG;
    std::vector<double> zero(d, 0.0);
    NormalDensity D(d);
    try {
        load_density(D, d, zero.data(), C);
    } catch (EMException &) {
        return FEM_ERR_NOT_PD;
    }
    try {
        *cond = D.ComputeConditionNumber();
    } catch (EMException &) {
        return FEM_ERR_ARG;
    }
    return FEM_OK;
}

extern "C" int fem_bhatt_dist(int d, const double *mu1, const double *C1, const double *mu2, const double *C2,
                              double *dist) {
    if (d < 1 || !mu1 || !C1 || !mu2 || !C2 || !dist) return FEM_ERR_ARG;
    NormalDensity A(d), B(d);
    BhattDist Bh(d);
    try {
        load_density(A, d, mu1, C1);
        load_density(B, d, mu2, C2);
        *dist = Bh.ComputeDist(A, B);
    } catch (EMException &) {
        return FEM_ERR_NOT_PD;
    }
    return FEM_OK;
}
