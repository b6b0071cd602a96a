// fem_tools.cpp -- synthetic mixture generator and k-means++ initializer
// (include/fastem_tools.h).  Host code, OpenMP.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

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

inline double sqdist(int d, const double *c, const float *x) {
    double s = 0.0;
    for (int j = 0; j < d; j++) {
        const double t = c[j] - (double)x[j];
        s += t * t;
    }
    return s;
}

constexpr int kParts = 64;  // fixed virtual partition for deterministic sums

inline void part_range(int64_t n, int p, int64_t &b, int64_t &e) {
    const int64_t q = n / kParts, r = n % kParts;
    b = p * q + (p < r ? p : r);
    e = b + q + (p < r ? 1 : 0);
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

namespace fem_host {

// Seeding (Forgy random partition, Clust/KMeansInitializer.cpp:33-49, or
// k-means++), up to max_sweeps Lloyd sweeps (Clust/KMAlgorithm.cpp:44-87 with
// MinRel 1e-4; the reference's MaxIter = max_sweeps + 1), then the mixture of
// the final clustering (EM2/EMInitializer.cpp:103-134).  rnd(rs) is the
// reference's Rand().
int kmeans_mixture(int64_t n, int d, int K, const float *X, double (*rnd)(void *), void *rs,
                   int seeding, int max_sweeps, double *w, double *mu, double *cov) {
    if (n < 1 || d < 1 || K < 1 || !X || !w || !mu || !cov) return -1;
    std::vector<double> C((size_t)K * d);
    if (seeding == 0) {
        std::vector<int64_t> cnt(K, 0);
        for (int64_t i = 0; i < n; i++) {
            const int part = (int)(rnd(rs) * K);
            cnt[part]++;
            for (int j = 0; j < d; j++) C[(size_t)part * d + j] += X[(size_t)i * d + j];
        }
        for (int k = 0; k < K; k++) {
            const double f = 1.0 / (double)cnt[k];
            for (int j = 0; j < d; j++) C[(size_t)k * d + j] *= f;
        }
    } else {
    std::vector<double> cd(n, std::numeric_limits<double>::infinity()), agg(n);
    // Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84
    const int64_t r0 = (int64_t)(rnd(rs) * (double)n);
    for (int j = 0; j < d; j++) C[j] = (double)X[(size_t)r0 * d + j];
    for (int c = 1; c < K; c++) {
        const double *pc = C.data() + (size_t)(c - 1) * d;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            const double s = sqdist(d, pc, X + (size_t)i * d);
            if (cd[i] > s) cd[i] = s;
        }
        double sum = 0.0;
        for (int64_t i = 0; i < n; i++) {
            sum += cd[i] * 1.0;
            agg[i] = sum;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) agg[i] /= sum;
        const double r = rnd(rs);
        for (int64_t i = 0; i < n; i++)
            if (r <= agg[i]) {
                for (int j = 0; j < d; j++) C[(size_t)c * d + j] = (double)X[(size_t)i * d + j];
                break;
            }
    }
    }
    std::vector<int> cls(n);
    std::vector<double> psum((size_t)kParts * K * d);
    std::vector<int64_t> pcnt((size_t)kParts * K);
    std::vector<double> pfit(kParts);
    auto assign = [&](bool accumulate) {
        std::fill(psum.begin(), psum.end(), 0.0);
        std::fill(pcnt.begin(), pcnt.end(), 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int p = 0; p < kParts; p++) {
            int64_t b, e;
            part_range(n, p, b, e);
            double f = 0.0;
            for (int64_t i = b; i < e; i++) {
                const float *x = X + (size_t)i * d;
                double best = std::numeric_limits<double>::max();
                int bj = -1;
                for (int k = 0; k < K; k++) {
                    const double s = sqdist(d, C.data() + (size_t)k * d, x);
                    if (s < best) { best = s; bj = k; }
                }
                cls[i] = bj;
                f += best;
                if (accumulate) {
                    pcnt[(size_t)p * K + bj]++;
                    double *ce = psum.data() + ((size_t)p * K + bj) * d;
                    for (int j = 0; j < d; j++) ce[j] += (double)x[j];
                }
            }
            pfit[p] = f;
        }
        double fit = 0.0;
        for (int p = 0; p < kParts; p++) fit += pfit[p];
        return fit / (double)n;
    };
    // Clust/KMAlgorithm.cpp:44-87,205-255 (empty clusters repaired from Rand())
    double best = std::numeric_limits<double>::max();
    for (int it = 1; it <= max_sweeps; it++) {
        const double fit = assign(true);
        for (int p = 1; p < kParts; p++) {
            for (size_t q = 0; q < (size_t)K * d; q++) psum[q] += psum[(size_t)p * K * d + q];
            for (int k = 0; k < K; k++) pcnt[k] += pcnt[(size_t)p * K + k];
        }
        for (int k = 0; k < K; k++) {
            if (pcnt[k] > 0) {
                const double f = 1.0 / (double)pcnt[k];
                for (int j = 0; j < d; j++) C[(size_t)k * d + j] = psum[(size_t)k * d + j] * f;
            } else {
                const int64_t src = (int64_t)(rnd(rs) * (double)n);
                for (int j = 0; j < d; j++) C[(size_t)k * d + j] = (double)X[(size_t)src * d + j];
            }
        }
        const double rel = (best - fit) / best;
        if (fit < std::nextafter(fit, best)) best = fit;
        else break;
        if (rel < 1e-4) break;
        if (it == max_sweeps) break;
    }
    assign(false);
    // EM2/EMInitializer.cpp:103-134
    std::vector<double> pc((size_t)kParts * K * d * d, 0.0);
    std::fill(pcnt.begin(), pcnt.end(), 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < kParts; p++) {
        int64_t b, e;
        part_range(n, p, b, e);
        std::vector<double> row(d);
        for (int64_t i = b; i < e; i++) {
            const int c = cls[i];
            pcnt[(size_t)p * K + c]++;
            for (int j = 0; j < d; j++) row[j] = (double)X[(size_t)i * d + j] - C[(size_t)c * d + j];
            double *S = pc.data() + ((size_t)p * K + c) * d * d;
            for (int a = 0; a < d; a++)
                for (int bb = 0; bb < d; bb++) S[(size_t)a * d + bb] += row[a] * row[bb];
        }
    }
    for (int p = 1; p < kParts; p++) {
        for (size_t q = 0; q < (size_t)K * d * d; q++) pc[q] += pc[(size_t)p * K * d * d + q];
        for (int k = 0; k < K; k++) pcnt[k] += pcnt[(size_t)p * K + k];
    }
    int rc = 0;
    for (int k = 0; k < K; k++) {
        if (pcnt[k] < 2) rc = -1;
        const double f = 1.0 / ((double)pcnt[k] - 1.0);
        for (size_t q = 0; q < (size_t)d * d; q++) cov[(size_t)k * d * d + q] = pc[(size_t)k * d * d + q] * f;
        w[k] = (double)pcnt[k] / (double)n;
        for (int j = 0; j < d; j++) mu[(size_t)k * d + j] = C[(size_t)k * d + j];
    }
    return rc;
}

}  // namespace fem_host

static double mt_canonical(void *p) { return canonical(*static_cast<std::mt19937_64 *>(p)); }

extern "C" int fem_kmeanspp_init(int64_t n, int d, int K, const float *X, uint64_t seed,
                                 int lloyd_iters, double *w, double *mu, double *cov) {
    std::mt19937_64 rng(seed);
    return fem_host::kmeans_mixture(n, d, K, X, mt_canonical, &rng, 1, lloyd_iters, w, mu, cov);
}
