// setnet_init.cpp -- the learned initializer as one C entry per use
// (include/fastem_setnet.h): raw weight files, the proposal (subsample,
// normalise, BF16 forward, seeding over the point head's moved points, then
// the device Lloyd sweeps and cluster mixture of fem_kmeans_init_seeded_rng,
// k-means++ as the fallback) and the judged form that runs EM from the
// learned and the k-means++ starts and keeps the higher final
// log-likelihood.  No counterpart in the reference, whose initializers are
// EM2/EMInitializer.cpp's.
//
// All draws come from one caller-supplied Rand() stream, in this order: n_sub
// row indices, the k-means++ draws over the moved points, then the Lloyd
// repair draws (and, for the fallback or the judged form, the k-means++
// initializer's draws), so a seed fixes the whole initialisation.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "../../../include/fastem_setnet.h"
#include "../fem_internal.h"

#pragma clang fp contract(off)

namespace {

const char kMagic[8] = {'F', 'E', 'M', 'S', 'N', 'E', 'T', '1'};

inline double canonical(std::mt19937_64 &g) {
    const double r = (double)g() / 18446744073709551616.0;
    return r < 1.0 ? r : std::nextafter(1.0, 0.0);
}

double mt_rand(void *p) { return canonical(*static_cast<std::mt19937_64 *>(p)); }

int ctx_fail(fem_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    else fem::fail_global(code, m);
    return code;
}

// Starting centroids (K x d) proposed by the network for the context's data.
int propose_means(fem_setnet *net, fem_ctx *c, int n_sub, double (*rnd)(void *), void *st, double *C0) {
    int d = 0, K = 0;
    if (!net || !c || !rnd || !C0) return ctx_fail(c, FEM_ERR_ARG, "learned initializer: null argument");
    fem_setnet_shape(net, &d, &K);
    if (d != c->d || K != c->K)
        return ctx_fail(c, FEM_ERR_ARG, "learned initializer: the network's d and K differ from the data's");
    if (n_sub < K) return ctx_fail(c, FEM_ERR_ARG, "learned initializer: subsample smaller than K");
    // subsample: n_sub row indices (long)(Rand() * N), with replacement
    std::vector<int64_t> idx(n_sub);
    for (int j = 0; j < n_sub; j++) idx[j] = (int64_t)(rnd(st) * (double)c->n);
    std::vector<float> Xs((size_t)n_sub * d);
    if (int rc = fem::gather_rows_host(c, idx.data(), n_sub, Xs.data())) return rc;
    // normalise: column means (sequential sums), one scale = root mean
    // square deviation over all coordinates (rows outer, columns inner)
    std::vector<double> cen(d, 0.0);
    for (int i = 0; i < n_sub; i++)
        for (int j = 0; j < d; j++) cen[j] += (double)Xs[(size_t)i * d + j];
    for (int j = 0; j < d; j++) cen[j] /= (double)n_sub;
    double ss = 0.0;
    for (int i = 0; i < n_sub; i++)
        for (int j = 0; j < d; j++) {
            const double t = (double)Xs[(size_t)i * d + j] - cen[j];
            ss += t * t;
        }
    double sc = std::sqrt(ss / ((double)n_sub * (double)d));
    if (!(sc > 0.0)) sc = 1.0;
    std::vector<float> Z((size_t)n_sub * d);
    for (int i = 0; i < n_sub; i++)
        for (int j = 0; j < d; j++) Z[(size_t)i * d + j] = (float)(((double)Xs[(size_t)i * d + j] - cen[j]) / sc);
    // BF16-MFMA forward; the point head moves every point towards its
    // cluster's mean
    std::vector<float> qmu((size_t)K * d), qlv((size_t)K * d), m((size_t)n_sub * d);
    if (int rc = fem_setnet_forward_bf16(net, 1, n_sub, Z.data(), qmu.data(), qlv.data(), m.data()))
        return ctx_fail(c, rc, std::string("learned initializer: ") + fem_setnet_last_error(net));
    // k-means++ seeding over the moved points
    // (Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84 rule: first row
    // uniform, then the first row whose normalised prefix sum of D^2 reaches
    // the draw); the moved points have collapsed onto their clusters, so the
    // seeds fall one per cluster
    std::vector<double> P(m.begin(), m.end()), C((size_t)K * d, 0.0);
    std::vector<double> cd(n_sub, std::numeric_limits<double>::infinity()), agg(n_sub);
    const int64_t r0 = (int64_t)(rnd(st) * (double)n_sub);
    for (int j = 0; j < d; j++) C[j] = P[(size_t)r0 * d + j];
    for (int k = 1; k < K; k++) {
        const double *pc = C.data() + (size_t)(k - 1) * d;
        for (int i = 0; i < n_sub; i++) {
            double s = 0.0;
            for (int j = 0; j < d; j++) {
                const double t = pc[j] - P[(size_t)i * d + j];
                s += t * t;
            }
            if (cd[i] > s) cd[i] = s;
        }
        double sum = 0.0;
        for (int i = 0; i < n_sub; i++) {
            sum += cd[i];
            agg[i] = sum;
        }
        for (int i = 0; i < n_sub; i++) agg[i] /= sum;
        const double r = rnd(st);
        for (int i = 0; i < n_sub; i++)
            if (r <= agg[i]) {
                for (int j = 0; j < d; j++) C[(size_t)k * d + j] = P[(size_t)i * d + j];
                break;
            }
    }
    for (int k = 0; k < K; k++)
        for (int j = 0; j < d; j++) C0[(size_t)k * d + j] = cen[j] + sc * C[(size_t)k * d + j];
    return FEM_OK;
}

int propose(fem_setnet *net, fem_ctx *c, int n_sub, double (*rnd)(void *), void *st, int max_sweeps, double *w,
            double *mu, double *cov, int *used) {
    if (!w || !mu || !cov) return ctx_fail(c, FEM_ERR_ARG, "learned initializer: null output");
    std::vector<double> C0((size_t)(c ? c->K * c->d : 0));
    if (int rc = propose_means(net, c, n_sub, rnd, st, C0.data())) return rc;
    int rc = fem_kmeans_init_seeded_rng(c, C0.data(), rnd, st, max_sweeps, w, mu, cov);
    if (rc == FEM_OK) {
        // a cluster of fewer than d + 1 points has a singular covariance:
        // without regularisation (the context's g) EM could not start from
        // this mixture, so k-means++ instead
        bool full_rank = true;
        for (int k = 0; k < c->K; k++)
            if (std::llround(w[k] * (double)c->n) < (long long)c->d + 1) full_rank = false;
        if (full_rank || c->g > 0.0) {
            if (used) *used = 1;
            return FEM_OK;
        }
    } else if (rc != FEM_ERR_ARG) {
        return rc;
    }
    // a cluster of the proposal kept fewer than two points, or (g = 0) fewer
    // than d + 1: k-means++
    if (used) *used = 0;
    return fem_kmeans_init_rng(c, rnd, st, FEM_SEED_KMEANSPP, max_sweeps, w, mu, cov);
}

// EM from (w, mu, cov) on the context to the reference's stopping rule
// (EM2/EMAlgorithm.cpp:139-191: regularisation before the loop and after each
// M-step, e = |dLL/LL|, 0 on a decrease, stop when e < eps or after
// maxsweeps E-steps).  On success the converged parameters replace w, mu,
// cov and *ll is the last E-step's log-likelihood.
int run_em(fem_ctx *c, double *w, double *mu, double *cov, int maxsweeps, double eps, double g, double *ll) {
    const int K = c->K, d = c->d;
    if (int rc = fem_set_regularization(c, g)) return rc;
    std::vector<double> creg(cov, cov + (size_t)K * d * d);
    for (int k = 0; k < K; k++) {
        double *S = creg.data() + (size_t)k * d * d;
        double tr = 0.0;
        for (int i = 0; i < d; i++) tr += S[(size_t)i * d + i];
        const double diag = g * tr / (double)d;
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) S[(size_t)i * d + j] = S[(size_t)i * d + j] * (1.0 - g) + (i == j ? diag : 0.0);
    }
    if (int rc = fem_set_params(c, w, mu, creg.data())) return rc;
    if (int rc = fem_factor(c, nullptr, nullptr)) return rc;
    double prev = -std::numeric_limits<double>::max(), cur = prev;
    for (int it = 1; it <= maxsweeps; it++) {
        if (int rc = fem_estep(c, &cur)) return rc;
        if (int rc = fem_mstep(c, nullptr, nullptr, nullptr)) return rc;
        if (int rc = fem_factor(c, nullptr, nullptr)) return rc;
        double e = std::fabs((cur - prev) / cur);
        if (cur < prev) e = 0.0;
        prev = cur;
        if (e < eps) break;
    }
    *ll = cur;
    return fem_get_params(c, w, mu, cov);
}

}  // namespace

extern "C" {

int fem_setnet_save(fem_setnet *net, const char *path) {
    if (!net || !path) return FEM_ERR_ARG;
    int d = 0, K = 0;
    fem_setnet_shape(net, &d, &K);
    const int np = fem_setnet_num_params(net);
    std::vector<float> th(np);
    if (int rc = fem_setnet_get_params(net, th.data())) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return fem::fail_global(FEM_ERR_ARG, std::string("cannot write ") + path);
    const int32_t hdr[3] = {d, K, np};
    const bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 3, f) == 3 &&
                    std::fwrite(th.data(), 4, (size_t)np, f) == (size_t)np;
    const bool closed = std::fclose(f) == 0;
    return ok && closed ? FEM_OK : fem::fail_global(FEM_ERR_ARG, std::string("short write of ") + path);
}

int fem_setnet_load(fem_setnet **out, const char *path, int device) {
    if (!out || !path) return fem::fail_global(FEM_ERR_ARG, "fem_setnet_load: null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fem::fail_global(FEM_ERR_ARG, std::string("cannot open network file ") + path);
    char magic[8];
    int32_t hdr[3];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0 || std::fread(hdr, 4, 3, f) != 3 ||
        hdr[0] < 1 || hdr[1] < 1 || hdr[2] < 1) {
        std::fclose(f);
        return fem::fail_global(FEM_ERR_ARG, std::string("not a set-network file: ") + path);
    }
    std::vector<float> th((size_t)hdr[2]);
    const size_t got = std::fread(th.data(), 4, th.size(), f);
    std::fclose(f);
    if (got != th.size()) return fem::fail_global(FEM_ERR_ARG, std::string("truncated network file ") + path);
    fem_setnet *net = nullptr;
    if (int rc = fem_setnet_create(&net, hdr[0], hdr[1], 0, device))
        return fem::fail_global(rc, "fem_setnet_load: unsupported network shape");
    if (fem_setnet_num_params(net) != hdr[2]) {
        fem_setnet_destroy(net);
        return fem::fail_global(FEM_ERR_ARG, std::string("parameter count mismatch in ") + path);
    }
    if (int rc = fem_setnet_set_params(net, th.data())) {
        const std::string m = fem_setnet_last_error(net);
        fem_setnet_destroy(net);
        return fem::fail_global(rc, m);
    }
    *out = net;
    return FEM_OK;
}

int fem_setnet_propose_means_rng(fem_setnet *net, fem_ctx *ctx, int n_sub, double (*rnd)(void *), void *rnd_state,
                                 double *C0) {
    return propose_means(net, ctx, n_sub, rnd, rnd_state, C0);
}

int fem_setnet_propose_means(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, double *C0) {
    std::mt19937_64 rng(seed);
    return propose_means(net, ctx, n_sub, mt_rand, &rng, C0);
}

int fem_setnet_propose_rng(fem_setnet *net, fem_ctx *ctx, int n_sub, double (*rnd)(void *), void *rnd_state,
                           int max_sweeps, double *w, double *mu, double *cov, int *used) {
    return propose(net, ctx, n_sub, rnd, rnd_state, max_sweeps, w, mu, cov, used);
}

int fem_setnet_propose(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, int max_sweeps, double *w,
                       double *mu, double *cov, int *used) {
    std::mt19937_64 rng(seed);
    return propose(net, ctx, n_sub, mt_rand, &rng, max_sweeps, w, mu, cov, used);
}

int fem_setnet_judge_rng(fem_setnet *net, fem_ctx *c, int n_sub, double (*rnd)(void *), void *rnd_state,
                         int max_sweeps, int em_sweeps, double eps, double g, double *w, double *mu, double *cov,
                         int *used, double *ll) {
    if (!c || !w || !mu || !cov || em_sweeps < 1 || !(eps > 0.0) || !(g >= 0.0 && g <= 1.0))
        return ctx_fail(c, FEM_ERR_ARG, "fem_setnet_judge: bad argument");
    const size_t K = c->K, d = c->d;
    std::vector<double> wl(K), ml(K * d), cl(K * d * d), wk(K), mk(K * d), ck(K * d * d);
    const double lost = -std::numeric_limits<double>::infinity();
    double lll = lost, llk = lost;
    int ul = 0;
    // the EM below regularises with g: the proposal may keep rank-deficient
    // clusters then
    if (int r = fem_set_regularization(c, g)) return r;
    int rc = propose(net, c, n_sub, rnd, rnd_state, max_sweeps, wl.data(), ml.data(), cl.data(), &ul);
    if (rc != FEM_OK && rc != FEM_ERR_ARG) return rc;
    const bool have_l = rc == FEM_OK;
    rc = fem_kmeans_init_rng(c, rnd, rnd_state, FEM_SEED_KMEANSPP, max_sweeps, wk.data(), mk.data(), ck.data());
    if (rc != FEM_OK && rc != FEM_ERR_ARG) return rc;
    const bool have_k = rc == FEM_OK;
    // EM from each start; a start whose EM fails (not positive definite)
    // loses; other failures end the call
    if (have_l) {
        rc = run_em(c, wl.data(), ml.data(), cl.data(), em_sweeps, eps, g, &lll);
        if (rc != FEM_OK && rc != FEM_ERR_NOT_PD) return rc;
        if (rc != FEM_OK) lll = lost;
    }
    if (have_k) {
        rc = run_em(c, wk.data(), mk.data(), ck.data(), em_sweeps, eps, g, &llk);
        if (rc != FEM_OK && rc != FEM_ERR_NOT_PD) return rc;
        if (rc != FEM_OK) llk = lost;
    }
    if (ll) {
        ll[0] = lll;
        ll[1] = llk;
    }
    if (lll == lost && llk == lost) return ctx_fail(c, FEM_ERR_NOT_PD, "fem_setnet_judge: EM failed from both starts");
    const bool learned = lll >= llk;
    std::memcpy(w, learned ? wl.data() : wk.data(), sizeof(double) * K);
    std::memcpy(mu, learned ? ml.data() : mk.data(), sizeof(double) * K * d);
    std::memcpy(cov, learned ? cl.data() : ck.data(), sizeof(double) * K * d * d);
    if (used) *used = learned ? (ul ? 1 : 0) : 2;
    return fem_set_params(c, w, mu, cov);
}

int fem_setnet_judge(fem_setnet *net, fem_ctx *ctx, int n_sub, uint64_t seed, int max_sweeps, int em_sweeps,
                     double eps, double g, double *w, double *mu, double *cov, int *used, double *ll) {
    std::mt19937_64 rng(seed);
    return fem_setnet_judge_rng(net, ctx, n_sub, mt_rand, &rng, max_sweeps, em_sweeps, eps, g, w, mu, cov, used,
                                ll);
}

}  // extern "C"
