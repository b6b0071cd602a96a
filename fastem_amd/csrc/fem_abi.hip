// fem_abi.hip -- the C ABI of libfastem (include/fastem_abi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fem_internal.h"

namespace fem {

namespace {
constexpr size_t kStageBytes = 32u << 20;

struct Staging {
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Staging() {
        for (int i = 0; i < 2; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    hipError_t init(size_t bytes) {
        for (int i = 0; i < 2; i++) {
            if (hipError_t e = hipHostMalloc((void **)&buf[i], bytes, hipHostMallocDefault)) return e;
            if (hipError_t e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming)) return e;
        }
        return hipSuccess;
    }
};
}  // namespace

hipError_t copy_rows_h2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, int64_t rows,
                         hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (spitch == width && dpitch == width) {
        if (hipError_t e = hipMemcpyAsync(dst, src, width * (size_t)rows, hipMemcpyHostToDevice, s)) return e;
        return hipStreamSynchronize(s);
    }
    const int64_t per = std::max<int64_t>(1, (int64_t)(kStageBytes / dpitch));
    Staging st;
    if (hipError_t e = st.init((size_t)std::min(per, rows) * dpitch)) return e;
    const char *in = static_cast<const char *>(src);
    char *out = static_cast<char *>(dst);
    int it = 0;
    for (int64_t r0 = 0; r0 < rows; r0 += per, it++) {
        const int b = it & 1;
        if (it >= 2)
            if (hipError_t e = hipEventSynchronize(st.ev[b])) return e;
        const int64_t nr = std::min(per, rows - r0);
        for (int64_t r = 0; r < nr; r++) {
            char *row = st.buf[b] + (size_t)r * dpitch;
            std::memcpy(row, in + (size_t)(r0 + r) * spitch, width);
            if (dpitch > width) std::memset(row + width, 0, dpitch - width);
        }
        if (hipError_t e = hipMemcpyAsync(out + (size_t)r0 * dpitch, st.buf[b], (size_t)nr * dpitch,
                                          hipMemcpyHostToDevice, s))
            return e;
        if (hipError_t e = hipEventRecord(st.ev[b], s)) return e;
    }
    return hipStreamSynchronize(s);
}

hipError_t copy_rows_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, int64_t rows,
                         hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (spitch == width && dpitch == width) {
        if (hipError_t e = hipMemcpyAsync(dst, src, width * (size_t)rows, hipMemcpyDeviceToHost, s)) return e;
        return hipStreamSynchronize(s);
    }
    const int64_t per = std::max<int64_t>(1, (int64_t)(kStageBytes / spitch));
    Staging st;
    if (hipError_t e = st.init((size_t)std::min(per, rows) * spitch)) return e;
    const char *in = static_cast<const char *>(src);
    char *out = static_cast<char *>(dst);
    auto unpack = [&](int b, int64_t r0) {
        const int64_t nr = std::min(per, rows - r0);
        for (int64_t r = 0; r < nr; r++)
            std::memcpy(out + (size_t)(r0 + r) * dpitch, st.buf[b] + (size_t)r * spitch, width);
    };
    int it = 0;
    int64_t prev = -1;
    for (int64_t r0 = 0; r0 < rows; r0 += per, it++) {
        const int b = it & 1;
        const int64_t nr = std::min(per, rows - r0);
        if (hipError_t e = hipMemcpyAsync(st.buf[b], in + (size_t)r0 * spitch, (size_t)nr * spitch,
                                          hipMemcpyDeviceToHost, s))
            return e;
        if (hipError_t e = hipEventRecord(st.ev[b], s)) return e;
        if (prev >= 0) {
            if (hipError_t e = hipEventSynchronize(st.ev[b ^ 1])) return e;
            unpack(b ^ 1, prev);
        }
        prev = r0;
        // the next copy reuses buffer b^1 only after it was unpacked above
    }
    if (hipError_t e = hipEventSynchronize(st.ev[(it - 1) & 1])) return e;
    unpack((it - 1) & 1, prev);
    return hipStreamSynchronize(s);
}

}  // namespace fem

namespace {

thread_local std::string g_create_err;

int fail(fem_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

#define FEM_HIP(c, call)                                                              \
    do {                                                                              \
        hipError_t _e = (call);                                                       \
        if (_e != hipSuccess)                                                         \
            return fail((c), FEM_ERR_HIP,                                             \
                        std::string(#call) + ": " + hipGetErrorString(_e));           \
    } while (0)

template <typename T>
hipError_t dmalloc(T **p, size_t count) {
    return hipMalloc((void **)p, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
}

void free_all(fem_ctx *c) {
    void *ptrs[] = {c->X, c->Rt, c->Rt32, c->w, c->mu, c->cov, c->nk, c->logdet, c->info, c->coef,
                    c->U, c->stage64, c->stage32, c->stageb, c->stager, c->refc, c->ll_part, c->ll_dev,
                    c->mpart, c->cpart, c->kC, c->dmin, c->cls, c->kfit, c->kpick, c->kcand, c->kpot, c->korder,
                    c->Xs, c->perm, c->perm2, c->ref, c->ref2, c->slots, c->rel, c->sort_tmp, c->nskip};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_params) (void)hipEventDestroy(c->ev_params);
    if (c->params_host) (void)hipHostFree(c->params_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Buffers of the reference-centred BF16x3 E-step where it fits (unless
// FASTEM_ESTEPB_REF=0 keeps the per-component form throughout); the
// references start at component 0.
hipError_t alloc_ref(fem_ctx *c) {
    if (c->stager || !fem::estepr_fits(c->NB, c->K)) return hipSuccess;
    if (const char *v = std::getenv("FASTEM_ESTEPB_REF"))
        if (std::atoi(v) == 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)&c->stager, fem::estepr_words(c->NB, c->K) * sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->refc, (size_t)c->npad * sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->refc, 0, (size_t)c->npad * sizeof(int), c->stream);
    return e;
}

int check_ctx(fem_ctx *c) {
    if (!c) return fail(nullptr, FEM_ERR_ARG, "null context");
    return FEM_OK;
}

int set_device(fem_ctx *c) {
    FEM_HIP(c, hipSetDevice(c->device));
    return FEM_OK;
}

}  // namespace

int fem::fail_global(int code, const std::string &msg) { return fail(nullptr, code, msg); }

extern "C" {

const char *fem_version(void) { return "fastem-mi355x 0.1 (gfx950)"; }

const char *fem_last_error(const fem_ctx *c) {
    return c ? c->err.c_str() : g_create_err.c_str();
}

int fem_create(fem_ctx **out, int64_t n, int d, int K, int precision, int device) {
    if (!out) return fail(nullptr, FEM_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n < 1 || d < 1 || K < 1)
        return fail(nullptr, FEM_ERR_ARG, "n, d and K must be positive");
    if (d > fem::kMaxDim)
        return fail(nullptr, FEM_ERR_ARG, "dimension above 128 is not supported");
    if (precision != FEM_PREC_FP64 && precision != FEM_PREC_FP32 && precision != FEM_PREC_BF16X3)
        return fail(nullptr, FEM_ERR_ARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, FEM_ERR_HIP, "no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, FEM_ERR_ARG, "bad device index");

    fem_ctx *c = new fem_ctx();
    c->device = device;
    c->n = n;
    c->ntotal = n;
    c->d = d;
    c->NB = (d + 15) / 16;
    c->dpad = 16 * c->NB;
    c->K = K;
    c->KP = 16 * ((K + 15) / 16);
    c->prec = precision;
    c->npad = ((n + fem::kRowAlign - 1) / fem::kRowAlign) * fem::kRowAlign;
    const int ub = fem::upper_blocks(c->NB);
    c->est_blocks = (int)(c->npad / 64);  // >= workgroups of any E-step variant
    // FP64 and FP32 E-steps: four 16-point groups per wave measured faster
    // for d <= 32 (6-8% at d = 8, 16), two for larger d
    c->estep_groups = c->estep_groups32 = c->NB <= 2 ? 4 : 2;
    if (const char *e = std::getenv("FASTEM_ESTEP64_GROUPS")) c->estep_groups = std::atoi(e) == 4 ? 4 : 2;
    if (const char *e = std::getenv("FASTEM_ESTEP32_GROUPS")) c->estep_groups32 = std::atoi(e) == 4 ? 4 : 2;
    if (const char *e = std::getenv("FASTEM_ESTEPB_GROUPS")) c->estep_groupsb = std::atoi(e) == 2 ? 2 : 4;
    if (const char *e = std::getenv("FASTEM_COV_DENSE")) c->cov_sparse = std::atoi(e) == 0;
    if (const char *e = std::getenv("FASTEM_ESTEP_DENSE")) c->prune = std::atoi(e) == 0;
    c->prune = c->prune && K <= 64 && c->NB >= 2;
    c->m_chunks = (int)(c->npad / 4096);
    // covariance chunks: at most 256 chunks of a multiple of 4096 rows
    int64_t units = c->npad / 4096;
    int64_t per = (units + 255) / 256;
    c->c_chunk_rows = per * 4096;
    c->c_chunks = (int)((c->npad + c->c_chunk_rows - 1) / c->c_chunk_rows);

    auto bail = [&](hipError_t e, const char *what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        free_all(c);
        delete c;
        return fail(nullptr, FEM_ERR_HIP, m);
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    for (auto &ev : c->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(e, "hipEventCreate");
    const size_t nX = (size_t)c->npad * c->dpad;
    const size_t nR = (size_t)c->KP * c->npad;
    if ((e = dmalloc(&c->X, nX)) != hipSuccess) return bail(e, "hipMalloc(X)");
    const bool f64 = precision == FEM_PREC_FP64;
    if (f64 && (e = dmalloc(&c->Rt, nR)) != hipSuccess) return bail(e, "hipMalloc(R)");
    if (!f64 && (e = dmalloc(&c->Rt32, nR)) != hipSuccess) return bail(e, "hipMalloc(R32)");
    if (!f64 && (e = dmalloc(&c->stage32, (size_t)K * fem::stage32_words(c->NB))) != hipSuccess)
        return bail(e, "hipMalloc(fp32 factors)");
    if (precision == FEM_PREC_BF16X3 &&
        (e = dmalloc(&c->stageb, (size_t)K * fem::estepb_stage(c->NB))) != hipSuccess)
        return bail(e, "hipMalloc(bf16 factors)");
    if (precision == FEM_PREC_BF16X3 && (e = alloc_ref(c)) != hipSuccess) return bail(e, "hipMalloc(bf16 references)");
    if ((e = dmalloc(&c->w, K)) != hipSuccess ||
        (e = dmalloc(&c->mu, (size_t)K * d)) != hipSuccess ||
        (e = dmalloc(&c->cov, (size_t)K * d * d)) != hipSuccess ||
        (e = dmalloc(&c->nk, K)) != hipSuccess || (e = dmalloc(&c->logdet, K)) != hipSuccess ||
        (e = dmalloc(&c->info, K)) != hipSuccess || (e = dmalloc(&c->coef, K)) != hipSuccess ||
        (e = dmalloc(&c->korder, K)) != hipSuccess)
        return bail(e, "hipMalloc(params)");
    if ((e = dmalloc(&c->U, (size_t)K * c->dpad * c->dpad)) != hipSuccess ||
        (e = dmalloc(&c->stage64, (size_t)K * fem::stage64_words(c->NB))) != hipSuccess)
        return bail(e, "hipMalloc(factors)");
    if ((e = dmalloc(&c->ll_part, c->est_blocks)) != hipSuccess ||
        (e = dmalloc(&c->ll_dev, 1)) != hipSuccess ||
        (e = dmalloc(&c->mpart, (size_t)c->m_chunks * c->KP * (c->dpad + 1))) != hipSuccess ||
        (e = dmalloc(&c->cpart, (size_t)c->c_chunks * c->KP * ub * 256)) != hipSuccess)
        return bail(e, "hipMalloc(partials)");
    // the pruned FP64 E-step's per-slot components (none yet: -1) and its
    // skip count; the slot order's buffers come with the first sort
    if (c->prune && ((e = dmalloc(&c->ref, (size_t)c->npad)) != hipSuccess ||
                     (e = dmalloc(&c->nskip, 1)) != hipSuccess ||
                     (e = hipMemsetAsync(c->ref, 0xFF, (size_t)c->npad * sizeof(int), c->stream)) != hipSuccess))
        return bail(e, "hipMalloc(pruning)");
    // the E-step's component order: the index order until the first
    // factorisation ranks the weights
    std::vector<int> ident(K);
    for (int k = 0; k < K; k++) ident[k] = k;
    if ((e = hipMemsetAsync(c->X, 0, nX * sizeof(float), c->stream)) != hipSuccess ||
        (f64 && (e = hipMemsetAsync(c->Rt, 0, nR * sizeof(double), c->stream)) != hipSuccess) ||
        (!f64 && (e = hipMemsetAsync(c->Rt32, 0, nR * sizeof(float), c->stream)) != hipSuccess) ||
        (e = hipMemcpyAsync(c->korder, ident.data(), sizeof(int) * K, hipMemcpyHostToDevice, c->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return bail(e, "initial clear");
    *out = c;
    return FEM_OK;
}

int fem_destroy(fem_ctx *c) {
    if (!c) return FEM_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return FEM_OK;
}

static int upload(fem_ctx *c, const float *X, int64_t ld, hipMemcpyKind kind) {
    if (int r = check_ctx(c)) return r;
    if (!X || ld < c->d) return fail(c, FEM_ERR_ARG, "bad data pointer or row stride");
    if (int r = set_device(c)) return r;
    if (kind == hipMemcpyHostToDevice)
        FEM_HIP(c, fem::copy_rows_h2d(c->X, (size_t)c->dpad * sizeof(float), X, (size_t)ld * sizeof(float),
                                      (size_t)c->d * sizeof(float), c->n, c->stream));
    else
        FEM_HIP(c, hipMemcpy2DAsync(c->X, (size_t)c->dpad * sizeof(float), X, (size_t)ld * sizeof(float),
                                    (size_t)c->d * sizeof(float), (size_t)c->n, kind, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    c->ref_ok = false;
    fem::reset_order(c);
    return FEM_OK;
}

int fem_upload_data(fem_ctx *c, const float *X, int64_t ld) {
    return upload(c, X, ld, hipMemcpyHostToDevice);
}

int fem_upload_data_device(fem_ctx *c, const float *dX, int64_t ld) {
    return upload(c, dX, ld, hipMemcpyDeviceToDevice);
}

int fem_set_precision(fem_ctx *c, int precision) {
    if (int r = check_ctx(c)) return r;
    if (precision != FEM_PREC_FP64 && precision != FEM_PREC_FP32 && precision != FEM_PREC_BF16X3)
        return fail(c, FEM_ERR_ARG, "unknown precision");
    if (int r = set_device(c)) return r;
    const size_t nR = (size_t)c->KP * c->npad;
    if (precision == FEM_PREC_FP64 && !c->Rt) {
        FEM_HIP(c, hipMalloc((void **)&c->Rt, nR * sizeof(double)));
        FEM_HIP(c, hipMemsetAsync(c->Rt, 0, nR * sizeof(double), c->stream));
    }
    if (precision != FEM_PREC_FP64) {
        if (!c->Rt32) {
            FEM_HIP(c, hipMalloc((void **)&c->Rt32, nR * sizeof(float)));
            FEM_HIP(c, hipMemsetAsync(c->Rt32, 0, nR * sizeof(float), c->stream));
        }
        if (!c->stage32)
            FEM_HIP(c, hipMalloc((void **)&c->stage32, sizeof(float) * c->K * fem::stage32_words(c->NB)));
    }
    if (precision == FEM_PREC_BF16X3) {
        if (!c->stageb)
            FEM_HIP(c, hipMalloc((void **)&c->stageb, sizeof(unsigned int) * c->K * fem::estepb_stage(c->NB)));
        FEM_HIP(c, alloc_ref(c));
    }
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    c->prec = precision;
    c->factored = false;
    c->have_post = false;
    c->ref_ok = false;
    return FEM_OK;
}

int fem_set_total_rows(fem_ctx *c, int64_t ntotal) {
    if (int r = check_ctx(c)) return r;
    if (ntotal < 1) return fail(c, FEM_ERR_ARG, "ntotal must be positive");
    c->ntotal = ntotal;
    return FEM_OK;
}

int fem_set_regularization(fem_ctx *c, double g) {
    if (int r = check_ctx(c)) return r;
    if (!(g >= 0.0 && g <= 1.0)) return fail(c, FEM_ERR_ARG, "regularisation must be in [0,1]");
    c->g = g;
    return FEM_OK;
}

int fem_set_params(fem_ctx *c, const double *w, const double *mu, const double *cov) {
    if (int r = check_ctx(c)) return r;
    if (!w || !mu || !cov) return fail(c, FEM_ERR_ARG, "null parameter pointer");
    if (int r = set_device(c)) return r;
    const size_t K = c->K, d = c->d;
    FEM_HIP(c, hipMemcpyAsync(c->w, w, K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->mu, mu, K * d * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FEM_HIP(c, hipMemcpyAsync(c->cov, cov, K * d * d * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    c->have_params = true;
    c->factored = false;
    c->ref_ok = false;
    // the slot order stays (any order gives the same results); it is
    // re-sorted after the next pruned E-steps
    c->esteps_since_sort = 0;
    c->sort_every = 1;
    return FEM_OK;
}

int fem_get_params(fem_ctx *c, double *w, double *mu, double *cov) {
    if (int r = check_ctx(c)) return r;
    if (int r = set_device(c)) return r;
    const size_t K = c->K, d = c->d;
    if (w) FEM_HIP(c, hipMemcpyAsync(w, c->w, K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (mu)
        FEM_HIP(c, hipMemcpyAsync(mu, c->mu, K * d * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (cov)
        FEM_HIP(c, hipMemcpyAsync(cov, c->cov, K * d * d * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    return FEM_OK;
}

int fem_get_params_start(fem_ctx *c) {
    if (int r = check_ctx(c)) return r;
    if (int r = set_device(c)) return r;
    const size_t K = c->K, d = c->d;
    if (!c->params_host)
        FEM_HIP(c, hipHostMalloc((void **)&c->params_host, (K + K * d + K * d * d) * sizeof(double),
                                 hipHostMallocDefault));
    if (!c->ev_params) FEM_HIP(c, hipEventCreateWithFlags(&c->ev_params, hipEventDisableTiming));
    double *h = c->params_host;
    FEM_HIP(c, hipMemcpyAsync(h, c->w, K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipMemcpyAsync(h + K, c->mu, K * d * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipMemcpyAsync(h + K + K * d, c->cov, K * d * d * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    FEM_HIP(c, hipEventRecord(c->ev_params, c->stream));
    c->params_pending = true;
    return FEM_OK;
}

int fem_get_params_finish(fem_ctx *c, double *w, double *mu, double *cov) {
    if (int r = check_ctx(c)) return r;
    if (!c->params_pending) return fail(c, FEM_ERR_STATE, "fem_get_params_finish without fem_get_params_start");
    if (int r = set_device(c)) return r;
    FEM_HIP(c, hipEventSynchronize(c->ev_params));
    c->params_pending = false;
    const size_t K = c->K, d = c->d;
    const double *h = c->params_host;
    if (w) std::memcpy(w, h, K * sizeof(double));
    if (mu) std::memcpy(mu, h + K, K * d * sizeof(double));
    if (cov) std::memcpy(cov, h + K + K * d, K * d * d * sizeof(double));
    return FEM_OK;
}

int fem_factor(fem_ctx *c, double *logdet, int *info) {
    if (int r = check_ctx(c)) return r;
    if (!c->have_params) return fail(c, FEM_ERR_STATE, "fem_factor before fem_set_params");
    if (int r = set_device(c)) return r;
    FEM_HIP(c, hipEventRecord(c->ev[6], c->stream));
    FEM_HIP(c, fem::launch_factor(c));
    FEM_HIP(c, hipEventRecord(c->ev[7], c->stream));
    int *hinfo = info;
    std::string tmp;
    int *buf = nullptr;
    if (!hinfo) {
        buf = new int[c->K];
        hinfo = buf;
    }
    hipError_t e = hipMemcpyAsync(hinfo, c->info, c->K * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && logdet)
        e = hipMemcpyAsync(logdet, c->logdet, c->K * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        delete[] buf;
        return fail(c, FEM_ERR_HIP, std::string("fem_factor: ") + hipGetErrorString(e));
    }
    (void)hipEventElapsedTime(&c->t_factor, c->ev[6], c->ev[7]);
    int bad = -1;
    for (int k = 0; k < c->K; k++)
        if (hinfo[k] != 0) {
            bad = k;
            break;
        }
    if (bad >= 0) {
        char m[160];
        snprintf(m, sizeof m, "Cholesky decomposition failed for component %d (column %d)", bad,
                 hinfo[bad] - 1);
        delete[] buf;
        c->factored = false;
        return fail(c, FEM_ERR_NOT_PD, m);
    }
    delete[] buf;
    c->factored = true;
    return FEM_OK;
}

int fem_estep(fem_ctx *c, double *loglik) {
    if (int r = check_ctx(c)) return r;
    if (!c->factored) return fail(c, FEM_ERR_STATE, "fem_estep needs a successful fem_factor");
    if (int r = set_device(c)) return r;
    FEM_HIP(c, hipEventRecord(c->ev[0], c->stream));
    c->last_form = c->prec != FEM_PREC_BF16X3 ? c->prec : c->stage_ref ? FEM_ESTEP_BF16X3_REF : FEM_ESTEP_BF16X3;
    FEM_HIP(c, fem::launch_estep(c));
    FEM_HIP(c, hipEventRecord(c->ev[1], c->stream));
    double ll = 0.0;
    FEM_HIP(c, hipMemcpyAsync(&ll, c->ll_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    (void)hipEventElapsedTime(&c->t_estep, c->ev[0], c->ev[1]);
    if (loglik) *loglik = ll;
    c->have_post = true;
    return FEM_OK;
}

int fem_mstep(fem_ctx *c, double *w, double *mu, double *cov) {
    if (int r = check_ctx(c)) return r;
    if (!c->have_post) return fail(c, FEM_ERR_STATE, "fem_mstep needs posteriors from fem_estep");
    if (int r = set_device(c)) return r;
    FEM_HIP(c, hipEventRecord(c->ev[2], c->stream));
    FEM_HIP(c, fem::launch_means(c));
    FEM_HIP(c, hipEventRecord(c->ev[3], c->stream));
    FEM_HIP(c, fem::launch_cov(c));
    FEM_HIP(c, hipEventRecord(c->ev[4], c->stream));
    c->factored = false;
    if (int r = fem_get_params(c, w, mu, cov)) return r;  // synchronises
    (void)hipEventElapsedTime(&c->t_means, c->ev[2], c->ev[3]);
    (void)hipEventElapsedTime(&c->t_cov, c->ev[3], c->ev[4]);
    return FEM_OK;
}

int fem_get_posteriors(fem_ctx *c, double *R, int64_t ld) {
    if (int r = check_ctx(c)) return r;
    if (!R || ld < c->K) return fail(c, FEM_ERR_ARG, "bad posterior pointer or stride");
    if (!c->have_post) return fail(c, FEM_ERR_STATE, "no posteriors yet");
    if (int r = set_device(c)) return r;
    double *tmp = nullptr;
    FEM_HIP(c, hipMalloc((void **)&tmp, (size_t)c->n * c->K * sizeof(double)));
    hipError_t e = fem::launch_transpose_post(c, tmp, c->K);
    if (e == hipSuccess)
        e = fem::copy_rows_d2h(R, (size_t)ld * sizeof(double), tmp, (size_t)c->K * sizeof(double),
                               (size_t)c->K * sizeof(double), c->n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(c, FEM_ERR_HIP, std::string("fem_get_posteriors: ") + hipGetErrorString(e));
    return FEM_OK;
}

int fem_get_classes(fem_ctx *c, int *cls) {
    if (int r = check_ctx(c)) return r;
    if (!cls) return fail(c, FEM_ERR_ARG, "bad class pointer");
    if (!c->have_post) return fail(c, FEM_ERR_STATE, "no posteriors yet");
    if (int r = set_device(c)) return r;
    int *tmp = nullptr;
    FEM_HIP(c, hipMalloc((void **)&tmp, (size_t)c->n * sizeof(int)));
    hipError_t e = fem::launch_classes(c, tmp);
    int *pt = nullptr;  // point order of the slots' classes
    if (e == hipSuccess && c->sorted) {
        e = hipMalloc((void **)&pt, (size_t)c->n * sizeof(int));
        if (e == hipSuccess) e = fem::launch_unpermute_int(c, tmp, pt);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(cls, pt ? pt : tmp, (size_t)c->n * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (pt) (void)hipFree(pt);
    if (e != hipSuccess) return fail(c, FEM_ERR_HIP, std::string("fem_get_classes: ") + hipGetErrorString(e));
    return FEM_OK;
}

int fem_last_estep_skipped(fem_ctx *c, double *frac) {
    if (int r = check_ctx(c)) return r;
    if (!frac) return fail(c, FEM_ERR_ARG, "null pointer");
    *frac = 0.0;
    if (!c->nskip || !c->refs_valid || c->prec != FEM_PREC_FP64) return FEM_OK;
    if (int r = set_device(c)) return r;
    unsigned long long sk = 0;
    FEM_HIP(c, hipMemcpyAsync(&sk, c->nskip, sizeof(sk), hipMemcpyDeviceToHost, c->stream));
    FEM_HIP(c, hipStreamSynchronize(c->stream));
    const double evals = (double)(c->npad / (16 * c->estep_groups)) * c->K;  // waves x components
    *frac = (double)sk / evals;
    return FEM_OK;
}

int fem_last_estep_form(fem_ctx *c, int *form) {
    if (int r = check_ctx(c)) return r;
    if (!form) return fail(c, FEM_ERR_ARG, "null form pointer");
    *form = c->last_form;
    return FEM_OK;
}

int fem_last_times(fem_ctx *c, float *estep_ms, float *means_ms, float *cov_ms, float *factor_ms) {
    if (int r = check_ctx(c)) return r;
    if (estep_ms) *estep_ms = c->t_estep;
    if (means_ms) *means_ms = c->t_means;
    if (cov_ms) *cov_ms = c->t_cov;
    if (factor_ms) *factor_ms = c->t_factor;
    return FEM_OK;
}

}  // extern "C"
