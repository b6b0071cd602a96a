// fem_batch.hip -- many independent FP64 GMM fits of uneven (n, d, K) advanced
// together: one persistent launch per EM phase walks a work list that covers
// every active fit (E-step tiles, means chunks, covariance chunks x component
// groups, per-component finalisers, factorisations), a per-fit convergence
// kernel applies the reference's stopping rule on the device
// (EM2/EMAlgorithm.cpp:166-183), and one EM iteration of all fits is captured
// in a hipGraph that the host replays until no fit is active.  Fits that have
// converged or failed are skipped by every phase.  The per-unit bodies are the
// ones of the single-fit kernels (fem_tiles.h), so each fit's arithmetic and
// summation order are those of a single-fit run with the same tiles.
//
// The autotuner (fem_batch_tune) times each phase under several persistent
// grid sizes and E-step point-group counts on the batch itself and keeps the
// fastest per phase.
#include <hip/hip_runtime.h>

#include <map>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fem_cov64_kernel.h"
#include "fem_internal.h"
#include "fem_tiles.h"

namespace {

// Device-visible description of one fit: shapes, chunking and the offsets of
// its arrays in the pooled buffers.
struct BFit {
    int64_t n, npad, ntotal;
    int d, dpad, NB, K, KP, npairs, ub;
    int m_chunks, c_chunks;
    int eg, cmul;                       // tile shape: E-step point groups (2, 4), chunk multiplier (1, 2, 4)
    int64_t c_chunk_base, c_chunk_rows; // covariance chunk rows: base (<= 256 chunks) times cmul
    int64_t x;                          // float pool
    int64_t r;                          // double pool: posteriors KP x npad
    int64_t w, mu, cov, nk, logdet, coef, U, stage64, llp, mpart, cpart;  // double pool
    int64_t info;                       // int pool
};

// Per-fit iteration state (EM2/EMAlgorithm.cpp:152-183 semantics).
struct BState {
    double ll;       // LL of the last E-step
    double prev;     // previous LL (starts at -DBL_MAX)
    int active;      // still iterating
    int sweeps;      // E-steps done; the reference's IterCounter = sweeps + 1 at exit
    int status;      // 0 running, 1 converged (e < eps), 2 sweep cap, -1 Cholesky failure
    int pad;
};

enum { ST_RUN = 0, ST_CONV = 1, ST_CAP = 2, ST_FAIL = -1 };

struct Item {
    int fit, a, b, pad;
};

}  // namespace

struct fem_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int F = 0;
    std::vector<BFit> fits;
    BFit *dfits = nullptr;
    BState *dstate = nullptr;
    float *xpool = nullptr;
    double *dpool = nullptr;
    int *ipool = nullptr;
    int64_t nx = 0, nd = 0, ni = 0;
    // work lists
    // E-step tiles of the fits with 2 (estA) and 4 (estB) point groups per
    // wave, each list ordered by the fits' 16-feature blocks NB (the items of
    // NB = nb at [off[nb - 1], off[nb])), covariance (fit, component group,
    // chunk) items, and the shape-independent lists
    Item *estA = nullptr, *estB = nullptr, *means = nullptr, *cov = nullptr, *comp = nullptr, *covred = nullptr;
    int n_estA = 0, n_estB = 0, n_means = 0, n_cov = 0, n_comp = 0, n_covred = 0;
    int offA[5] = {0, 0, 0, 0, 0}, offB[5] = {0, 0, 0, 0, 0}, offC[5] = {0, 0, 0, 0, 0};
    int maxNB = 1;
    int cus = 256;
    // tuned launch grids (the E-step and covariance ones as multiples of the
    // CU count, per NB range of a list)
    int mult_estA = 2, mult_estB = 2, grid_means = 0, mult_cov = 2;
    // run parameters
    int maxsweeps = 100;
    double eps = 1e-5, g = 0.0;
    // covariances over the nonzero posteriors only (FASTEM_COV_DENSE=1: every point)
    bool cov_sparse = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
    std::string tune_report;
};

namespace fem {

// ---------------------------------------------------------------------------
// Phase kernels.  Every loop index is workgroup-uniform.

__global__ __launch_bounds__(256) void kb_factor(const BFit *__restrict__ F, BState *__restrict__ S,
                                                 double *__restrict__ dp, int *__restrict__ ip,
                                                 const Item *__restrict__ items, int nitems, int gate) {
    extern __shared__ double A[];
    __shared__ int fail;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (gate && !S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        factor_unit(itm.a, f.d, f.dpad, f.NB, f.npairs, dp + f.cov, dp + f.w, dp + f.mu, dp + f.logdet,
                    ip + f.info, dp + f.coef, dp + f.U, dp + f.stage64, nullptr, A, &fail);
    }
}

// RegularizeCovariances before the loop (EM2/EMAlgorithm.cpp:145,237-244):
// C <- C (1 - g) + (g tr(C) / d) I for every component of every fit, the
// trace summed in ascending diagonal order.  One workgroup per component.
__global__ __launch_bounds__(256) void kb_regularize(const BFit *__restrict__ F, double *__restrict__ dp,
                                                     const Item *__restrict__ items, int nitems, double g) {
    __shared__ double diag;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        const BFit &f = F[itm.fit];
        double *C = dp + f.cov + (size_t)itm.a * f.d * f.d;
        if (threadIdx.x == 0) {
            double tr = 0.0;
            for (int i = 0; i < f.d; i++) tr += C[(size_t)i * f.d + i];
            diag = g * tr / (double)f.d;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < f.d * f.d; e += 256) {
            const int i = e / f.d, j = e % f.d;
            C[e] = C[e] * (1.0 - g) + (i == j ? diag : 0.0);
        }
        __syncthreads();
    }
}

// One persistent E-step kernel per (NB, P): its items are the tiles of fits
// with NB 16-feature blocks, so each instance gets the registers and LDS of
// its own shape.  One kernel switching over NB took the largest shape's 255
// registers (plus AGPR spills) and one wave per SIMD for every fit; per NB,
// capped at two waves per SIMD (169-250 registers, no spill), the 256-fit
// benchmark batch ran 10% faster (three waves spilled 8-92 registers in the
// persistent loop and gained half as much).
template <int NB, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void kb_estep(
    const BFit *__restrict__ F, const BState *__restrict__ S, const float *__restrict__ xp, double *__restrict__ dp,
    const Item *__restrict__ items, int nitems) {
    extern __shared__ double lds[];  // estep64_lds<NB>()
    __shared__ double wsum[4];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        const int64_t tile = itm.a;
        estep64_unit<NB, P>(tile, xp + f.x, f.n, f.npad, f.K, dp + f.stage64, dp + f.coef, dp + f.r,
                            dp + f.llp + tile, lds, wsum);
    }
}

// LL of each active fit: fixed-order sum of its E-step tile partials.
__global__ __launch_bounds__(256) void kb_ll(const BFit *__restrict__ F, BState *__restrict__ S,
                                             const double *__restrict__ dp) {
    __shared__ double s[256];
    const int fi = blockIdx.x;
    if (!S[fi].active) return;
    const BFit &f = F[fi];
    const double r = block_sum(dp + f.llp, f.npad / (64 * f.eg), s);
    if (threadIdx.x == 0) S[fi].ll = r;
}

__global__ __launch_bounds__(256) void kb_means(const BFit *__restrict__ F, const BState *__restrict__ S,
                                                const float *__restrict__ xp, double *__restrict__ dp,
                                                const Item *__restrict__ items, int nitems) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        double *out = dp + f.mpart + (size_t)itm.a * f.KP * (f.dpad + 1);
        switch (f.NB) {
            case 1: means64_unit<1>(itm.a, xp + f.x, dp + f.r, f.npad, f.KP, out); break;
            case 2: means64_unit<2>(itm.a, xp + f.x, dp + f.r, f.npad, f.KP, out); break;
            case 3: means64_unit<3>(itm.a, xp + f.x, dp + f.r, f.npad, f.KP, out); break;
            case 4: means64_unit<4>(itm.a, xp + f.x, dp + f.r, f.npad, f.KP, out); break;
        }
    }
}

__global__ __launch_bounds__(256) void kb_means_finalize(const BFit *__restrict__ F, const BState *__restrict__ S,
                                                         double *__restrict__ dp, const Item *__restrict__ items,
                                                         int nitems) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        // 64 features per call (its column tile ct); any d
        for (int ct = 0; ct * 64 < f.d; ct++)
            means_finalize_unit(itm.a, ct, f.d, f.dpad, f.KP, f.m_chunks, dp + f.mpart, dp + f.mu, dp + f.nk, 0);
    }
}

// Per NB, as the E-step: one kernel switching over NB took the largest
// shape's 196 registers and two waves per SIMD for every fit; the single-NB
// instances need 41-104 registers below d = 49.
template <int NB>
__global__ __launch_bounds__(256) void kb_cov(const BFit *__restrict__ F, const BState *__restrict__ S,
                                              const float *__restrict__ xp, double *__restrict__ dp,
                                              const Item *__restrict__ items, int nitems) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        cov64_unit<NB>(itm.a, itm.b, xp + f.x, dp + f.r, f.npad, f.K, f.KP, f.d, dp + f.mu, dp + f.nk,
                       f.c_chunk_rows, dp + f.cpart);
    }
}

// The same over each component's nonzero posteriors (the single fit's
// k_cov64_sparse body, d <= 64), with its 48 KB of LDS lists.
template <int NB>
__global__ __launch_bounds__(256) void kb_cov_sparse(const BFit *__restrict__ F, const BState *__restrict__ S,
                                                     const float *__restrict__ xp, double *__restrict__ dp,
                                                     const Item *__restrict__ items, int nitems) {
    __shared__ int lidx[4 * kCovSeg];
    __shared__ double lrv[4 * kCovSeg];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        cov64_sparse<NB, 1, 0>(itm.a, itm.b, xp + f.x, dp + f.r, f.npad, f.K, f.d, dp + f.mu, dp + f.nk,
                               f.c_chunk_rows, f.KP, dp + f.cpart, lidx, lrv);
    }
}

__global__ __launch_bounds__(256) void kb_cov_reduce(const BFit *__restrict__ F, const BState *__restrict__ S,
                                                     double *__restrict__ dp, const Item *__restrict__ items,
                                                     int nitems) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        cov_reduce_unit(itm.a, itm.b, f.d, f.NB, f.KP, f.c_chunks, dp + f.cpart, dp + f.cov);
    }
}

__global__ __launch_bounds__(256) void kb_cov_finish(const BFit *__restrict__ F, const BState *__restrict__ S,
                                                     double *__restrict__ dp, const Item *__restrict__ items,
                                                     int nitems, double g) {
    __shared__ double str;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item itm = items[it];
        if (!S[itm.fit].active) continue;
        const BFit &f = F[itm.fit];
        cov_finish_unit(itm.a, f.d, dp + f.nk, g, (double)f.ntotal, dp + f.cov, dp + f.w, &str);
    }
}

// After the factorisation of the new parameters: Cholesky failure stops the
// fit (the reference throws), else e = |dLL/LL| (0 on a decrease) against eps
// and the sweep cap.
__global__ void kb_converge(const BFit *__restrict__ F, BState *__restrict__ S, const int *__restrict__ ip,
                            int nfits, int maxsweeps, double eps) {
    const int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= nfits) return;
    BState st = S[fi];
    if (!st.active) return;
    const BFit &f = F[fi];
    st.sweeps += 1;
    bool bad = false;
    for (int k = 0; k < f.K; k++) bad |= ip[f.info + k] != 0;
    if (bad) {
        st.active = 0;
        st.status = ST_FAIL;
    } else {
        double e = fabs((st.ll - st.prev) / st.ll);
        if (st.ll < st.prev) e = 0.0;
        st.prev = st.ll;
        if (e < eps) {
            st.active = 0;
            st.status = ST_CONV;
        } else if (st.sweeps >= maxsweeps) {
            st.active = 0;
            st.status = ST_CAP;
        }
    }
    S[fi] = st;
}

// Initial state: active unless the first factorisation failed.
__global__ void kb_start(const BFit *__restrict__ F, BState *__restrict__ S, const int *__restrict__ ip,
                         int nfits, int maxsweeps) {
    const int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= nfits) return;
    const BFit &f = F[fi];
    BState st;
    st.ll = -1.7976931348623157e308;
    st.prev = -1.7976931348623157e308;
    st.sweeps = 0;
    st.pad = 0;
    bool bad = false;
    for (int k = 0; k < f.K; k++) bad |= ip[f.info + k] != 0;
    st.active = (!bad && maxsweeps > 0) ? 1 : 0;
    st.status = bad ? ST_FAIL : (maxsweeps > 0 ? ST_RUN : ST_CAP);
    S[fi] = st;
}

}  // namespace fem

// ---------------------------------------------------------------------------
namespace {

int berr(fem_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    return code;
}

#define BH(b, call)                                                                              \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess) return berr((b), FEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

int grid_of(int mult, int cus, int nitems) { return std::max(1, std::min(nitems, mult * cus)); }

// The E-step of the n items of fits with nb 16-feature blocks under P point
// groups per wave, as one persistent launch of mult x CUs workgroups.
template <int P>
void launch_estep_nb(fem_batch *b, int nb, const Item *items, int n, int mult) {
    if (n <= 0) return;
    const int grid = grid_of(mult, b->cus, n);
    // four point groups exist only for NB <= 2 (build_shape_lists keeps the
    // lists so), so those instances are not built for NB = 3, 4
    switch (nb) {
#define FEM_KB_ESTEP(NBV)                                                                                           \
    case NBV:                                                                                                     \
        if constexpr (P == 2 || NBV <= 2)                                                                         \
            hipLaunchKernelGGL((fem::kb_estep<NBV, P>), dim3(grid), dim3(256),                                    \
                               (size_t)fem::estep64_lds<NBV>() * sizeof(double), b->stream, b->dfits, b->dstate,   \
                               b->xpool, b->dpool, items, n);                                                      \
        break;
        FEM_KB_ESTEP(1)
        FEM_KB_ESTEP(2)
        FEM_KB_ESTEP(3)
        FEM_KB_ESTEP(4)
#undef FEM_KB_ESTEP
    }
}

// The covariance items of fits with nb 16-feature blocks, and a whole
// covariance list (ordered by NB) as one launch per NB.
void launch_cov_nb(fem_batch *b, int nb, const Item *items, int n, int mult) {
    if (n <= 0) return;
    const int grid = grid_of(mult, b->cus, n);
    if (b->cov_sparse) {
        switch (nb) {
            case 1: hipLaunchKernelGGL(fem::kb_cov_sparse<1>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
            case 2: hipLaunchKernelGGL(fem::kb_cov_sparse<2>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
            case 3: hipLaunchKernelGGL(fem::kb_cov_sparse<3>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
            case 4: hipLaunchKernelGGL(fem::kb_cov_sparse<4>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
        }
        return;
    }
    switch (nb) {
        case 1: hipLaunchKernelGGL(fem::kb_cov<1>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
        case 2: hipLaunchKernelGGL(fem::kb_cov<2>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
        case 3: hipLaunchKernelGGL(fem::kb_cov<3>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
        case 4: hipLaunchKernelGGL(fem::kb_cov<4>, dim3(grid), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool, items, n); break;
    }
}

void launch_cov_list(fem_batch *b, const Item *items, const int *off, int mult) {
    for (int nb = 1; nb <= 4; nb++) launch_cov_nb(b, nb, items + off[nb - 1], off[nb] - off[nb - 1], mult);
}

// A whole E-step list (ordered by NB, ranges off[0..4]): one launch per NB.
void launch_estep_list(fem_batch *b, int P, const Item *items, const int *off, int mult) {
    for (int nb = 1; nb <= 4; nb++) {
        if (P == 4) launch_estep_nb<4>(b, nb, items + off[nb - 1], off[nb] - off[nb - 1], mult);
        else launch_estep_nb<2>(b, nb, items + off[nb - 1], off[nb] - off[nb - 1], mult);
    }
}

// one EM iteration of all active fits (the captured sequence)
hipError_t launch_iteration(fem_batch *b) {
    const size_t lds_f = (size_t)64 * 64 * sizeof(double);
    launch_estep_list(b, 2, b->estA, b->offA, b->mult_estA);
    launch_estep_list(b, 4, b->estB, b->offB, b->mult_estB);
    hipLaunchKernelGGL(fem::kb_ll, dim3(b->F), dim3(256), 0, b->stream, b->dfits, b->dstate, b->dpool);
    hipLaunchKernelGGL(fem::kb_means, dim3(b->grid_means), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool,
                       b->dpool, b->means, b->n_means);
    hipLaunchKernelGGL(fem::kb_means_finalize, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), 0, b->stream,
                       b->dfits, b->dstate, b->dpool, b->comp, b->n_comp);
    launch_cov_list(b, b->cov, b->offC, b->mult_cov);
    hipLaunchKernelGGL(fem::kb_cov_reduce, dim3(grid_of(8, b->cus, b->n_covred)), dim3(256), 0, b->stream,
                       b->dfits, b->dstate, b->dpool, b->covred, b->n_covred);
    hipLaunchKernelGGL(fem::kb_cov_finish, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), 0, b->stream,
                       b->dfits, b->dstate, b->dpool, b->comp, b->n_comp, b->g);
    hipLaunchKernelGGL(fem::kb_factor, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), lds_f, b->stream,
                       b->dfits, b->dstate, b->dpool, b->ipool, b->comp, b->n_comp, 1);
    hipLaunchKernelGGL(fem::kb_converge, dim3((b->F + 63) / 64), dim3(64), 0, b->stream, b->dfits, b->dstate,
                       b->ipool, b->F, b->maxsweeps, b->eps);
    return hipGetLastError();
}

void free_batch(fem_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *ptrs[] = {b->dfits, b->dstate, b->xpool, b->dpool, b->ipool, b->estA, b->estB,
                    b->means, b->cov, b->comp, b->covred};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

template <typename T>
hipError_t upload_vec(T **dst, const std::vector<T> &v) {
    hipError_t e = hipMalloc((void **)dst, std::max<size_t>(1, v.size()) * sizeof(T));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Large work lists interleaved across fits (ordered by item index, then
// fit) so that a persistent grid sees a mix of fits.
void interleave(std::vector<Item> &v) {
    std::stable_sort(v.begin(), v.end(), [](const Item &x, const Item &y) {
        if (x.a != y.a) return x.a < y.a;
        return x.fit < y.fit;
    });
}

// E-step tiles of fit f under P point groups per wave; covariance items of
// fit f under its current chunk rows.
void estep_items(std::vector<Item> &v, int i, const BFit &f, int P) {
    for (int64_t t = 0; t < f.npad / (64 * P); t++) v.push_back({i, (int)t, 0, 0});
}
void cov_items(std::vector<Item> &v, int i, const BFit &f) {
    for (int ch = 0; ch < f.c_chunks; ch++)
        for (int cg = 0; cg < f.KP / 4; cg++)
            if (4 * cg < f.K) v.push_back({i, cg, ch, 0});
}

// Orders an E-step list by the fits' NB (stable: the interleaving of fits is
// kept within an NB) and records the range of each NB in off[0..4].
void by_nb(const fem_batch *b, std::vector<Item> &v, int *off) {
    std::stable_sort(v.begin(), v.end(),
                     [&](const Item &x, const Item &y) { return b->fits[x.fit].NB < b->fits[y.fit].NB; });
    off[0] = 0;
    for (int nb = 1; nb <= 4; nb++) {
        int n = 0;
        for (const Item &it : v) n += b->fits[it.fit].NB == nb;
        off[nb] = off[nb - 1] + n;
    }
}

void set_chunks(BFit &f) {
    f.c_chunk_rows = f.c_chunk_base * f.cmul;
    f.c_chunks = (int)((f.npad + f.c_chunk_rows - 1) / f.c_chunk_rows);
}

template <typename T>
hipError_t replace_list(T **dst, int *count, const std::vector<T> &v) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    *count = (int)v.size();
    return upload_vec(dst, v);
}

// (Re)builds the lists that depend on the fits' tile shapes from b->fits
// and uploads the descriptors.
hipError_t build_shape_lists(fem_batch *b) {
    std::vector<Item> a, c4, cv;
    for (int i = 0; i < b->F; i++) {
        BFit &f = b->fits[i];
        set_chunks(f);
        if (f.NB > 2) f.eg = 2;  // four point groups only for d <= 32
        estep_items(f.eg == 4 ? c4 : a, i, f, f.eg);
        cov_items(cv, i, f);
    }
    interleave(a);
    interleave(c4);
    by_nb(b, a, b->offA);
    by_nb(b, c4, b->offB);
    by_nb(b, cv, b->offC);
    hipError_t e;
    if ((e = replace_list(&b->estA, &b->n_estA, a)) != hipSuccess) return e;
    if ((e = replace_list(&b->estB, &b->n_estB, c4)) != hipSuccess) return e;
    if ((e = replace_list(&b->cov, &b->n_cov, cv)) != hipSuccess) return e;
    return hipMemcpy(b->dfits, b->fits.data(), sizeof(BFit) * b->F, hipMemcpyHostToDevice);
}

}  // namespace

extern "C" {

int fem_batch_create(fem_batch **out, int nfits, const int64_t *n, const int *d, const int *K, int device) {
    if (!out || nfits < 1 || !n || !d || !K) return FEM_ERR_ARG;
    *out = nullptr;
    fem_batch *b = new fem_batch();
    b->device = device;
    b->F = nfits;
    auto bail = [&](int code, const std::string &m) {
        free_batch(b);
        delete b;
        (void)m;
        return code;
    };
    for (int i = 0; i < nfits; i++)
        if (n[i] < 1 || d[i] < 1 || d[i] > 64 || K[i] < 1) return bail(FEM_ERR_ARG, "bad fit shape");
    if (hipSetDevice(device) != hipSuccess) return bail(FEM_ERR_HIP, "hipSetDevice");
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    b->cus = cus;
    if (const char *e = std::getenv("FASTEM_COV_DENSE")) b->cov_sparse = std::atoi(e) == 0;
    std::vector<Item> means, comp, covred;
    for (int i = 0; i < nfits; i++) {
        BFit f{};
        f.n = n[i];
        f.ntotal = n[i];
        f.npad = ((n[i] + fem::kRowAlign - 1) / fem::kRowAlign) * fem::kRowAlign;
        f.d = d[i];
        f.NB = (d[i] + 15) / 16;
        f.dpad = 16 * f.NB;
        f.K = K[i];
        f.KP = 16 * ((K[i] + 15) / 16);
        f.npairs = fem::estep_pairs(f.NB);
        f.ub = fem::upper_blocks(f.NB);
        f.m_chunks = (int)(f.npad / 4096);
        const int64_t units = f.npad / 4096, per = (units + 255) / 256;
        f.eg = 2;
        f.cmul = 1;
        f.c_chunk_base = per * 4096;
        set_chunks(f);  // the most chunks of any multiplier: cpart is sized for it
        f.x = b->nx;
        b->nx += f.npad * f.dpad;
        auto take = [&](int64_t cnt) {
            const int64_t o = b->nd;
            b->nd += (cnt + 31) / 32 * 32;
            return o;
        };
        f.r = take((int64_t)f.KP * f.npad);
        f.w = take(f.K);
        f.mu = take((int64_t)f.K * f.d);
        f.cov = take((int64_t)f.K * f.d * f.d);
        f.nk = take(f.K);
        f.logdet = take(f.K);
        f.coef = take(f.K);
        f.U = take((int64_t)f.K * f.dpad * f.dpad);
        f.stage64 = take((int64_t)f.K * fem::stage64_words(f.NB));
        f.llp = take(f.npad / 64);
        f.mpart = take((int64_t)f.m_chunks * f.KP * (f.dpad + 1));
        f.cpart = take((int64_t)f.c_chunks * f.KP * f.ub * 256);
        f.info = b->ni;
        b->ni += f.K;
        b->fits.push_back(f);
        b->maxNB = std::max(b->maxNB, f.NB);
        for (int c = 0; c < f.m_chunks; c++) means.push_back({i, c, 0, 0});
        for (int k = 0; k < f.K; k++) comp.push_back({i, k, 0, 0});
        for (int k = 0; k < f.K; k++)
            for (int u = 0; u < f.ub; u++) covred.push_back({i, k, u, 0});
    }
    interleave(means);
    b->n_means = (int)means.size();
    b->n_comp = (int)comp.size();
    b->n_covred = (int)covred.size();
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = upload_vec(&b->dfits, b->fits)) != hipSuccess ||
        (e = hipMalloc((void **)&b->dstate, sizeof(BState) * nfits)) != hipSuccess ||
        (e = hipMalloc((void **)&b->xpool, sizeof(float) * std::max<int64_t>(1, b->nx))) != hipSuccess ||
        (e = hipMalloc((void **)&b->dpool, sizeof(double) * std::max<int64_t>(1, b->nd))) != hipSuccess ||
        (e = hipMalloc((void **)&b->ipool, sizeof(int) * std::max<int64_t>(1, b->ni))) != hipSuccess ||
        (e = upload_vec(&b->means, means)) != hipSuccess || (e = build_shape_lists(b)) != hipSuccess ||
        (e = upload_vec(&b->comp, comp)) != hipSuccess || (e = upload_vec(&b->covred, covred)) != hipSuccess ||
        // cleared on the batch's own stream, and finished before return: the
        // stream is non-blocking, so a clear queued on the null stream could
        // land after the uploads and parameters the caller enqueues next
        (e = hipMemsetAsync(b->xpool, 0, sizeof(float) * std::max<int64_t>(1, b->nx), b->stream)) != hipSuccess ||
        (e = hipMemsetAsync(b->dpool, 0, sizeof(double) * std::max<int64_t>(1, b->nd), b->stream)) != hipSuccess ||
        (e = hipMemsetAsync(b->ipool, 0, sizeof(int) * std::max<int64_t>(1, b->ni), b->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(b->stream)) != hipSuccess)
        return bail(FEM_ERR_HIP, hipGetErrorString(e));
    b->mult_estA = b->mult_estB = 2;
    b->grid_means = grid_of(2, cus, b->n_means);
    b->mult_cov = 2;
    *out = b;
    return FEM_OK;
}

int fem_batch_destroy(fem_batch *b) {
    if (!b) return FEM_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
    return FEM_OK;
}

const char *fem_batch_last_error(const fem_batch *b) { return b ? b->err.c_str() : ""; }

int fem_batch_upload(fem_batch *b, int fit, const float *X, int64_t ld) {
    if (!b || fit < 0 || fit >= b->F || !X) return berr(b, FEM_ERR_ARG, "bad fit or pointer");
    const BFit &f = b->fits[fit];
    if (ld < f.d) return berr(b, FEM_ERR_ARG, "bad row stride");
    BH(b, hipSetDevice(b->device));
    BH(b, fem::copy_rows_h2d(b->xpool + f.x, (size_t)f.dpad * sizeof(float), X, (size_t)ld * sizeof(float),
                             (size_t)f.d * sizeof(float), f.n, b->stream));
    return FEM_OK;
}

int fem_batch_set_params(fem_batch *b, int fit, const double *w, const double *mu, const double *cov) {
    if (!b || fit < 0 || fit >= b->F || !w || !mu || !cov) return berr(b, FEM_ERR_ARG, "bad fit or pointer");
    const BFit &f = b->fits[fit];
    BH(b, hipSetDevice(b->device));
    BH(b, hipMemcpyAsync(b->dpool + f.w, w, sizeof(double) * f.K, hipMemcpyHostToDevice, b->stream));
    BH(b, hipMemcpyAsync(b->dpool + f.mu, mu, sizeof(double) * f.K * f.d, hipMemcpyHostToDevice, b->stream));
    BH(b, hipMemcpyAsync(b->dpool + f.cov, cov, sizeof(double) * f.K * f.d * f.d, hipMemcpyHostToDevice, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    return FEM_OK;
}

int fem_batch_get_result(fem_batch *b, int fit, double *w, double *mu, double *cov, double *ll, int *iter_counter,
                         int *status) {
    if (!b || fit < 0 || fit >= b->F) return berr(b, FEM_ERR_ARG, "bad fit");
    const BFit &f = b->fits[fit];
    BH(b, hipSetDevice(b->device));
    BState st;
    BH(b, hipMemcpyAsync(&st, b->dstate + fit, sizeof(BState), hipMemcpyDeviceToHost, b->stream));
    if (w) BH(b, hipMemcpyAsync(w, b->dpool + f.w, sizeof(double) * f.K, hipMemcpyDeviceToHost, b->stream));
    if (mu) BH(b, hipMemcpyAsync(mu, b->dpool + f.mu, sizeof(double) * f.K * f.d, hipMemcpyDeviceToHost, b->stream));
    if (cov)
        BH(b, hipMemcpyAsync(cov, b->dpool + f.cov, sizeof(double) * f.K * f.d * f.d, hipMemcpyDeviceToHost,
                             b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    if (ll) *ll = st.ll;
    if (iter_counter) *iter_counter = st.sweeps + 1;
    if (status) *status = st.status;
    return FEM_OK;
}

// Regularises (g > 0) and factorises every fit's current parameters and
// marks the fits active.
static int batch_start(fem_batch *b) {
    const size_t lds_f = (size_t)64 * 64 * sizeof(double);
    if (b->g > 0.0)
        hipLaunchKernelGGL(fem::kb_regularize, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), 0, b->stream,
                           b->dfits, b->dpool, b->comp, b->n_comp, b->g);
    hipLaunchKernelGGL(fem::kb_factor, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), lds_f, b->stream, b->dfits,
                       b->dstate, b->dpool, b->ipool, b->comp, b->n_comp, 0);
    hipLaunchKernelGGL(fem::kb_start, dim3((b->F + 63) / 64), dim3(64), 0, b->stream, b->dfits, b->dstate, b->ipool,
                       b->F, b->maxsweeps);
    BH(b, hipGetLastError());
    return FEM_OK;
}

static int count_active(fem_batch *b, int *active) {
    std::vector<BState> st(b->F);
    BH(b, hipMemcpyAsync(st.data(), b->dstate, sizeof(BState) * b->F, hipMemcpyDeviceToHost, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    int a = 0;
    for (auto &s : st) a += s.active;
    *active = a;
    return FEM_OK;
}

// Runs EM on all fits from their current parameters: regularisation
// coefficient g (applied to the initial covariances, then after every M-step,
// as TrainNoCopy does), stopping rule e < eps or maxsweeps E-steps.  use_graph != 0 replays a
// captured iteration, else launches the same kernels eagerly.  Blocks until
// every fit has stopped.
int fem_batch_run(fem_batch *b, int maxsweeps, double eps, double g, int use_graph) {
    if (!b || maxsweeps < 0 || !(eps > 0.0) || !(g >= 0.0 && g <= 1.0)) return berr(b, FEM_ERR_ARG, "bad argument");
    BH(b, hipSetDevice(b->device));
    const bool relaunch = b->maxsweeps != maxsweeps || b->eps != eps || b->g != g;
    b->maxsweeps = maxsweeps;
    b->eps = eps;
    b->g = g;
    if (relaunch && b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    if (int r = batch_start(b)) return r;
    if (use_graph && !b->exec) {
        BH(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_iteration(b);
        hipGraph_t gr = nullptr;
        hipError_t ce = hipStreamEndCapture(b->stream, &gr);
        if (le != hipSuccess) return berr(b, FEM_ERR_HIP, std::string("capture: ") + hipGetErrorString(le));
        if (ce != hipSuccess) return berr(b, FEM_ERR_HIP, std::string("end capture: ") + hipGetErrorString(ce));
        b->graph = gr;
        BH(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    int active = 1;
    int launched = 0;
    while (active > 0 && launched < maxsweeps) {
        const int burst = std::min(4, maxsweeps - launched);
        for (int i = 0; i < burst; i++) {
            if (use_graph) BH(b, hipGraphLaunch(b->exec, b->stream));
            else BH(b, launch_iteration(b));
        }
        launched += burst;
        if (int r = count_active(b, &active)) return r;
    }
    BH(b, hipStreamSynchronize(b->stream));
    return FEM_OK;
}

// Autotuner.  From the current parameters it picks, per class of fits with
// the same (16-feature blocks, padded K) -- the shape the kernels' work
// depends on -- the E-step point groups per wave (2 or 4) and the
// covariance chunk rows (1, 2 or 4 times the base length of at most 256
// chunks), each by timing that class's own work items under every candidate
// (mean of three launches after a warm one); then the persistent grids of the
// combined lists ({1, 2, 4} x CUs for the E-steps, {1, 2, 4, 8} x CUs for the
// means and covariance passes).  Parameters are restored afterwards.  The
// report is written to `report` (may be NULL).
int fem_batch_tune(fem_batch *b, char *report, int report_len) {
    if (!b) return FEM_ERR_ARG;
    BH(b, hipSetDevice(b->device));
    // the timed phases rewrite the means (and scratch); keep the parameters
    std::vector<std::vector<double>> saved(b->F);
    for (int i = 0; i < b->F; i++) {
        const BFit &f = b->fits[i];
        const int64_t cnt = f.cov + (int64_t)f.K * f.d * f.d - f.w;
        saved[i].resize(cnt);
        BH(b, hipMemcpy(saved[i].data(), b->dpool + f.w, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    }
    const int save_max = b->maxsweeps;
    b->maxsweeps = 1 << 30;
    if (int r = batch_start(b)) return r;
    hipEvent_t e0, e1;
    BH(b, hipEventCreate(&e0));
    BH(b, hipEventCreate(&e1));
    auto timeit = [&](auto &&fn) -> float {
        fn();  // warm
        (void)hipEventRecord(e0, b->stream);
        for (int r = 0; r < 3; r++) fn();
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms / 3.0f;
    };
    std::string rep;
    char line[256];
    std::map<std::pair<int, int>, std::vector<int>> classes;
    for (int i = 0; i < b->F; i++) classes[{b->fits[i].NB, b->fits[i].KP}].push_back(i);
    Item *tmp = nullptr;
    int ntmp = 0;
    struct FreeTmp {
        Item **p;
        ~FreeTmp() {
            if (*p) (void)hipFree(*p);
        }
    } free_tmp{&tmp};
    // E-step point groups per class (four only for d <= 32, as the
    // single-fit E-step: at d = 48, 64 four groups' accumulators and point
    // tiles do not fit the registers of three waves per SIMD)
    for (auto &cl : classes) {
        float best = 1e30f;
        int bestP = 2;
        const int nb = cl.first.first;
        for (int P : {2, 4}) {
            if (P == 4 && nb > 2) continue;
            std::vector<Item> v;
            for (int i : cl.second) estep_items(v, i, b->fits[i], P);
            interleave(v);
            BH(b, replace_list(&tmp, &ntmp, v));
            const float ms = timeit([&]() {
                if (P == 4) launch_estep_nb<4>(b, nb, tmp, ntmp, 2);
                else launch_estep_nb<2>(b, nb, tmp, ntmp, 2);
            });
            snprintf(line, sizeof line, "class d<=%d K<=%d (%zu fits): estep groups=%d %.3f ms\n", 16 * cl.first.first,
                     cl.first.second, cl.second.size(), P, ms);
            rep += line;
            if (ms < best) {
                best = ms;
                bestP = P;
            }
        }
        for (int i : cl.second) b->fits[i].eg = bestP;
    }
    BH(b, build_shape_lists(b));
    // one E-step and means pass with the chosen shapes, so that the timed
    // covariance passes see this iteration's means
    launch_estep_list(b, 2, b->estA, b->offA, 2);
    launch_estep_list(b, 4, b->estB, b->offB, 2);
    hipLaunchKernelGGL(fem::kb_ll, dim3(b->F), dim3(256), 0, b->stream, b->dfits, b->dstate, b->dpool);
    hipLaunchKernelGGL(fem::kb_means, dim3(grid_of(2, b->cus, b->n_means)), dim3(256), 0, b->stream, b->dfits,
                       b->dstate, b->xpool, b->dpool, b->means, b->n_means);
    hipLaunchKernelGGL(fem::kb_means_finalize, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), 0, b->stream,
                       b->dfits, b->dstate, b->dpool, b->comp, b->n_comp);
    BH(b, hipGetLastError());
    // the descriptors are rewritten below by copies outside the stream
    BH(b, hipStreamSynchronize(b->stream));
    // covariance chunk rows per class
    for (auto &cl : classes) {
        float best = 1e30f;
        int bestM = 1;
        for (int mul : {1, 2, 4}) {
            std::vector<Item> v;
            for (int i : cl.second) {
                b->fits[i].cmul = mul;
                set_chunks(b->fits[i]);
                cov_items(v, i, b->fits[i]);
            }
            BH(b, hipMemcpy(b->dfits, b->fits.data(), sizeof(BFit) * b->F, hipMemcpyHostToDevice));
            BH(b, replace_list(&tmp, &ntmp, v));
            const float ms = timeit([&]() { launch_cov_nb(b, cl.first.first, tmp, ntmp, 2); });
            snprintf(line, sizeof line, "class d<=%d K<=%d: cov chunk rows=%lld %.3f ms\n", 16 * cl.first.first,
                     cl.first.second, (long long)b->fits[cl.second[0]].c_chunk_rows, ms);
            rep += line;
            if (ms < best) {
                best = ms;
                bestM = mul;
            }
        }
        for (int i : cl.second) b->fits[i].cmul = bestM;
    }
    BH(b, build_shape_lists(b));
    // persistent grids of the combined lists
    auto pick_grid = [&](const char *name, std::initializer_list<int> mults, int n, int *grid, auto &&launch) {
        float best = 1e30f;
        for (int mult : mults) {
            const int g = grid_of(mult, b->cus, n);
            const float ms = timeit([&]() { launch(g); });
            snprintf(line, sizeof line, "%s grid=%dxCU: %.3f ms\n", name, mult, ms);
            rep += line;
            if (ms < best) {
                best = ms;
                *grid = g;
            }
        }
    };
    // E-step lists: one multiple of the CU count for all of a list's NB ranges
    auto pick_mult = [&](const char *name, int P, const Item *items, const int *off, int n, int *mult) {
        if (n <= 0) return;
        float best = 1e30f;
        for (int m : {1, 2, 4}) {
            const float ms = timeit([&]() { launch_estep_list(b, P, items, off, m); });
            snprintf(line, sizeof line, "%s grid=%dxCU per NB: %.3f ms\n", name, m, ms);
            rep += line;
            if (ms < best) {
                best = ms;
                *mult = m;
            }
        }
    };
    pick_mult("estep groups=2", 2, b->estA, b->offA, b->n_estA, &b->mult_estA);
    pick_mult("estep groups=4", 4, b->estB, b->offB, b->n_estB, &b->mult_estB);
    hipLaunchKernelGGL(fem::kb_ll, dim3(b->F), dim3(256), 0, b->stream, b->dfits, b->dstate, b->dpool);
    pick_grid("means", {1, 2, 4, 8}, b->n_means, &b->grid_means, [&](int g) {
        hipLaunchKernelGGL(fem::kb_means, dim3(g), dim3(256), 0, b->stream, b->dfits, b->dstate, b->xpool, b->dpool,
                           b->means, b->n_means);
    });
    hipLaunchKernelGGL(fem::kb_means_finalize, dim3(grid_of(4, b->cus, b->n_comp)), dim3(256), 0, b->stream,
                       b->dfits, b->dstate, b->dpool, b->comp, b->n_comp);
    if (b->n_cov > 0) {
        float best = 1e30f;
        for (int m : {1, 2, 4, 8}) {
            const float ms = timeit([&]() { launch_cov_list(b, b->cov, b->offC, m); });
            snprintf(line, sizeof line, "cov grid=%dxCU per NB: %.3f ms\n", m, ms);
            rep += line;
            if (ms < best) {
                best = ms;
                b->mult_cov = m;
            }
        }
    }
    for (auto &cl : classes) {
        const BFit &f = b->fits[cl.second[0]];
        snprintf(line, sizeof line, "chosen for class d<=%d K<=%d: estep groups=%d, cov chunk rows=%lld\n",
                 16 * cl.first.first, cl.first.second, f.eg, (long long)f.c_chunk_rows);
        rep += line;
    }
    snprintf(line, sizeof line, "chosen grids: estep %dxCU/%dxCU per NB, means %d, cov %dxCU per NB (CUs=%d)\n",
             b->mult_estA, b->mult_estB, b->grid_means, b->mult_cov, b->cus);
    rep += line;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BH(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->F; i++)
        BH(b, hipMemcpy(b->dpool + b->fits[i].w, saved[i].data(), sizeof(double) * saved[i].size(),
                        hipMemcpyHostToDevice));
    b->maxsweeps = save_max;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    b->tune_report = rep;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return FEM_OK;
}

}  // extern "C"
