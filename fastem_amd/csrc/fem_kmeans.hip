// fem_kmeans.hip -- k-means initialisation on the device data of a context:
// Forgy seeding (Clust/KMeansInitializer.cpp:33-49), k-means++ seeding
// (Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84) or k-means||
// seeding (Clust/KMeansOrOrInitializer.cpp:143-174), Lloyd sweeps (Clust/KMAlgorithm.cpp:44-87,205-255) and the mixture of the
// final clustering (EM2/EMInitializer.cpp:103-134).
//
// Squared distances are FP64 sums in ascending feature order with no
// multiply-add contraction (double centroid minus float row widened) and the
// argmin keeps the lowest index on ties; the assignment pass takes its
// argmin from FP64 MFMA distances only where a rounding bound proves it
// (k_assign_mm).  The k-means++ D^2 draw follows the sequential FP64 prefix
// sums: decided on the device where a rounding bound proves the sequential
// rule's row (k_kpp_pick), by those sums on the host otherwise.  So seeds and
// assignment counts equal the FP64 oracle's.  Centroid sums
// and counts come from the MFMA means pass applied to the one-hot assignment
// (products by 1.0 are exact; sums in FP64, fixed order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "fem_internal.h"
#include "fem_tiles.h"

namespace fem {

#pragma clang fp contract(off)
// Squared distance between a double centroid and a row held in registers
// (DP floats, zero beyond d); ascending feature order, no contraction.
template <int DP>
__device__ inline double sqdist_reg(const double *c, const float (&x)[DP], int d) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < DP; j++) {
        if (j < d) {
            const double t = c[j] - (double)x[j];
            s = s + t * t;
        }
    }
    return s;
}

template <int DP>
__device__ inline void load_row(const float *__restrict__ X, int64_t i, int dpad, float (&x)[DP]) {
#pragma unroll
    for (int j = 0; j < DP; j += 4) {
        const float4 v = *reinterpret_cast<const float4 *>(X + i * dpad + j);
        x[j] = v.x;
        x[j + 1] = v.y;
        x[j + 2] = v.z;
        x[j + 3] = v.w;
    }
}

// k-means++ distance update: dmin[i] = min(dmin[i], |x_i - c|^2) (or the
// distance itself on reset), and bsum[block] = the workgroup's sum of the
// updated dmin (fixed tree order).
template <int DP>
__global__ __launch_bounds__(256) void k_kpp_update(const float *__restrict__ X, int64_t n, int dpad, int d,
                                                    const double *__restrict__ cen, double *__restrict__ dmin,
                                                    int reset, double *__restrict__ bsum) {
    __shared__ double lc[DP];
    __shared__ double red[256];
    for (int j = threadIdx.x; j < d; j += 256) lc[j] = cen[j];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        float x[DP];
        load_row<DP>(X, i, dpad, x);
        const double s = sqdist_reg<DP>(lc, x, d);
        v = (reset || dmin[i] > s) ? s : dmin[i];
        dmin[i] = v;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// Greedy k-means++ trial: for each of the L candidate centres (L x d in
// cand), pot[block * L + l] = the workgroup's sum over its rows of
// min(dmin_i, |x_i - c_l|^2) (fixed tree order), the potential the seeding
// would have with c_l added.
template <int DP>
__global__ __launch_bounds__(256) void k_kpp_trial(const float *__restrict__ X, int64_t n, int dpad, int d,
                                                   const double *__restrict__ cand, int L,
                                                   const double *__restrict__ dmin, double *__restrict__ pot) {
    __shared__ double lc[kMaxTrials][DP];
    __shared__ double red[256];
    for (int e = threadIdx.x; e < L * d; e += 256) lc[e / d][e % d] = cand[e];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float x[DP];
    double dm = 0.0;
    if (i < n) {
        load_row<DP>(X, i, dpad, x);
        dm = dmin[i];
    }
    for (int l = 0; l < L; l++) {
        double v = 0.0;
        if (i < n) {
            const double s = sqdist_reg<DP>(lc[l], x, d);
            v = dm > s ? s : dm;
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) pot[(size_t)blockIdx.x * L + l] = red[0];
        __syncthreads();
    }
}

// The k-means++ D^2 draw (Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:
// 33-84): the first row i with r <= agg_i / agg_{n-1}, agg the prefix sums of
// dmin, decided on the device where that provably agrees with the host's
// sequential FP64 sums.  Any summation order of m non-negative terms is
// within (m - 1) u of the exact sum relatively (u = 2^-53), so every prefix
// sum formed here and every one the sequential loop forms is within
// (n + terms here) u of the exact one; bnd bounds the resulting difference of
// the quotients.  The device finds the first row whose prefix reaches
// (r - bnd) S; if it also reaches (r + bnd) S, the sequential rule picks the
// same row, and if none reaches (r - bnd) S, it picks none.  out = {row (n:
// none), 1} then; {-1, 0} asks the host to decide with the sequential sums
// (a prefix within bnd of r S, or S not positive and finite).  One workgroup;
// bsum: nb = ceil(n / 256) workgroup sums from k_kpp_update.
__global__ __launch_bounds__(256) void k_kpp_pick(const double *__restrict__ bsum, int nb,
                                                  const double *__restrict__ dmin, int64_t n, double r,
                                                  int64_t *__restrict__ out) {
    __shared__ double ts[256], tp[256];
    __shared__ double S_s;
    __shared__ int first;
    const int t = threadIdx.x;
    const int per = (nb + 255) / 256;
    const int b0 = min(t * per, nb), b1 = min(b0 + per, nb);
    double s = 0.0;
    for (int b = b0; b < b1; b++) s += bsum[b];
    ts[t] = s;
    __syncthreads();
    if (t == 0) {
        double a = 0.0;
        for (int u = 0; u < 256; u++) {
            tp[u] = a;
            a += ts[u];
        }
        S_s = a;
        first = 0x7fffffff;
    }
    __syncthreads();
    const double S = S_s;
    if (!(S > 0.0) || !(S < __builtin_inf())) {
        if (t == 0) out[0] = -1, out[1] = 0;
        return;
    }
    // terms per device prefix: tree 8 + range 2 per + offsets 256 + block 256
    const double bnd = 8.0 * ((double)n + 2.0 * per + 1024.0) * 0x1p-53;
    const double lo = (r - bnd) * S, hi = (r + bnd) * S;
    double a = tp[t];
    int fb = 0x7fffffff;
    for (int b = b0; b < b1; b++) {
        a += bsum[b];
        if (a >= lo) {
            fb = b;
            break;
        }
    }
    atomicMin(&first, fb);
    __syncthreads();
    if (t != 0) return;
    const int b = first;
    if (b == 0x7fffffff) {
        out[0] = n, out[1] = 1;
        return;
    }
    const int tb = b / per;
    double ex = tp[tb];
    for (int q = tb * per; q < b; q++) ex += bsum[q];
    const int64_t i0 = (int64_t)b * 256, i1 = min(n, i0 + 256);
    for (int64_t i = i0; i < i1; i++) {
        ex += dmin[i];
        if (ex >= lo) {
            out[0] = i, out[1] = (ex >= hi) ? 1 : 0;
            return;
        }
    }
    out[0] = -1, out[1] = 0;
}

// One point per thread, its row in registers; centroids staged in LDS.
// Writes the one-hot assignment into the component-major Rt (all KP rows),
// cls[i], and a fixed-order per-workgroup sum of the minimum squared distances.
template <int DP>
__global__ __launch_bounds__(256) void k_assign(const float *__restrict__ X, int64_t n, int64_t npad, int dpad,
                                                int d, int K, int KP, const double *__restrict__ C,
                                                double *__restrict__ Rt, int *__restrict__ cls,
                                                double *__restrict__ fit_part) {
    extern __shared__ double lc[];  // K * d
    __shared__ double red[256];
    for (int e = threadIdx.x; e < K * d; e += 256) lc[e] = C[e];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double best = 0.0;
    if (i < npad) {
        int bj = -1;
        if (i < n) {
            float x[DP];
            load_row<DP>(X, i, dpad, x);
            best = std::numeric_limits<double>::max();
            for (int k = 0; k < K; k++) {
                const double s = sqdist_reg<DP>(lc + (size_t)k * d, x, d);
                if (s < best) {
                    best = s;
                    bj = k;
                }
            }
            cls[i] = bj;
        }
        for (int k = 0; k < KP; k++) Rt[(size_t)k * npad + i] = (k == bj) ? 1.0 : 0.0;
    }
    red[threadIdx.x] = (i < n) ? best : 0.0;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) fit_part[blockIdx.x] = red[0];
}

// The same assignment with the distances' cross terms on FP64 MFMA.  Phase 1
// (the FP64 E-step's tile: lane (g, c) holds features 4 lb + g of point c of
// each of its P groups; A = 16 centroids x 4 features from LDS, B = the point
// tile): a_k = (|x|^2 + |c_k|^2) - 2 x.c_k for every centroid, and per point
// the smallest a (lowest index on ties) and the second smallest.  Every a_k
// is within E = 2 (4 DP + 16) u (|x|^2 + max_k |c_k|^2) of the exact
// distance s_k that sqdist_reg computes (u = 2^-53: the rounding of the two
// norms, of the MFMA dot product over DP features and of the final sum,
// plus that of sqdist_reg itself; both bounded by (2 DP + 8) u (|x|^2 +
// |c_k|^2), doubled for margin), so a gap above 2E between the smallest and
// the second smallest a proves that exact argmin is the MFMA one, strictly.
// Phase 2 (one point per thread): the exact distance to that centroid, or,
// for points without such a gap, the exact scan over all centroids as in
// k_assign; outputs as k_assign's, so both kernels give identical results.
// LDS: KP centroid rows of DP + 4 doubles (zero beyond d and K), then the KP
// squared norms.
template <int DP>
__global__ __launch_bounds__(256) void k_assign_mm(const float *__restrict__ X, int64_t n, int64_t npad,
                                                   int dpad, int d, int K, int KP, const double *__restrict__ C,
                                                   double *__restrict__ Rt, int *__restrict__ cls,
                                                   double *__restrict__ fit_part) {
    constexpr int NL = DP / 4;                 // 4-feature strips
    constexpr int P = DP <= 64 ? 4 : 2;        // 16-point groups per wave and pass
    constexpr int H = 4 / P;                   // passes: 64 points per wave
    constexpr int LD = DP + 4;
    constexpr double kBound = 2.0 * (4.0 * DP + 16.0) * 0x1p-53;
    extern __shared__ double lc[];
    __shared__ double red[256];
    __shared__ int sb[256];  // per point: the proven argmin, or -1
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, c = lane & 15;
    double *ccl = lc + (size_t)KP * LD;
    for (int e = tid; e < KP * LD; e += 256) {
        const int k = e / LD, j = e % LD;
        lc[e] = (k < K && j < d) ? C[(size_t)k * d + j] : 0.0;
    }
    __syncthreads();
    for (int k = tid; k < KP; k += 256) {
        double s = 0.0;
        for (int j = 0; j < d; j++) s += lc[(size_t)k * LD + j] * lc[(size_t)k * LD + j];
        ccl[k] = s;
    }
    __syncthreads();
    double ccmax = 0.0;
    for (int k = 0; k < K; k++) ccmax = fmax(ccmax, ccl[k]);
    const int64_t blk0 = (int64_t)blockIdx.x * 256;

    for (int h = 0; h < H; h++) {
        const int loc0 = wave * 64 + h * 16 * P;
        float xr[P][NL];
#pragma unroll
        for (int p = 0; p < P; p++)
#pragma unroll
            for (int lb = 0; lb < NL; lb++) xr[p][lb] = X[(blk0 + loc0 + 16 * p + c) * DP + 4 * lb + g];
        double xx[P], best[P], second[P];
        int bidx[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            double s = 0.0;
#pragma unroll
            for (int lb = 0; lb < NL; lb++) s += (double)xr[p][lb] * (double)xr[p][lb];
            xx[p] = group_sum4(s);
            best[p] = second[p] = __builtin_inf();
            bidx[p] = 0x7fffffff;
        }
        for (int kb = 0; kb < KP / 16; kb++) {
            v4d acc[P];
#pragma unroll
            for (int p = 0; p < P; p++) acc[p] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int lb = 0; lb < NL; lb++) {
                const double a = lc[(size_t)(16 * kb + c) * LD + 4 * lb + g];  // A[c][g]
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, (double)xr[p][lb], acc[p], 0, 0, 0);
            }
            // element v of lane (g, c) holds x_c . c_k for k = 16 kb + 4 v + g
            // (the FP64 16x16x4 accumulator layout)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int k = 16 * kb + 4 * v + g;
                if (k < K) {
                    const double ck = ccl[k];
#pragma unroll
                    for (int p = 0; p < P; p++) {
                        const double dist = (xx[p] + ck) - 2.0 * acc[p][v];
                        if (dist < best[p]) {
                            second[p] = best[p];
                            best[p] = dist;
                            bidx[p] = k;
                        } else if (dist < second[p]) {
                            second[p] = dist;
                        }
                    }
                }
            }
        }
        // merge the four lane groups of each point (a set function: the
        // smallest value with the lowest index, and the second smallest)
#pragma unroll
        for (int p = 0; p < P; p++) {
#pragma unroll
            for (int o = 16; o < 64; o <<= 1) {
                const double ob = __shfl_xor(best[p], o), os = __shfl_xor(second[p], o);
                const int oi = __shfl_xor(bidx[p], o);
                if (ob < best[p] || (ob == best[p] && oi < bidx[p])) {
                    second[p] = fmin(best[p], os);
                    best[p] = ob;
                    bidx[p] = oi;
                } else {
                    second[p] = fmin(second[p], ob);
                }
            }
            if (g == 0) {
                const bool proven = second[p] - best[p] > 2.0 * kBound * (xx[p] + ccmax);
                sb[loc0 + 16 * p + c] = proven ? bidx[p] : -1;
            }
        }
    }
    __syncthreads();

    const int64_t i = blk0 + tid;
    double bestd = 0.0;
    if (i < npad) {
        int bj = -1;
        if (i < n) {
            float x[DP];
            load_row<DP>(X, i, dpad, x);
            const int cand = sb[tid];
            if (cand >= 0) {
                bj = cand;
                bestd = sqdist_reg<DP>(lc + (size_t)cand * LD, x, d);
            } else {
                bestd = std::numeric_limits<double>::max();
                for (int k = 0; k < K; k++) {
                    const double s = sqdist_reg<DP>(lc + (size_t)k * LD, x, d);
                    if (s < bestd) {
                        bestd = s;
                        bj = k;
                    }
                }
            }
            cls[i] = bj;
        }
        for (int k = 0; k < KP; k++) Rt[(size_t)k * npad + i] = (k == bj) ? 1.0 : 0.0;
    }
    red[tid] = (i < n) ? bestd : 0.0;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) red[tid] += red[tid + st];
        __syncthreads();
    }
    if (tid == 0) fit_part[blockIdx.x] = red[0];
}

// Squared distance of two float rows as Clust/CentroidVector.h:79-87 forms
// it: each difference and square rounded to float, the sum in double,
// ascending feature order.
template <int DP>
__device__ inline double sqdist_ff(const float *c, const float (&x)[DP], int d) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < DP; j++) {
        if (j < d) {
            const float t = x[j] - c[j];
            const float t2 = t * t;
            s = s + (double)t2;
        }
    }
    return s;
}

// k-means|| closest-center update (Clust/KMeansOrOrInitializer.cpp:38-70):
// nc new float centers (row-major, stride d) staged in LDS, numbered from
// `base`; a row moves to a center only if strictly closer, centers in
// order, so ties keep the earlier center.  reset starts from +inf / -1.
template <int DP>
__global__ __launch_bounds__(256) void k_oror_update(const float *__restrict__ X, int64_t n, int dpad, int d,
                                                     const float *__restrict__ cen, int nc, int base,
                                                     double *__restrict__ cd, int *__restrict__ cc, int reset) {
    extern __shared__ float lcf[];  // nc * d
    for (int e = threadIdx.x; e < nc * d; e += 256) lcf[e] = cen[e];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float x[DP];
    load_row<DP>(X, i, dpad, x);
    double best = reset ? __builtin_inf() : cd[i];
    int bi = reset ? -1 : cc[i];
    for (int j = 0; j < nc; j++) {
        const double s = sqdist_ff<DP>(lcf + (size_t)j * d, x, d);
        if (s < best) {
            best = s;
            bi = base + j;
        }
    }
    cd[i] = best;
    cc[i] = bi;
}

// rows idx[0..m) of X packed densely (m x d floats)
__global__ __launch_bounds__(256) void k_gather_rows(const float *__restrict__ X, int dpad, int d,
                                                     const int64_t *__restrict__ idx, int m,
                                                     float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * d) return;
    const int64_t r = e / d, q = e % d;
    out[e] = X[idx[r] * dpad + q];
}
#pragma clang fp contract(on)

template <typename F>
hipError_t by_dpad(int dpad, F &&f) {
    switch (dpad) {
        case 16: return f(std::integral_constant<int, 16>{});
        case 32: return f(std::integral_constant<int, 32>{});
        case 48: return f(std::integral_constant<int, 48>{});
        case 64: return f(std::integral_constant<int, 64>{});
        case 80: return f(std::integral_constant<int, 80>{});
        case 96: return f(std::integral_constant<int, 96>{});
        case 112: return f(std::integral_constant<int, 112>{});
        case 128: return f(std::integral_constant<int, 128>{});
    }
    return hipErrorInvalidValue;
}

}  // namespace fem

namespace {

inline double canonical(std::mt19937_64 &g) {
    const double r = (double)g() / 18446744073709551616.0;
    return r < 1.0 ? r : std::nextafter(1.0, 0.0);
}

struct KmErr {
    int code;
    std::string msg;
};

#define KM_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) throw KmErr{FEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)}; \
    } while (0)

void ensure_scratch(fem_ctx *c) {
    if (!c->kC) KM_HIP(hipMalloc((void **)&c->kC, sizeof(double) * c->K * c->d));
    if (!c->dmin) KM_HIP(hipMalloc((void **)&c->dmin, sizeof(double) * c->npad));
    if (!c->cls) KM_HIP(hipMalloc((void **)&c->cls, sizeof(int) * c->npad));
    if (!c->kfit) KM_HIP(hipMalloc((void **)&c->kfit, sizeof(double) * (c->npad / 256)));
    if (!c->kpick) KM_HIP(hipMalloc((void **)&c->kpick, sizeof(int64_t) * 2));
    if (!c->kcand) KM_HIP(hipMalloc((void **)&c->kcand, sizeof(double) * fem::kMaxTrials * c->d));
    if (!c->kpot) KM_HIP(hipMalloc((void **)&c->kpot, sizeof(double) * (c->npad / 256) * fem::kMaxTrials));
    if (!c->Rt) {
        KM_HIP(hipMalloc((void **)&c->Rt, sizeof(double) * c->KP * c->npad));
        c->owns_rt64 = true;
    }
}

void fetch_row(fem_ctx *c, int64_t i, double *dst) {
    std::vector<float> row(c->d);
    KM_HIP(hipMemcpyAsync(row.data(), c->X + i * c->dpad, sizeof(float) * c->d, hipMemcpyDeviceToHost,
                          c->stream));
    KM_HIP(hipStreamSynchronize(c->stream));
    for (int j = 0; j < c->d; j++) dst[j] = (double)row[j];
}

// assignment with centroids C (host, K x d): returns the mean squared
// distance; leaves the one-hot in Rt and counts/sums computed when asked.
double assign(fem_ctx *c, const std::vector<double> &C, std::vector<double> *sums, std::vector<double> *counts) {
    KM_HIP(hipMemcpyAsync(c->kC, C.data(), sizeof(double) * c->K * c->d, hipMemcpyHostToDevice, c->stream));
    const int blocks = (int)(c->npad / 256);
    const size_t lds = sizeof(double) * c->K * c->d;
    // MFMA form where its LDS (centroid rows padded by 4, their norms) fits
    const size_t lds_mm = sizeof(double) * (size_t)c->KP * (c->dpad + 5);
    KM_HIP(fem::by_dpad(c->dpad, [&](auto DPc) -> hipError_t {
        constexpr int DP = decltype(DPc)::value;
        if (lds_mm <= 150 * 1024) {
            if (lds_mm > 64 * 1024) {
                hipError_t e = hipFuncSetAttribute((const void *)fem::k_assign_mm<DP>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_mm);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(fem::k_assign_mm<DP>, dim3(blocks), dim3(256), lds_mm, c->stream, c->X, c->n,
                               c->npad, c->dpad, c->d, c->K, c->KP, c->kC, c->Rt, c->cls, c->kfit);
            return hipGetLastError();
        }
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute((const void *)fem::k_assign<DP>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(fem::k_assign<DP>, dim3(blocks), dim3(256), lds, c->stream, c->X, c->n, c->npad,
                           c->dpad, c->d, c->K, c->KP, c->kC, c->Rt, c->cls, c->kfit);
        return hipGetLastError();
    }));
    std::vector<double> fp(blocks);
    KM_HIP(hipMemcpyAsync(fp.data(), c->kfit, sizeof(double) * blocks, hipMemcpyDeviceToHost, c->stream));
    if (sums) {
        KM_HIP(fem::launch_means64_partials(c));
        KM_HIP(fem::launch_means_finalize(c, c->kC, c->nk, 1));
        KM_HIP(hipMemcpyAsync(sums->data(), c->kC, sizeof(double) * c->K * c->d, hipMemcpyDeviceToHost,
                              c->stream));
        KM_HIP(hipMemcpyAsync(counts->data(), c->nk, sizeof(double) * c->K, hipMemcpyDeviceToHost, c->stream));
    }
    KM_HIP(hipStreamSynchronize(c->stream));
    double fit = 0.0;
    for (double v : fp) fit += v;
    return fit / (double)c->n;
}

// Lloyd sweeps from centroids C (NaiveKMA::RunKMeans, MinRel 1e-4) and the
// mixture of the final clustering (EM2/EMInitializer.cpp:103-134): counts,
// covariances about the centroids divided by (n_k - 1), weights n_k / n.
// Empty clusters are re-seeded with a row drawn by `draw`.  Throws KmErr.
template <typename Draw>
int lloyd_and_mixture(fem_ctx *c, std::vector<double> &C, Draw &&draw, int max_sweeps, double *w, double *mu,
                      double *cov) {
    const int K = c->K, d = c->d;
    const int64_t n = c->n;
    // the assignments and the mixture's sums are per point: back to point
    // order (the pruned E-step's slot order is rebuilt afterwards)
    fem::reset_order(c);
    double best = std::numeric_limits<double>::max();
    std::vector<double> sums((size_t)K * d), counts(K);
    for (int it = 1; it <= max_sweeps; it++) {
        const double fit = assign(c, C, &sums, &counts);
        for (int k = 0; k < K; k++) {
            if (counts[k] > 0) {
                const double f = 1.0 / counts[k];
                for (int j = 0; j < d; j++) C[(size_t)k * d + j] = sums[(size_t)k * d + j] * f;
            } else {
                const int64_t src = (int64_t)(draw() * (double)n);
                fetch_row(c, src, C.data() + (size_t)k * d);
            }
        }
        const double rel = (best - fit) / best;
        if (fit < std::nextafter(fit, best)) best = fit;
        else break;
        if (rel < 1e-4) break;
        if (it == max_sweeps) break;
    }
    // mixture of the final clustering: counts, then covariances about the
    // centroids divided by (n_k - 1), weights n_k / n
    assign(c, C, &sums, &counts);
    int rc = FEM_OK;
    std::vector<double> nk1(K);
    for (int k = 0; k < K; k++) {
        nk1[k] = counts[k] - 1.0;
        if (counts[k] < 2) rc = FEM_ERR_ARG;
    }
    KM_HIP(hipMemcpyAsync(c->mu, C.data(), sizeof(double) * K * d, hipMemcpyHostToDevice, c->stream));
    KM_HIP(hipMemcpyAsync(c->nk, nk1.data(), sizeof(double) * K, hipMemcpyHostToDevice, c->stream));
    KM_HIP(fem::launch_cov64_partials(c));
    KM_HIP(fem::launch_cov_finalize(c, 0.0));
    for (int k = 0; k < K; k++) w[k] = counts[k] / (double)n;
    KM_HIP(hipMemcpyAsync(c->w, w, sizeof(double) * K, hipMemcpyHostToDevice, c->stream));
    KM_HIP(hipMemcpyAsync(cov, c->cov, sizeof(double) * K * d * d, hipMemcpyDeviceToHost, c->stream));
    KM_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(mu, C.data(), sizeof(double) * K * d);
    c->have_params = true;
    c->factored = false;
    c->have_post = false;
    c->ref_ok = false;
    if (rc != FEM_OK) c->err = "k-means left a cluster with fewer than two points";
    return rc;
}

#pragma clang fp contract(off)
// k-means|| (Clust/KMeansOrOrInitializer.cpp:143-174) with oversampling 1.4
// and 5 rounds (EM2/EMInitializer.cpp:158-161), in the order of a one-thread
// run of the reference: one draw per row per round in row order, selected
// rows become centers in row order.  The closest-center distances and
// indices are kept on the device; each round's cost is summed sequentially
// on the host from a copy of the distances, as are the draws (the shared
// Rand() stream).  The candidate centers, weighted by how many rows each is
// closest to, then go through the weighted k-means++ of
// Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:23-84 on the host (a few
// hundred candidates).  Writes K x d centroids into C.
template <typename Draw>
void kmeansoror_seed(fem_ctx *c, std::vector<double> &C, Draw &&draw) {
    const int K = c->K, d = c->d;
    const int64_t n = c->n;
    const double os = 1.4;
    const int rounds = os <= 0.1 ? 10 : 5;
    const int chunk = std::max(1, 16384 / d);  // centers per launch: 64 KB of LDS
    const int blocks = (int)((n + 255) / 256);
    struct Dev {
        void *p = nullptr;
        ~Dev() {
            if (p) (void)hipFree(p);
        }
    } dcen, didx;
    KM_HIP(hipMalloc(&dcen.p, sizeof(float) * (size_t)chunk * d));
    KM_HIP(hipMalloc(&didx.p, sizeof(int64_t) * (size_t)chunk));
    double *cd = nullptr;
    KM_HIP(hipHostMalloc((void **)&cd, sizeof(double) * n, hipHostMallocDefault));
    struct Pinned {
        double *p;
        ~Pinned() { (void)hipHostFree(p); }
    } pin{cd};
    std::vector<int64_t> rows;
    // Inserts rows[from, to) as centers numbered from `from`.
    auto insert = [&](size_t from, size_t to, int reset) {
        for (size_t b = from; b < to; b += chunk) {
            const int nc = (int)std::min<size_t>(chunk, to - b);
            KM_HIP(hipMemcpyAsync(didx.p, rows.data() + b, sizeof(int64_t) * nc, hipMemcpyHostToDevice,
                                  c->stream));
            hipLaunchKernelGGL(fem::k_gather_rows, dim3((nc * d + 255) / 256), dim3(256), 0, c->stream, c->X,
                               c->dpad, d, (const int64_t *)didx.p, nc, (float *)dcen.p);
            KM_HIP(hipGetLastError());
            KM_HIP(fem::by_dpad(c->dpad, [&](auto DPc) -> hipError_t {
                constexpr int DP = decltype(DPc)::value;
                hipLaunchKernelGGL(fem::k_oror_update<DP>, dim3(blocks), dim3(256), sizeof(float) * nc * d,
                                   c->stream, c->X, n, c->dpad, d, (const float *)dcen.p, nc, (int)b, c->dmin,
                                   c->cls, reset && b == from ? 1 : 0);
                return hipGetLastError();
            }));
        }
    };
    auto cost_of = [&]() {
        KM_HIP(hipMemcpyAsync(cd, c->dmin, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        KM_HIP(hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (int64_t i = 0; i < n; i++) s += cd[i];
        return s;
    };
    rows.push_back((int64_t)(draw() * (double)n));
    insert(0, 1, 1);
    double cost = cost_of();
    for (int r = 0; r < rounds; r++) {
        const size_t base = rows.size();
        for (int64_t i = 0; i < n; i++) {
            const double prob = os * (double)K * cd[i] / cost;
            if (draw() < prob) rows.push_back(i);
        }
        insert(base, rows.size(), 0);
        cost = cost_of();
    }
    // weights: rows closest to each candidate (Clust/KMeansOrOrInitializer.cpp:114-126)
    const int64_t m = (int64_t)rows.size();
    std::vector<int> cls(n);
    KM_HIP(hipMemcpyAsync(cls.data(), c->cls, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    KM_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> wt(m, 0.0);
    for (int64_t i = 0; i < n; i++) wt[cls[i]] += 1.0;
    // candidate rows to the host
    std::vector<float> cand((size_t)m * d);
    for (int64_t b = 0; b < m; b += chunk) {
        const int nc = (int)std::min<int64_t>(chunk, m - b);
        KM_HIP(hipMemcpyAsync(didx.p, rows.data() + b, sizeof(int64_t) * nc, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(fem::k_gather_rows, dim3((nc * d + 255) / 256), dim3(256), 0, c->stream, c->X, c->dpad,
                           d, (const int64_t *)didx.p, nc, (float *)dcen.p);
        KM_HIP(hipGetLastError());
        KM_HIP(hipMemcpyAsync(cand.data() + (size_t)b * d, dcen.p, sizeof(float) * nc * d, hipMemcpyDeviceToHost,
                              c->stream));
        KM_HIP(hipStreamSynchronize(c->stream));
    }
    // weighted k-means++ over the candidates; distances: double centroid
    // minus float row, ascending feature order
    std::vector<double> cd2(m, std::numeric_limits<double>::infinity()), agg(m);
    const int64_t r0 = (int64_t)(draw() * (double)m);
    for (int j = 0; j < d; j++) C[j] = (double)cand[(size_t)r0 * d + j];
    for (int k = 1; k < K; k++) {
        const double *pc = C.data() + (size_t)(k - 1) * d;
        for (int64_t j = 0; j < m; j++) {
            double s = 0.0;
            for (int q = 0; q < d; q++) {
                const double t = pc[q] - (double)cand[(size_t)j * d + q];
                s += t * t;
            }
            if (cd2[j] > s) cd2[j] = s;
        }
        double sum = 0.0;
        for (int64_t j = 0; j < m; j++) {
            sum += cd2[j] * wt[j];
            agg[j] = sum;
        }
        for (int64_t j = 0; j < m; j++) agg[j] /= sum;
        const double rr = draw();
        for (int64_t j = 0; j < m; j++) {
            if (rr <= agg[j]) {
                for (int q = 0; q < d; q++) C[(size_t)k * d + q] = (double)cand[(size_t)j * d + q];
                break;
            }
        }
    }
}
#pragma clang fp contract(on)

}  // namespace

extern "C" int fem_kmeans_init_rng(fem_ctx *c, double (*rnd)(void *), void *rnd_state, int seeding,
                                   int max_sweeps, double *w, double *mu, double *cov);

static double mt_rand(void *p) { return canonical(*static_cast<std::mt19937_64 *>(p)); }

extern "C" int fem_kmeans_init(fem_ctx *c, uint64_t seed, int seeding, int max_sweeps, double *w, double *mu,
                               double *cov) {
    std::mt19937_64 rng(seed);
    return fem_kmeans_init_rng(c, mt_rand, &rng, seeding, max_sweeps, w, mu, cov);
}

extern "C" int fem_kmeans_init_rng(fem_ctx *c, double (*rnd)(void *), void *rnd_state, int seeding,
                                   int max_sweeps, double *w, double *mu, double *cov) {
    if (!c) return FEM_ERR_ARG;
    if (!rnd || !w || !mu || !cov || max_sweeps < 0 ||
        (seeding != FEM_SEED_FORGY && seeding != FEM_SEED_KMEANSPP && seeding != FEM_SEED_KMEANSOROR &&
         seeding != FEM_SEED_KMEANSPP_GREEDY)) {
        c->err = "fem_kmeans_init: bad argument";
        return FEM_ERR_ARG;
    }
    try {
        KM_HIP(hipSetDevice(c->device));
        ensure_scratch(c);
        const int K = c->K, d = c->d;
        const int64_t n = c->n;
        auto draw = [&]() { return rnd(rnd_state); };
        std::vector<double> C((size_t)K * d, 0.0);
        if (seeding == FEM_SEED_KMEANSPP || seeding == FEM_SEED_KMEANSPP_GREEDY) {
            // Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84; the
            // greedy form draws L = 2 + floor(ln K) candidates per seed by the
            // same rule and keeps the one of least potential
            const bool greedy = seeding == FEM_SEED_KMEANSPP_GREEDY;
            const int L = greedy ? std::min(fem::kMaxTrials, 2 + (int)std::log((double)K)) : 1;
            const int64_t r0 = (int64_t)(draw() * (double)n);
            fetch_row(c, r0, C.data());
            double *dm = nullptr;
            KM_HIP(hipHostMalloc((void **)&dm, sizeof(double) * n, hipHostMallocDefault));
            std::vector<double> agg(n);
            struct Pinned {
                double *p;
                ~Pinned() { (void)hipHostFree(p); }
            } pin{dm};
            const int blocks = (int)((n + 255) / 256);
            // FASTEM_KPP_DRAW=host decides every draw with the host's
            // sequential sums (the path the device falls back to)
            const char *kd = std::getenv("FASTEM_KPP_DRAW");
            const bool host_draw = kd && std::strcmp(kd, "host") == 0;
            int agg_step = 0;  // seed step whose prefix sums agg holds (0: none)
            double sum = 0.0;
            // the row of the D^2 rule for the uniform variate r at seed step k
            // (n: none)
            auto pick_row = [&](int k, double r) -> int64_t {
                if (!host_draw) {
                    hipLaunchKernelGGL(fem::k_kpp_pick, dim3(1), dim3(256), 0, c->stream, c->kfit, blocks, c->dmin,
                                       n, r, c->kpick);
                    KM_HIP(hipGetLastError());
                    int64_t pick[2];
                    KM_HIP(hipMemcpyAsync(pick, c->kpick, sizeof pick, hipMemcpyDeviceToHost, c->stream));
                    KM_HIP(hipStreamSynchronize(c->stream));
                    if (pick[1] == 1) return pick[0];
                }
                // undecided on the device: the sequential FP64 sums
                if (agg_step != k) {
                    KM_HIP(hipMemcpyAsync(dm, c->dmin, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
                    KM_HIP(hipStreamSynchronize(c->stream));
                    sum = 0.0;
                    for (int64_t i = 0; i < n; i++) {
                        sum += dm[i] * 1.0;
                        agg[i] = sum;
                    }
                    agg_step = k;
                }
                // first i with r <= agg[i] / sum: agg is non-decreasing and
                // rounded division by sum > 0 keeps that order, so a binary
                // search over the quotients finds the same row as the
                // reference's scan over the normalised prefix sums
                int64_t lo = 0, hi = n;  // answer in [lo, hi); hi = n: none
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (r <= agg[mid] / sum) hi = mid;
                    else lo = mid + 1;
                }
                return lo;
            };
            std::vector<double> cand((size_t)fem::kMaxTrials * d), pot((size_t)blocks * fem::kMaxTrials);
            for (int k = 1; k < K; k++) {
                KM_HIP(hipMemcpyAsync(c->kC, C.data() + (size_t)(k - 1) * d, sizeof(double) * d,
                                      hipMemcpyHostToDevice, c->stream));
                KM_HIP(fem::by_dpad(c->dpad, [&](auto DPc) -> hipError_t {
                    constexpr int DP = decltype(DPc)::value;
                    hipLaunchKernelGGL(fem::k_kpp_update<DP>, dim3(blocks), dim3(256), 0, c->stream, c->X, n,
                                       c->dpad, d, c->kC, c->dmin, k == 1 ? 1 : 0, c->kfit);
                    return hipGetLastError();
                }));
                if (!greedy) {
                    const int64_t i = pick_row(k, draw());
                    if (i < n) fetch_row(c, i, C.data() + (size_t)k * d);
                    continue;
                }
                // greedy: L draws (each consumes one variate, in order), the
                // candidates' potentials sum(min(dmin, |x - c|^2)) on the
                // device, summed over the workgroups in order on the host;
                // the least potential wins, ties to the earlier draw
                int nc = 0;
                for (int l = 0; l < L; l++) {
                    const int64_t i = pick_row(k, draw());
                    if (i < n) fetch_row(c, i, cand.data() + (size_t)(nc++) * d);
                }
                if (nc == 0) continue;
                KM_HIP(hipMemcpyAsync(c->kcand, cand.data(), sizeof(double) * nc * d, hipMemcpyHostToDevice,
                                      c->stream));
                KM_HIP(fem::by_dpad(c->dpad, [&](auto DPc) -> hipError_t {
                    constexpr int DP = decltype(DPc)::value;
                    hipLaunchKernelGGL(fem::k_kpp_trial<DP>, dim3(blocks), dim3(256), 0, c->stream, c->X, n,
                                       c->dpad, d, c->kcand, nc, c->dmin, c->kpot);
                    return hipGetLastError();
                }));
                KM_HIP(hipMemcpyAsync(pot.data(), c->kpot, sizeof(double) * blocks * nc, hipMemcpyDeviceToHost,
                                      c->stream));
                KM_HIP(hipStreamSynchronize(c->stream));
                int best = 0;
                double bestp = 0.0;
                for (int l = 0; l < nc; l++) {
                    double p = 0.0;
                    for (int bl = 0; bl < blocks; bl++) p += pot[(size_t)bl * nc + l];
                    if (l == 0 || p < bestp) best = l, bestp = p;
                }
                std::copy(cand.begin() + (size_t)best * d, cand.begin() + (size_t)(best + 1) * d,
                          C.begin() + (size_t)k * d);
            }
        } else if (seeding == FEM_SEED_FORGY) {
            // Clust/KMeansInitializer.cpp:33-49: K rows at (long)(Rand()*N),
            // in component order, repeats allowed
            for (int k = 0; k < K; k++) {
                const int64_t r = (int64_t)(draw() * (double)n);
                fetch_row(c, r, C.data() + (size_t)k * d);
            }
        } else {
            kmeansoror_seed(c, C, draw);
        }
        return lloyd_and_mixture(c, C, draw, max_sweeps, w, mu, cov);
    } catch (KmErr &e) {
        c->err = e.msg;
        return e.code;
    }
}

// Lloyd sweeps and the cluster mixture from caller-supplied centroids
// (K x d, e.g. the learned initializer's proposals); empty clusters are
// re-seeded from rows drawn with rnd(rnd_state), or a mt19937_64(seed) stream.
extern "C" int fem_kmeans_init_seeded_rng(fem_ctx *c, const double *C0, double (*rnd)(void *), void *rnd_state,
                                          int max_sweeps, double *w, double *mu, double *cov) {
    if (!c) return FEM_ERR_ARG;
    if (!C0 || !rnd || !w || !mu || !cov || max_sweeps < 0) {
        c->err = "fem_kmeans_init_seeded: bad argument";
        return FEM_ERR_ARG;
    }
    try {
        KM_HIP(hipSetDevice(c->device));
        ensure_scratch(c);
        std::vector<double> C(C0, C0 + (size_t)c->K * c->d);
        auto draw = [&]() { return rnd(rnd_state); };
        return lloyd_and_mixture(c, C, draw, max_sweeps, w, mu, cov);
    } catch (KmErr &e) {
        c->err = e.msg;
        return e.code;
    }
}

extern "C" int fem_kmeans_init_seeded(fem_ctx *c, const double *C0, uint64_t seed, int max_sweeps, double *w,
                                      double *mu, double *cov) {
    std::mt19937_64 rng(seed);
    return fem_kmeans_init_seeded_rng(c, C0, mt_rand, &rng, max_sweeps, w, mu, cov);
}

// Rows idx[0..m) of the context's data copied to host memory (m x d floats).
int fem::gather_rows_host(fem_ctx *c, const int64_t *idx, int m, float *out) {
    if (!c || m < 0 || (m > 0 && (!idx || !out))) return FEM_ERR_ARG;
    if (m == 0) return FEM_OK;
    for (int j = 0; j < m; j++)
        if (idx[j] < 0 || idx[j] >= c->n) {
            c->err = "gather_rows_host: row index out of range";
            return FEM_ERR_ARG;
        }
    int64_t *didx = nullptr;
    float *dout = nullptr;
    auto done = [&](int rc, const char *what, hipError_t e) {
        if (didx) (void)hipFree(didx);
        if (dout) (void)hipFree(dout);
        if (rc != FEM_OK) c->err = std::string(what) + ": " + hipGetErrorString(e);
        return rc;
    };
    hipError_t e;
    if ((e = hipSetDevice(c->device)) != hipSuccess) return done(FEM_ERR_HIP, "hipSetDevice", e);
    if ((e = hipMalloc((void **)&didx, sizeof(int64_t) * m)) != hipSuccess ||
        (e = hipMalloc((void **)&dout, sizeof(float) * (size_t)m * c->d)) != hipSuccess)
        return done(FEM_ERR_HIP, "hipMalloc", e);
    if ((e = hipMemcpyAsync(didx, idx, sizeof(int64_t) * m, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return done(FEM_ERR_HIP, "copy", e);
    hipLaunchKernelGGL(fem::k_gather_rows, dim3((unsigned)(((int64_t)m * c->d + 255) / 256)), dim3(256), 0,
                       c->stream, c->X, c->dpad, c->d, didx, m, dout);
    if ((e = hipGetLastError()) != hipSuccess) return done(FEM_ERR_HIP, "gather", e);
    if ((e = hipMemcpyAsync(out, dout, sizeof(float) * (size_t)m * c->d, hipMemcpyDeviceToHost, c->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return done(FEM_ERR_HIP, "copy", e);
    return done(FEM_OK, "", hipSuccess);
}
