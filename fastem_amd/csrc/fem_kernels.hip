// fem_kernels.hip -- CDNA4 (gfx950) kernels of the EM engine, FP64 parity form.
//
// Stage (d) k_factor   : batched Cholesky / log-det / U = L^-T, one workgroup
//                        per component (EM2/NormalDensity.cpp:43-65).
// Stage (a) k_estep64  : fused E-step.  For a tile of points and each
//                        component: Y^T = U_k^T X^T - U_k^T mu_k on
//                        v_mfma_f64_16x16x4f64 (triangular blocks only),
//                        squared norms, log-densities, then log-sum-exp over K
//                        and posteriors (EM2/MatrixEM.cpp:39-87).
// Stage (b) k_means64  : R^T X and N_k per chunk of points on MFMA
//                        (EM2/MatrixEM.cpp:319-344).
// Stage (c) k_cov64    : weighted SYRK per component and chunk on MFMA
//                        (EM2/MatrixEM.cpp:53-67,90-99,346-375).
// Finalisers reduce the per-chunk partials in a fixed order in FP64.
//
// MFMA 16x16x4 operand layout (wave64, lane l): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k.  Accumulator element v of lane l holds
// D[4*v + l/16][l%16] for v_mfma_f64_16x16x4f64 and D[4*(l/16) + v][l%16] for
// v_mfma_f32_16x16x4f32 (checked on gfx950 with a lane-layout probe).
#include "fem_internal.h"

#include <cmath>

namespace fem {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __host__ constexpr int pairs_of(int NB) {
    int p = 0;
    for (int lb = 0; lb < 4 * NB; lb++) p += NB - lb / 4;
    return p;
}

__device__ inline double wave_xor_sum_g(double v) {
    // sum over the four 16-lane groups (lanes l, l^16, l^32, l^48)
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// ---------------------------------------------------------------------------
// Stage (d).  Lower-symmetrise, right-looking Cholesky in LDS (the same
// operation sequence per element as column Cholesky-Crout), log-det, and the
// columns of L^-1 by forward substitution, stored as rows of U = L^-T.  Then
// the E-step operands: fragments of U in MFMA A-operand order and m = U^T mu.
__global__ __launch_bounds__(256) void k_factor(int K, int d, int dpad, int NB, int npairs,
                                                const double *__restrict__ cov,
                                                const double *__restrict__ w,
                                                const double *__restrict__ mu,
                                                double *__restrict__ logdet, int *__restrict__ info,
                                                double *__restrict__ coef, double *__restrict__ U,
                                                double *__restrict__ ufrag, float *__restrict__ ufrag32,
                                                double *__restrict__ minit, float *__restrict__ mus32) {
    extern __shared__ double A[];  // d x d
    __shared__ int fail;
    const int k = blockIdx.x, tid = threadIdx.x;
    const double *C = cov + (size_t)k * d * d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        int i = e / d, j = e % d;
        A[e] = (i >= j) ? C[(size_t)i * d + j] : C[(size_t)j * d + i];
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < d; j++) {
        if (tid == 0) {
            double p = A[j * d + j];
            if (!(p > 0.0)) fail = j + 1;
            else A[j * d + j] = sqrt(p);
        }
        __syncthreads();
        if (fail) break;
        const double ljj = A[j * d + j];
        for (int i = j + 1 + tid; i < d; i += blockDim.x) A[i * d + j] = A[i * d + j] / ljj;
        __syncthreads();
        const int m = d - j - 1;
        for (int e = tid; e < m * m; e += blockDim.x) {
            int i = j + 1 + e / m, c = j + 1 + e % m;
            if (c <= i) A[i * d + c] = A[i * d + c] - A[i * d + j] * A[c * d + j];
        }
        __syncthreads();
    }
    if (fail) {
        if (tid == 0) {
            info[k] = fail;
            logdet[k] = NAN;
            coef[k] = NAN;
        }
        return;
    }
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < d; i++) s += log(A[i * d + i]);
        info[k] = 0;
        logdet[k] = 2.0 * s;
        const double lc = -0.5 * (double)d * log(2.0 * M_PI);
        coef[k] = log(w[k]) + lc - 0.5 * (2.0 * s);
    }
    double *Uk = U + (size_t)k * dpad * dpad;
    for (int e = tid; e < dpad * dpad; e += blockDim.x) Uk[e] = 0.0;
    __syncthreads();
    if (tid < d) {
        const int c = tid;  // column c of L^-1 = row c of U
        for (int i = c; i < d; i++) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int q = c; q < i; q++) t -= A[i * d + q] * Uk[(size_t)c * dpad + q];
            Uk[(size_t)c * dpad + i] = t / A[i * d + i];
        }
    }
    __syncthreads();
    for (int e = tid; e < npairs * 64; e += blockDim.x) {
        int p = e >> 6, lane = e & 63;
        int lb = 0, rem = p;
        while (rem >= NB - lb / 4) { rem -= NB - lb / 4; lb++; }
        int jb = lb / 4 + rem;
        double v = Uk[(size_t)(4 * lb + (lane >> 4)) * dpad + 16 * jb + (lane & 15)];
        ufrag[((size_t)k * npairs + p) * 64 + lane] = v;
        if (ufrag32) ufrag32[((size_t)k * npairs + p) * 64 + lane] = (float)v;
    }
    for (int j = tid; j < dpad; j += blockDim.x) {
        double s = 0.0;
        if (j < d)
            for (int f = 0; f <= j; f++) s += Uk[(size_t)f * dpad + j] * mu[(size_t)k * d + f];
        minit[(size_t)k * dpad + j] = s;
        if (mus32) {
            double m = (j < d) ? mu[(size_t)k * d + j] : 0.0;
            float hi = (float)m;
            mus32[(size_t)k * 2 * dpad + j] = hi;
            mus32[(size_t)k * 2 * dpad + dpad + j] = (float)(m - (double)hi);
        }
    }
}

hipError_t launch_factor(fem_ctx *c) {
    const int npairs = estep_pairs(c->NB);
    const size_t lds = (size_t)c->d * c->d * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_factor,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_factor, dim3(c->K), dim3(256), lds, c->stream, c->K, c->d, c->dpad,
                       c->NB, npairs, c->cov, c->w, c->mu, c->logdet, c->info, c->coef, c->U,
                       c->ufrag, c->ufrag32, c->minit, c->mus32);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stage (a), FP64.  Workgroup = 4 waves; each wave owns P groups of 16 points.
// Lane (g = l/16, c = l%16) holds, for point c of each group, features
// 4*lb + g (lb = 0..4NB-1) as FP32, widened to FP64 per use: the B operand of
// every MFMA.  For component k the A operands (fragments of U_k^T) and
// m_k = U_k^T mu_k are double-buffered in LDS: the global loads of component
// k+1 are issued into registers before the MFMAs of component k and stored to
// the idle buffer after them (one barrier per component).  Accumulators start
// at -m_k so that D = U^T (x - mu).
template <int NB, int P>
__global__ __launch_bounds__(256) void k_estep64(const float *__restrict__ X, int64_t n,
                                                 int64_t npad, int K,
                                                 const double *__restrict__ ufrag,
                                                 const double *__restrict__ minit,
                                                 const double *__restrict__ coef,
                                                 double *__restrict__ Rt,
                                                 double *__restrict__ ll_part) {
    constexpr int DP = 16 * NB;
    constexpr int NPAIR = pairs_of(NB);
    constexpr int STAGE = NPAIR * 64 + DP;
    constexpr int PER = (STAGE + 255) / 256;
    extern __shared__ double lds[];  // 2 * STAGE
    __shared__ double wsum[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int64_t base = (int64_t)blockIdx.x * (64 * P) + wave * (16 * P);

    float xr[P][4 * NB];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++) xr[p][lb] = X[(base + 16 * p + c) * DP + 4 * lb + g];

    auto stage_src = [&](int k, int e) -> double {
        return (e < NPAIR * 64) ? ufrag[(size_t)k * NPAIR * 64 + e]
                                : minit[(size_t)k * DP + (e - NPAIR * 64)];
    };
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int e = tid + 256 * q;
        if (e < STAGE) lds[e] = stage_src(0, e);
    }
    __syncthreads();

    double mx[P];
#pragma unroll
    for (int p = 0; p < P; p++) mx[p] = -INFINITY;

    for (int k = 0; k < K; k++) {
        const double *lu = lds + (k & 1) * STAGE;
        const double *lm = lu + NPAIR * 64;
        const bool more = k + 1 < K;
        double pre[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = tid + 256 * q;
            pre[q] = (more && e < STAGE) ? stage_src(k + 1, e) : 0.0;
        }
        const double ck = coef[k];
        v4d acc[P][NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            v4d init;
            init[0] = -lm[16 * jb + 0 + g];
            init[1] = -lm[16 * jb + 4 + g];
            init[2] = -lm[16 * jb + 8 + g];
            init[3] = -lm[16 * jb + 12 + g];
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][jb] = init;
        }
        int pr = 0;
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++) {
            double b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
#pragma unroll
            for (int jb = lb / 4; jb < NB; jb++) {
                const double a = lu[pr * 64 + lane];
                pr++;
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[p], acc[p][jb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            double q = 0.0;
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
#pragma unroll
                for (int v = 0; v < 4; v++) q += acc[p][jb][v] * acc[p][jb][v];
            q = wave_xor_sum_g(q);
            const double s = -0.5 * q + ck;
            mx[p] = fmax(mx[p], s);
            if ((k & 3) == g) Rt[(size_t)k * npad + base + 16 * p + c] = s;
        }
        if (more) {
            double *dst = lds + ((k + 1) & 1) * STAGE;
#pragma unroll
            for (int q = 0; q < PER; q++) {
                const int e = tid + 256 * q;
                if (e < STAGE) dst[e] = pre[q];
            }
        }
        __syncthreads();
    }

    double llsum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int64_t pt = base + 16 * p + c;
        double se = 0.0;
        for (int k = g; k < K; k += 4) se += exp(Rt[(size_t)k * npad + pt] - mx[p]);
        se = wave_xor_sum_g(se);
        const double lse = log(se) + mx[p];
        const bool valid = pt < n;
        for (int k = g; k < K; k += 4) {
            double *r = Rt + (size_t)k * npad + pt;
            *r = valid ? exp(*r - lse) : 0.0;
        }
        if (g == 0 && valid) llsum += lse;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) llsum += __shfl_xor(llsum, o);
    if (lane == 0) wsum[wave] = llsum;
    __syncthreads();
    if (tid == 0) ll_part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// Deterministic sum of n doubles into *out: fixed strided partials, then a
// fixed tree.  One workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_reduce_sum(const double *__restrict__ parts, int64_t n,
                                                    double *__restrict__ out) {
    __shared__ double s[256];
    double a = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) a += parts[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

template <int NB, int P>
static hipError_t estep64_t(fem_ctx *c) {
    constexpr int NPAIR = pairs_of(NB);
    const size_t lds = (size_t)2 * (NPAIR * 64 + 16 * NB) * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_estep64<NB, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = (int)(c->npad / (64 * P));
    hipLaunchKernelGGL((k_estep64<NB, P>), dim3(blocks), dim3(256), lds, c->stream, c->X, c->n,
                       c->npad, c->K, c->ufrag, c->minit, c->coef, c->Rt, c->ll_part);
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, c->stream, c->ll_part,
                       (int64_t)blocks, c->ll_dev);
    return hipGetLastError();
}

// Point groups per wave of the FP64 E-step: 2 or 4 (4 only for d <= 64).
hipError_t launch_estep(fem_ctx *c) {
    if (c->prec != FEM_PREC_FP64) return launch_estep32(c);
    const bool p4 = c->estep_groups == 4;
    switch (c->NB) {
        case 1: return p4 ? estep64_t<1, 4>(c) : estep64_t<1, 2>(c);
        case 2: return p4 ? estep64_t<2, 4>(c) : estep64_t<2, 2>(c);
        case 3: return p4 ? estep64_t<3, 4>(c) : estep64_t<3, 2>(c);
        case 4: return p4 ? estep64_t<4, 4>(c) : estep64_t<4, 2>(c);
        case 5: return estep64_t<5, 2>(c);
        case 6: return estep64_t<6, 2>(c);
        case 7: return estep64_t<7, 2>(c);
        case 8: return estep64_t<8, 2>(c);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage (b), FP64.  Workgroup = one chunk of kMeansChunk points.  Wave w owns
// component blocks kb = w, w+4, ...; per step of 4 points it loads the R
// fragment once (A[i][k] = R[16kb+i][i0+k]) and one X fragment per feature
// block (B[k][j] = X[i0+k][16jb+j]), issuing NB MFMAs; the A values also sum
// to N_k.  Partials: mpart[chunk][KP][DP+1].
constexpr int kMeansChunk = 4096;

template <int NB>
__global__ __launch_bounds__(256) void k_means64(const float *__restrict__ X,
                                                 const double *__restrict__ Rt, int64_t npad,
                                                 int KP, double *__restrict__ mpart) {
    constexpr int DP = 16 * NB;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int chunk = blockIdx.x;
    const int64_t i_begin = (int64_t)chunk * kMeansChunk;
    const int KB = KP / 16;
    double *out = mpart + (size_t)chunk * KP * (DP + 1);
    for (int kb = wave; kb < KB; kb += 4) {
        const double *rrow = Rt + (size_t)(16 * kb + c) * npad + i_begin + g;
        const float *xrow = X + (i_begin + g) * DP + c;
        v4d acc[NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) acc[jb] = (v4d){0.0, 0.0, 0.0, 0.0};
        double ns = 0.0;
#pragma unroll 4
        for (int s = 0; s < kMeansChunk; s += 4) {
            const double a = rrow[s];
            ns += a;
#pragma unroll
            for (int jb = 0; jb < NB; jb++) {
                const double bv = (double)xrow[(size_t)s * DP + 16 * jb];
                acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[jb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(size_t)(16 * kb + 4 * v + g) * (DP + 1) + 16 * jb + c] = acc[jb][v];
        ns = wave_xor_sum_g(ns);
        if (g == 0) out[(size_t)(16 * kb + c) * (DP + 1) + DP] = ns;
    }
}

// mu_k = M_k / N_k.  Grid (K, column tiles); each thread sums one column over
// the chunks with 4 interleaved accumulators combined in a fixed order.
__device__ inline double chunk_sum4(const double *__restrict__ p, size_t stride, int chunks) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int ch = 0;
    for (; ch + 4 <= chunks; ch += 4) {
        s0 += p[(size_t)ch * stride];
        s1 += p[(size_t)(ch + 1) * stride];
        s2 += p[(size_t)(ch + 2) * stride];
        s3 += p[(size_t)(ch + 3) * stride];
    }
    for (; ch < chunks; ch++) s0 += p[(size_t)ch * stride];
    return (s0 + s1) + (s2 + s3);
}

// raw != 0 writes the sums M_k themselves (k-means centroid sums).
__global__ __launch_bounds__(128) void k_means_finalize(int K, int d, int dpad, int KP, int chunks,
                                                        const double *__restrict__ mpart,
                                                        double *__restrict__ mu,
                                                        double *__restrict__ nk, int raw) {
    const int k = blockIdx.x, tid = threadIdx.x;
    const size_t stride = (size_t)KP * (dpad + 1);
    const double sn = chunk_sum4(mpart + (size_t)k * (dpad + 1) + dpad, stride, chunks);
    if (blockIdx.y == 0 && tid == 0) nk[k] = sn;
    const int j = blockIdx.y * 128 + tid;
    if (j < d) {
        const double s = chunk_sum4(mpart + (size_t)k * (dpad + 1) + j, stride, chunks);
        mu[(size_t)k * d + j] = raw ? s : s / sn;
    }
}

template <int NB>
static hipError_t means_t(fem_ctx *c) {
    hipLaunchKernelGGL(k_means64<NB>, dim3(c->m_chunks), dim3(256), 0, c->stream, c->X, c->Rt,
                       c->npad, c->KP, c->mpart);
    return hipGetLastError();
}

hipError_t launch_means64_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return means_t<1>(c);
        case 2: return means_t<2>(c);
        case 3: return means_t<3>(c);
        case 4: return means_t<4>(c);
        case 5: return means_t<5>(c);
        case 6: return means_t<6>(c);
        case 7: return means_t<7>(c);
        case 8: return means_t<8>(c);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_means_finalize(fem_ctx *c, double *mu_out, double *nk_out, int raw) {
    hipLaunchKernelGGL(k_means_finalize, dim3(c->K, (c->d + 127) / 128), dim3(128), 0, c->stream, c->K,
                       c->d, c->dpad, c->KP, c->m_chunks, c->mpart, mu_out, nk_out, raw);
    return hipGetLastError();
}

hipError_t launch_means(fem_ctx *c) {
    hipError_t e = (c->prec == FEM_PREC_FP64) ? launch_means64_partials(c) : launch_means32_partials(c);
    if (e != hipSuccess) return e;
    return launch_means_finalize(c, c->mu, c->nk, 0);
}

// ---------------------------------------------------------------------------
// Stage (c), FP64.  Workgroup = 4 waves = 4 consecutive components over one
// chunk of points.  Per step of 4 points each lane forms
// T[i0+g][16fb+c] = (x - mu_k) * sqrt(r) / sqrt(N_k) for every feature block
// fb and issues one MFMA per upper block (ib <= jb): D += T_ib^T T_jb.
// Blocks are remapped so the workgroups that share a chunk of X run on the
// same XCD (dispatch is round-robin over 8 XCDs).
template <int NB>
__global__ __launch_bounds__(256) void k_cov64(const float *__restrict__ X,
                                               const double *__restrict__ Rt, int64_t npad,
                                               int K, int KP, int d,
                                               const double *__restrict__ mu,
                                               const double *__restrict__ nk, int64_t chunk_rows,
                                               double *__restrict__ cpart) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    const int cg = logical % KG, chunk = logical / KG;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    const int64_t i0 = (int64_t)chunk * chunk_rows;
    const int64_t i1 = (i0 + chunk_rows < npad) ? i0 + chunk_rows : npad;
    double muv[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) muv[fb] = (16 * fb + c < d) ? mu[(size_t)k * d + 16 * fb + c] : 0.0;
    const double inv = 1.0 / sqrt(nk[k]);
    v4d acc[UB];
#pragma unroll
    for (int u = 0; u < UB; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double *rrow = Rt + (size_t)k * npad;
    for (int64_t i = i0; i < i1; i += 8) {
        // issue the loads of two steps (8 points) before any arithmetic
        double rv[2];
        float xv[2][NB];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            rv[q] = rrow[i + 4 * q + g];
#pragma unroll
            for (int fb = 0; fb < NB; fb++) xv[q][fb] = X[(i + 4 * q + g) * DP + 16 * fb + c];
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double sw = sqrt(rv[q]) * inv;
            double t[NB];
#pragma unroll
            for (int fb = 0; fb < NB; fb++) t[fb] = ((double)xv[q][fb] - muv[fb]) * sw;
            int u = 0;
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = ib; jb < NB; jb++) {
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(t[ib], t[jb], acc[u], 0, 0, 0);
                    u++;
                }
        }
    }
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
#pragma unroll
    for (int u = 0; u < UB; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)u * 256 + (4 * v + g) * 16 + c] = acc[u][v];
}

// Sum the chunk partials (fixed order, 4 interleaved accumulators) of one
// 16x16 upper block per workgroup and write it to both triangles of cov.
__global__ __launch_bounds__(256) void k_cov_reduce(int K, int d, int NB, int KP, int chunks,
                                                    const double *__restrict__ cpart,
                                                    double *__restrict__ cov) {
    const int k = blockIdx.x, u = blockIdx.y, tid = threadIdx.x;
    const int UB = NB * (NB + 1) / 2;
    int ib = 0, rem = u;
    while (rem >= NB - ib) { rem -= NB - ib; ib++; }
    const int jb = ib + rem;
    const int a = 16 * ib + (tid >> 4), bb = 16 * jb + (tid & 15);
    if (a >= d || bb >= d || a > bb) return;
    const double s = chunk_sum4(cpart + ((size_t)k * UB + u) * 256 + tid, (size_t)KP * UB * 256, chunks);
    double *C = cov + (size_t)k * d * d;
    C[(size_t)a * d + bb] = s;
    C[(size_t)bb * d + a] = s;
}

// Regularise component k (EM2/EMAlgorithm.cpp:237-244) and set
// w_k = N_k / ntotal.  One workgroup per k.
__global__ __launch_bounds__(256) void k_cov_finish(int K, int d, const double *__restrict__ nk, double g,
                                                    double ntotal, double *__restrict__ cov,
                                                    double *__restrict__ w) {
    __shared__ double str;
    const int k = blockIdx.x, tid = threadIdx.x;
    double *C = cov + (size_t)k * d * d;
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < d; i++) tr += C[(size_t)i * d + i];
        str = tr;
        w[k] = nk[k] / ntotal;
    }
    __syncthreads();
    const double diag = g * str / (double)d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        const int a = e / d, bb = e % d;
        C[e] = C[e] * (1.0 - g) + (a == bb ? diag : 0.0);
    }
}

template <int NB>
static hipError_t cov_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks;
    hipLaunchKernelGGL(k_cov64<NB>, dim3(blocks), dim3(256), 0, c->stream, c->X, c->Rt, c->npad,
                       c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    return hipGetLastError();
}

hipError_t launch_cov64_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return cov_t<1>(c);
        case 2: return cov_t<2>(c);
        case 3: return cov_t<3>(c);
        case 4: return cov_t<4>(c);
        case 5: return cov_t<5>(c);
        case 6: return cov_t<6>(c);
        case 7: return cov_t<7>(c);
        case 8: return cov_t<8>(c);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_cov_finalize(fem_ctx *c, double g) {
    const int UB = c->NB * (c->NB + 1) / 2;
    hipLaunchKernelGGL(k_cov_reduce, dim3(c->K, UB), dim3(256), 0, c->stream, c->K, c->d, c->NB, c->KP,
                       c->c_chunks, c->cpart, c->cov);
    hipLaunchKernelGGL(k_cov_finish, dim3(c->K), dim3(256), 0, c->stream, c->K, c->d, c->nk, g,
                       (double)c->ntotal, c->cov, c->w);
    return hipGetLastError();
}

hipError_t launch_cov(fem_ctx *c) {
    hipError_t e = (c->prec == FEM_PREC_FP64) ? launch_cov64_partials(c) : launch_cov32_partials(c);
    if (e != hipSuccess) return e;
    return launch_cov_finalize(c, c->g);
}

// ---------------------------------------------------------------------------
__global__ void k_zero(double *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

hipError_t launch_zero(double *p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

__global__ void k_transpose_post(const double *__restrict__ Rt, int64_t npad, int64_t n, int K,
                                 double *__restrict__ out, int64_t ld) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < K; k++) out[i * ld + k] = Rt[(size_t)k * npad + i];
}

hipError_t launch_transpose_post(const fem_ctx *c, double *out_dev, int64_t ld) {
    if (c->prec != FEM_PREC_FP64) return launch_transpose_post32(c, out_dev, ld);
    hipLaunchKernelGGL(k_transpose_post, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0,
                       c->stream, c->Rt, c->npad, c->n, c->K, out_dev, ld);
    return hipGetLastError();
}

}  // namespace fem
