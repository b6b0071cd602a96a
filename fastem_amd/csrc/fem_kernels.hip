// fem_kernels.hip -- single-fit launchers of the FP64 EM kernels (gfx950).
// The per-workgroup bodies live in fem_tiles.h and are shared with the
// batched persistent kernels of fem_batch.hip.
//
// Stage (d) k_factor   : batched Cholesky / log-det / U = L^-T, one workgroup
//                        per component (EM2/NormalDensity.cpp:43-65).
// Stage (a) k_estep64  : fused E-step on v_mfma_f64_16x16x4f64, log-sum-exp
//                        and posteriors (EM2/MatrixEM.cpp:39-87).
// Stage (b) k_means64  : R^T X and N_k per chunk of points on MFMA
//                        (EM2/MatrixEM.cpp:319-344).
// Stage (c) k_cov64    : in fem_cov64.hip (built with AGPR accumulators, which
//                        suit it better than the VGPR form this file uses).
// Finalisers reduce the per-chunk partials in a fixed order in FP64.
#include "fem_internal.h"
#include "fem_tiles.h"

namespace fem {

static_assert(stage64_words_of(1) == stage64_words(1) && stage64_words_of(4) == stage64_words(4) &&
                  stage64_words_of(5) == stage64_words(5) && stage64_words_of(8) == stage64_words(8),
              "FP64 stage block size: host and device formulas");

__global__ __launch_bounds__(256) void k_factor(int K, int d, int dpad, int NB, int npairs,
                                                const double *__restrict__ cov,
                                                const double *__restrict__ w,
                                                const double *__restrict__ mu,
                                                double *__restrict__ logdet, int *__restrict__ info,
                                                double *__restrict__ coef, double *__restrict__ U,
                                                double *__restrict__ stage64, float *__restrict__ stage32) {
    extern __shared__ double A[];  // d x d
    __shared__ int fail;
    factor_unit(blockIdx.x, d, dpad, NB, npairs, cov, w, mu, logdet, info, coef, U, stage64, stage32, A, &fail);
}

// Components by descending weight, ties to the lower index (a weight that
// is not a non-negative number ranks last): the FP64 E-step's processing
// order.  K <= 128: one workgroup, a rank count per component.
__global__ __launch_bounds__(256) void k_order(int K, const double *__restrict__ w, int *__restrict__ order) {
    for (int k = threadIdx.x; k < K; k += 256) {
        const double wk = w[k] >= 0.0 ? w[k] : -1.0;
        int r = 0;
        for (int j = 0; j < K; j++) {
            const double wj = w[j] >= 0.0 ? w[j] : -1.0;
            r += (wj > wk) || (wj == wk && j < k);
        }
        order[r] = k;
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
                       c->stage64, c->stage32);
    if (c->korder)  // contexts without it (fem_estep_noalloc's) keep the index order
        hipLaunchKernelGGL(k_order, dim3(1), dim3(256), 0, c->stream, c->K, c->w, c->korder);
    if (c->prec == FEM_PREC_BF16X3) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_prep_bf16(c);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// d <= 64: registers capped for three waves per SIMD (the compiler's own
// allocation, 206 VGPRs at d = 64, leaves two; capped, a few values spill
// to scratch and the E-step measured 2-3% faster at d = 32 and 64).  Wider
// d keeps the compiler's allocation (the cap spilled heavily at d = 128).
template <int NB, int P, bool PRUNE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_estep64_w3(
    const float *__restrict__ X, int64_t n, int64_t npad, int K, const double *__restrict__ stage,
    const double *__restrict__ coef, double *__restrict__ Rt, double *__restrict__ ll_part,
    const int *__restrict__ order, int *__restrict__ ref, unsigned long long *__restrict__ nskip) {
    extern __shared__ double lds[];  // estep64_lds<NB>()
    __shared__ double wsum[4];
    __shared__ int lint[128];
    estep64_unit<NB, P, PRUNE>(blockIdx.x, X, n, npad, K, stage, coef, Rt, ll_part + blockIdx.x, lds, wsum, order,
                               ref, lint, nskip);
}

template <int NB, int P, bool PRUNE = false>
__global__ __launch_bounds__(256) void k_estep64(const float *__restrict__ X, int64_t n,
                                                 int64_t npad, int K,
                                                 const double *__restrict__ stage,
                                                 const double *__restrict__ coef,
                                                 double *__restrict__ Rt,
                                                 double *__restrict__ ll_part,
                                                 const int *__restrict__ order, int *__restrict__ ref,
                                                 unsigned long long *__restrict__ nskip) {
    extern __shared__ double lds[];  // estep64_lds<NB>()
    __shared__ double wsum[4];
    __shared__ int lint[128];
    estep64_unit<NB, P, PRUNE>(blockIdx.x, X, n, npad, K, stage, coef, Rt, ll_part + blockIdx.x, lds, wsum, order,
                               ref, lint, nskip);
}

// d = 65..80: LDS leaves room for two workgroups per CU; registers capped
// so that two waves per SIMD fit
template <int NB, int P, bool PRUNE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_estep64_w2(
    const float *__restrict__ X, int64_t n, int64_t npad, int K, const double *__restrict__ stage,
    const double *__restrict__ coef, double *__restrict__ Rt, double *__restrict__ ll_part,
    const int *__restrict__ order, int *__restrict__ ref, unsigned long long *__restrict__ nskip) {
    extern __shared__ double lds[];  // estep64_lds<NB>()
    __shared__ double wsum[4];
    __shared__ int lint[128];
    estep64_unit<NB, P, PRUNE>(blockIdx.x, X, n, npad, K, stage, coef, Rt, ll_part + blockIdx.x, lds, wsum, order,
                               ref, lint, nskip);
}

// ---------------------------------------------------------------------------
// The pruned FP64 E-step, one wave at a time (K <= 64, NB >= 2).  The points
// are in the sorted slot order (fem_sort.hip), so a wave's 16P points mostly
// share one component.  Each wave takes the components its points named in
// the last E-step (ref) first, ascending, then the others by descending
// weight, and for each forms only the first 16-row block from the global
// stage block (10 quads; accumulators start at -m_k): its squared norm q0
// bounds q from below (q's in-lane sum starts with q0's terms in the same
// order and adds non-negative ones; the cross-group sums are the same
// additions), so ub = coef_k - q0 / 2 bounds the log-density from above.
// When ub < running max - alive_range for every point of the wave, the
// plain kernel would neither store this log-density nor change the maximum
// (its posterior is exactly zero), and the component is skipped; otherwise
// the rest of U_k^T (x - mu_k) is formed with the plain kernel's products in
// its order (the pairs of the first four strips, then strip by strip), its
// fragments read from the global stage block.  Every log-density formed is
// the plain kernel's bit for bit, and the outputs do not depend on the
// order.  No LDS stage and no barrier per component: the waves run
// independently until the tile's LL.  ref receives each point's component
// of largest log-density (K for padding rows); *nskip counts the skipped
// (wave, component) pairs.  No reference counterpart (the reference
// evaluates every component, EM2/MatrixEM.cpp:69-87).
template <int NB, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_estep64_pruned(
    const float *__restrict__ X, int64_t n, int64_t npad, int K, const double *__restrict__ stage,
    const double *__restrict__ coef, double *__restrict__ Rt, double *__restrict__ ll_part,
    const int *__restrict__ order, int *__restrict__ ref, unsigned long long *__restrict__ nskip) {
    constexpr int DP = 16 * NB;
    constexpr int NOFF = offd_pairs_of(NB);
    constexpr int NDIAG = diag_quads_of(NB);
    constexpr int STAGE = estep64_stage<NB>();
    __shared__ double wsum[4];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, c = lane & 15;
    const int slot = (lane & 3) + 4 * g;  // this lane's 4x4x4 A slot
    const int64_t base = (int64_t)blockIdx.x * (64 * P) + wave * (16 * P);

    float xr[P][4 * NB];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++)
            xr[p][lb] = __builtin_nontemporal_load(&X[(base + 16 * p + c) * DP + 4 * lb + g]);

    // the components this wave's points named: a mask over lanes
    uint64_t named = 0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int r = ref[base + 16 * p + c];
        if (r >= 0 && r < K) named |= 1ull << r;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned lo = (unsigned)named, hi = (unsigned)(named >> 32);
        named |= ((uint64_t)(unsigned)__shfl_xor((int)hi, o) << 32) | (unsigned)__shfl_xor((int)lo, o);
    }
    // (readfirstlane returns int: widened through unsigned, so that bit 31
    // does not spread into the high word)
    named = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(named >> 32)) << 32) |
            (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)named);
    named &= K >= 64 ? ~0ull : (1ull << K) - 1;

    double mx[P];
    unsigned stored[P];  // bit k/4: log-density of component k stored
    int best[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        mx[p] = -INFINITY;
        stored[p] = 0u;
        best[p] = K;
    }
    unsigned skipped = 0;

    auto process = [&](int k) {
        const double *blk = stage + (size_t)k * STAGE;
        const double *lu = blk;
        const double *lq = blk + NOFF * 64;
        const double *lm = lq + NDIAG * 16;
        const double ck = coef[k];
        v4d acc[P][NB];
        {
            v4d init;
            init[0] = lm[0 + g];
            init[1] = lm[4 + g];
            init[2] = lm[8 + g];
            init[3] = lm[12 + g];
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][0] = init;
        }
        // quads of strip lb (quad qd0 first) and pairs of strip lb (pair pr0
        // first), from the global stage block
        auto quads = [&](int lb, int qd0) {
            double b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
            const int jd = lb / 4;
            int qd = qd0;
#pragma unroll
            for (int v = lb % 4; v < 4; v++, qd++) {
                const double a = lq[(qd >> 1) * 32 + slot * 2 + (qd & 1)];
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jd][v] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b[p], acc[p][jd][v], 0, 0, 0);
            }
        };
        auto pairs = [&](int lb, int pr0) {
            double b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
            int pr = pr0;
#pragma unroll
            for (int jb = lb / 4 + 1; jb < NB; jb++, pr++) {
                const double a = lu[(pr >> 1) * 128 + lane * 2 + (pr & 1)];
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[p], acc[p][jb], 0, 0, 0);
            }
        };
        auto quad0 = [](int lb) { return 10 * (lb / 4) + (lb % 4) * 4 - (lb % 4) * (lb % 4 - 1) / 2; };
        auto pair0 = [](int lb) {
            int r = 0;
            for (int l = 0; l < lb; l++) r += NB - 1 - l / 4;
            return r;
        };
#pragma unroll
        for (int lb = 0; lb < 4; lb++) quads(lb, quad0(lb));
        bool out = true;
#pragma unroll
        for (int p = 0; p < P; p++) {
            double q0 = 0.0;
#pragma unroll
            for (int v = 0; v < 4; v++) q0 += acc[p][0][v] * acc[p][0][v];
            q0 = wave_xor_sum_g(q0);
            out = out && (base + 16 * p + c >= n || -0.5 * q0 + ck < mx[p] - alive_range<double>());
        }
        if (__all(out)) {
            skipped++;
            return;
        }
#pragma unroll
        for (int jb = 1; jb < NB; jb++) {
            v4d init;
            init[0] = lm[16 * jb + 0 + g];
            init[1] = lm[16 * jb + 4 + g];
            init[2] = lm[16 * jb + 8 + g];
            init[3] = lm[16 * jb + 12 + g];
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][jb] = init;
        }
#pragma unroll
        for (int lb = 0; lb < 4; lb++) pairs(lb, pair0(lb));
#pragma unroll
        for (int lb = 4; lb < 4 * NB; lb++) {
            quads(lb, quad0(lb));
            pairs(lb, pair0(lb));
        }
        double sv[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            double q = 0.0;
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
#pragma unroll
                for (int v = 0; v < 4; v++) q += acc[p][jb][v] * acc[p][jb][v];
            q = wave_xor_sum_g(q);
            sv[p] = -0.5 * q + ck;
            if (sv[p] > mx[p]) best[p] = k;
            mx[p] = fmax(mx[p], sv[p]);
        }
        if ((k & 3) == g) {
#pragma unroll
            for (int p = 0; p < P; p++)
                if (!(sv[p] < mx[p] - alive_range<double>())) {
                    Rt[(size_t)k * npad + base + 16 * p + c] = sv[p];
                    stored[p] |= 1u << ((k >> 2) & 15);
                }
        }
    };
    // the named components ascending, then the others in `order`
    uint64_t rest = named;
    int it = 0;
    auto next = [&]() {
        if (rest) {
            const int k = __builtin_ctzll(rest);
            rest &= rest - 1;
            return k;
        }
        while (it < K) {
            const int k = order ? order[it] : it;
            it++;
            if (!((named >> k) & 1)) return k;
        }
        return -1;
    };
    for (int k = next(); k >= 0; k = next()) process(k);

    double llsum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int64_t pt = base + 16 * p + c;
        const bool valid = pt < n;
        const double lse = lse_posteriors(Rt, npad, pt, K, g, mx[p], valid, stored[p]);
        if (g == 0 && valid) llsum += lse;
        if (g == 0) ref[pt] = valid ? min(best[p], K - 1) : K;  // valid points sort first
    }
    if (lane == 0 && skipped) atomicAdd(nskip, (unsigned long long)skipped);
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
    const double r = block_sum(parts, n, s);
    if (threadIdx.x == 0) *out = r;
}

template <int NB, int P, bool PRUNE>
static hipError_t estep64_launch(fem_ctx *c, int *ref) {
    const size_t lds = (size_t)estep64_lds<NB>() * sizeof(double);
    const void *kern;
    if constexpr (NB <= 4) kern = (const void *)k_estep64_w3<NB, P, PRUNE>;
    else if constexpr (NB == 5) kern = (const void *)k_estep64_w2<NB, P, PRUNE>;
    else kern = (const void *)k_estep64<NB, P, PRUNE>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = (int)(c->npad / (64 * P));
    if constexpr (NB <= 4)
        hipLaunchKernelGGL((k_estep64_w3<NB, P, PRUNE>), dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->n,
                           c->npad, c->K, c->stage64, c->coef, c->Rt, c->ll_part, c->korder, ref, c->nskip);
    else if constexpr (NB == 5)
        hipLaunchKernelGGL((k_estep64_w2<NB, P, PRUNE>), dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->n,
                           c->npad, c->K, c->stage64, c->coef, c->Rt, c->ll_part, c->korder, ref, c->nskip);
    else
        hipLaunchKernelGGL((k_estep64<NB, P, PRUNE>), dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->n,
                           c->npad, c->K, c->stage64, c->coef, c->Rt, c->ll_part, c->korder, ref, c->nskip);
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, c->stream, c->ll_part,
                       (int64_t)blocks, c->ll_dev);
    return hipGetLastError();
}

// The pruned kernel (and the sorted slot order, sort_points) where it
// applies: K <= 64, d > 16, not switched off; else the plain kernel.
template <int NB, int P>
static hipError_t estep64_t(fem_ctx *c) {
    if constexpr (NB >= 2) {
        if (c->prune && c->ref && c->K <= 64) {
            if (hipError_t e = sort_points(c)) return e;
            if (hipError_t e = hipMemsetAsync(c->nskip, 0, sizeof(unsigned long long), c->stream)) return e;
            hipError_t e;
            if constexpr (NB <= 4) {
                // one wave at a time, fragments from the global stage blocks
                const int blocks = (int)(c->npad / (64 * P));
                hipLaunchKernelGGL((k_estep64_pruned<NB, P>), dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->n,
                                   c->npad, c->K, c->stage64, c->coef, c->Rt, c->ll_part, c->korder, c->ref,
                                   c->nskip);
                hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, c->stream, c->ll_part, (int64_t)blocks,
                                   c->ll_dev);
                e = hipGetLastError();
            } else {
                // the LDS-staged kernel with the workgroup's components first
                e = estep64_launch<NB, P, true>(c, c->ref);
            }
            if (e == hipSuccess) c->refs_valid = true;
            return e;
        }
    }
    return estep64_launch<NB, P, false>(c, nullptr);
}

// Point groups per wave of the FP64 E-step: 2 or 4 (4 only for d <= 64).
hipError_t launch_estep(fem_ctx *c) {
    if (c->prec == FEM_PREC_BF16X3) return launch_estepb(c);
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
template <int NB>
__global__ __launch_bounds__(256) void k_means64(const float *__restrict__ X,
                                                 const double *__restrict__ Rt, int64_t npad,
                                                 int KP, double *__restrict__ mpart) {
    means64_unit<NB>(blockIdx.x, X, Rt, npad, KP, mpart + (size_t)blockIdx.x * KP * (16 * NB + 1));
}

// raw != 0 writes the sums M_k themselves (k-means centroid sums).
__global__ __launch_bounds__(256) void k_means_finalize(int K, int d, int dpad, int KP, int chunks,
                                                        const double *__restrict__ mpart,
                                                        double *__restrict__ mu,
                                                        double *__restrict__ nk, int raw) {
    means_finalize_unit(blockIdx.x, blockIdx.y, d, dpad, KP, chunks, mpart, mu, nk, raw);
}

template <int NB>
static hipError_t means_t(fem_ctx *c) {
    hipLaunchKernelGGL(k_means64<NB>, dim3(c->m_chunks), dim3(256), 0, c->stream, c->Xe(), c->Rt,
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
    hipLaunchKernelGGL(k_means_finalize, dim3(c->K, (c->d + 63) / 64), dim3(256), 0, c->stream, c->K,
                       c->d, c->dpad, c->KP, c->m_chunks, c->mpart, mu_out, nk_out, raw);
    return hipGetLastError();
}

hipError_t launch_means(fem_ctx *c) {
    hipError_t e = (c->prec == FEM_PREC_FP64)   ? launch_means64_partials(c)
                   : (c->prec == FEM_PREC_FP32) ? launch_means32_partials(c)
                                                : launch_meansb_partials(c);
    if (e != hipSuccess) return e;
    return launch_means_finalize(c, c->mu, c->nk, 0);
}

__global__ __launch_bounds__(256) void k_cov_reduce(int K, int d, int NB, int KP, int chunks,
                                                    const double *__restrict__ cpart,
                                                    double *__restrict__ cov) {
    cov_reduce_unit(blockIdx.x, blockIdx.y, d, NB, KP, chunks, cpart, cov);
}

__global__ __launch_bounds__(256) void k_cov_finish(int K, int d, const double *__restrict__ nk, double g,
                                                    double ntotal, double *__restrict__ cov,
                                                    double *__restrict__ w) {
    __shared__ double str;
    cov_finish_unit(blockIdx.x, d, nk, g, ntotal, cov, w, &str);
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
    hipError_t e = (c->prec == FEM_PREC_FP64)   ? launch_cov64_partials(c)
                   : (c->prec == FEM_PREC_FP32) ? launch_cov32_partials(c)
                                                : launch_covb_partials(c);
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

// Slot i's posteriors to row perm[i] (i before the points were sorted).
__global__ void k_transpose_post(const double *__restrict__ Rt, int64_t npad, int64_t n, int K,
                                 double *__restrict__ out, int64_t ld, const int *__restrict__ perm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t row = perm ? perm[i] : i;
    for (int k = 0; k < K; k++) out[row * ld + k] = Rt[(size_t)k * npad + i];
}

// Class of each row: the first component of largest posterior, compared with
// > from -1 as EMAlgorithm::RetrieveClasses does (EM2/EMAlgorithm.cpp:107-122).
template <typename T>
__global__ void k_classes(const T *__restrict__ Rt, int64_t npad, int64_t n, int K, int *__restrict__ cls) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double best = -1.0;
    int bj = -1;
    for (int k = 0; k < K; k++) {
        const double r = (double)Rt[(size_t)k * npad + i];
        if (r > best) {
            best = r;
            bj = k;
        }
    }
    cls[i] = bj;
}

hipError_t launch_classes(const fem_ctx *c, int *cls_dev) {
    const dim3 grid((unsigned)((c->n + 255) / 256));
    if (c->prec == FEM_PREC_FP64)
        hipLaunchKernelGGL(k_classes<double>, grid, dim3(256), 0, c->stream, c->Rt, c->npad, c->n, c->K, cls_dev);
    else
        hipLaunchKernelGGL(k_classes<float>, grid, dim3(256), 0, c->stream, c->Rt32, c->npad, c->n, c->K, cls_dev);
    return hipGetLastError();
}

hipError_t launch_transpose_post(const fem_ctx *c, double *out_dev, int64_t ld) {
    if (c->prec != FEM_PREC_FP64) return launch_transpose_post32(c, out_dev, ld);
    hipLaunchKernelGGL(k_transpose_post, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0,
                       c->stream, c->Rt, c->npad, c->n, c->K, out_dev, ld, c->sorted ? c->perm : nullptr);
    return hipGetLastError();
}

}  // namespace fem
