// fem_kernels32.hip -- FP32 precision form of the E-step and M-step
// contractions on v_mfma_f32_16x16x4f32 (gfx950).  Posteriors are FP32; the
// per-chunk partial sums are flushed from FP32 accumulators into FP64 at fixed
// intervals and all cross-chunk reductions, the Cholesky and the log-det stay
// FP64 (shared with the FP64 form).
//
// Accumulator element v of lane l of v_mfma_f32_16x16x4f32 holds
// D[4*(l/16) + v][l%16]; A[i][k] comes from lane i + 16k, B[k][j] from lane
// j + 16k.
//
// Centring: with mu = mu_hi + mu_lo split into two floats, the E-step forms
// U^T (x - mu_hi) - U^T mu_lo (the second term precomputed in FP64 and used as
// the accumulator seed), so the centred value is accurate to FP32 rounding of
// the centred value rather than of |mu| (means reach |mu| ~ 20 sigma); the
// covariance pass centres with mu_hi alone (see k_cov32).
#include "fem_internal.h"
#include "fem_tiles.h"

#include <cmath>

namespace fem {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __host__ constexpr int pairs32(int NB) {
    int p = 0;
    for (int lb = 0; lb < 4 * NB; lb++) p += NB - lb / 4;
    return p;
}


// ---------------------------------------------------------------------------
// Stage (a), FP32.  Workgroup = 4 waves x P groups of 16 points.  The
// component's stage block (U_k fragments, mu_hi, -U_k^T mu_lo; written by the
// factorisation) is double-buffered in LDS: component k+1's block is read
// into registers while component k is processed and written to the idle
// buffer before the one barrier per component.  The B operand is
// x - mu_hi and the accumulators start at -U_k^T mu_lo, so D = U_k^T (x - mu).
// Components are processed in the order `order` (by descending weight) and a
// log-density is stored only while within alive_range of its point's
// running maximum, as in the FP64 E-step (fem_tiles.h); the storing lane
// keeps the bits of what it stored and the log-sum-exp sweep's lanes take
// their components' bits from it.
template <int NB, int P>
__global__ __launch_bounds__(256) void k_estep32(const float *__restrict__ X, int64_t n, int64_t npad,
                                                 int K, const float *__restrict__ stage,
                                                 const double *__restrict__ coef,
                                                 float *__restrict__ Rt, double *__restrict__ ll_part,
                                                 const int *__restrict__ order) {
    constexpr int DP = 16 * NB;
    constexpr int NPAIR = pairs32(NB);
    constexpr int STAGE = NPAIR * 64 + 2 * DP;  // a multiple of 4
    constexpr int STAGE4 = STAGE / 4;
    constexpr int PER = (STAGE4 + 255) / 256;
    extern __shared__ __align__(16) float lds32[];  // 2 * STAGE
    __shared__ double wsum[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int64_t base = (int64_t)blockIdx.x * (64 * P) + wave * (16 * P);

    float xr[P][4 * NB];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++)  // read once: non-temporal (L2 keeps the stage blocks)
            xr[p][lb] = __builtin_nontemporal_load(&X[(base + 16 * p + c) * DP + 4 * lb + g]);

    // stage blocks move in 16-byte pieces; piece indices past the block are
    // clamped to its last piece (same value to the same address)
    const float4 *stage4 = reinterpret_cast<const float4 *>(stage);
    float4 *lds4 = reinterpret_cast<float4 *>(lds32);
    {
        const float4 *first4 = stage4 + (size_t)order[0] * STAGE4;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = min(tid + 256 * q, STAGE4 - 1);
            lds4[e] = first4[e];
        }
    }
    __syncthreads();

    float mx = -INFINITY;   // of point (g mod P, c)
    uint64_t stored = 0;    // bit k: its log-density stored (K <= 64)

    for (int it = 0; it < K; it++) {
        const int k = order[it];
        const float *lu = lds32 + (it & 1) * STAGE;
        const float *lmh = lu + NPAIR * 64;
        const float *lml = lmh + DP;
        const bool more = it + 1 < K;
        const float4 *next4 = stage4 + (size_t)(more ? order[it + 1] : k) * STAGE4;
        float4 pre[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) pre[q] = next4[min(tid + 256 * q, STAGE4 - 1)];
        const float ck = (float)coef[k];
        v4f acc[P][NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            const float4 m = *reinterpret_cast<const float4 *>(lml + 16 * jb + 4 * g);
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][jb] = (v4f){m.x, m.y, m.z, m.w};
        }
        float mh[4 * NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const float4 m = *reinterpret_cast<const float4 *>(lmh + 4 * NB * g + 4 * q);
            mh[4 * q] = m.x, mh[4 * q + 1] = m.y, mh[4 * q + 2] = m.z, mh[4 * q + 3] = m.w;
        }
        int pr = 0;
        float4 a4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++) {
            float b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = xr[p][lb] - mh[lb];
#pragma unroll
            for (int jb = lb / 4; jb < NB; jb++) {
                if ((pr & 3) == 0) a4 = *reinterpret_cast<const float4 *>(lu + (pr >> 2) * 256 + lane * 4);
                const float a = (pr & 3) == 0 ? a4.x : (pr & 3) == 1 ? a4.y : (pr & 3) == 2 ? a4.z : a4.w;
                pr++;
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[p], acc[p][jb], 0, 0, 0);
            }
        }
        // squared norms summed over the lane groups with lane swaps, so that
        // lane (g, c) holds the log-density of point (g mod P, c)
        float q[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            q[p] = 0.f;
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
#pragma unroll
                for (int v = 0; v < 4; v++) q[p] += acc[p][jb][v] * acc[p][jb][v];
        }
        float sv;
        if constexpr (P == 4)
            sv = -0.5f * group_sums4_spread(q) + ck;
        else
            sv = -0.5f * group_sums2_spread(q) + ck;
        mx = fmaxf(mx, sv);
        {
            // unconditional (after the last component the idle buffer is
            // dead) and ahead of the branches below, so that the loads of the
            // next block stay in this basic block, ahead of the MFMAs
            float4 *dst = lds4 + ((it + 1) & 1) * STAGE4;
#pragma unroll
            for (int q = 0; q < PER; q++) dst[min(tid + 256 * q, STAGE4 - 1)] = pre[q];
        }
        if (g < P && (K > 64 || !(sv < mx - alive_range<float>()))) {
            Rt[(size_t)k * npad + base + 16 * g + c] = sv;
            stored |= 1ull << (k & 63);
        }
        // (the barrier also orders these stores before the sweep below,
        // whose lanes read entries other lanes stored)
        __syncthreads();
    }

    double llsum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int64_t pt = base + 16 * p + c;
        const bool valid = pt < n;
        const float mxp = __shfl(mx, 16 * p + c);
        const float lse = lse_posteriors(Rt, npad, pt, K, g, mxp, valid, sweep_bits(stored, 16 * p + c, g));
        if (g == 0 && valid) llsum += (double)lse;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) llsum += __shfl_xor(llsum, o);
    if (lane == 0) wsum[wave] = llsum;
    __syncthreads();
    if (tid == 0) ll_part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ __launch_bounds__(256) void k_reduce_sum32(const double *__restrict__ parts, int64_t n,
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
static hipError_t estep32_t(fem_ctx *c) {
    constexpr int STAGE = pairs32(NB) * 64 + 2 * 16 * NB;
    const size_t lds = (size_t)2 * STAGE * sizeof(float);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_estep32<NB, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = (int)(c->npad / (64 * P));
    hipLaunchKernelGGL((k_estep32<NB, P>), dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->n, c->npad,
                       c->K, c->stage32, c->coef, c->Rt32, c->ll_part, c->korder);
    hipLaunchKernelGGL(k_reduce_sum32, dim3(1), dim3(256), 0, c->stream, c->ll_part, (int64_t)blocks,
                       c->ll_dev);
    return hipGetLastError();
}

// Point groups per wave: estep_groups32 (fem_create picks 4 for d <= 32, else 2;
// 4 is available for d <= 64).
hipError_t launch_estep32(fem_ctx *c) {
    const bool p2 = c->estep_groups32 == 2;
    switch (c->NB) {
        case 1: return p2 ? estep32_t<1, 2>(c) : estep32_t<1, 4>(c);
        case 2: return p2 ? estep32_t<2, 2>(c) : estep32_t<2, 4>(c);
        case 3: return p2 ? estep32_t<3, 2>(c) : estep32_t<3, 4>(c);
        case 4: return p2 ? estep32_t<4, 2>(c) : estep32_t<4, 4>(c);
        case 5: return estep32_t<5, 2>(c);
        case 6: return estep32_t<6, 2>(c);
        case 7: return estep32_t<7, 2>(c);
        case 8: return estep32_t<8, 2>(c);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage (b), FP32 products: as k_means64 (four R values per lane read at once
// per 16 points, a lane's NB X values of a point in one load, feature NB j + jb
// in column j of block jb), FP32 accumulators flushed into FP64 every 1024
// points.
template <int NB>
__global__ __launch_bounds__(256) void k_means32(const float *__restrict__ X, const float *__restrict__ Rt,
                                                 int64_t npad, int KP, double *__restrict__ mpart) {
    constexpr int DP = 16 * NB;
    constexpr int CH = 4096, FLUSH = 1024;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int chunk = blockIdx.x;
    const int64_t i_begin = (int64_t)chunk * CH;
    const int KB = KP / 16;
    double *out = mpart + (size_t)chunk * KP * (DP + 1);
    for (int kb = wave; kb < KB; kb += 4) {
        const float *rrow = Rt + (size_t)(16 * kb + c) * npad + i_begin + 4 * g;
        const float *xrow = X + (i_begin + 4 * g) * DP + NB * c;
        double accd[NB][4];
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++) accd[jb][v] = 0.0;
        double nsd = 0.0;
        for (int s0 = 0; s0 < CH; s0 += FLUSH) {
            v4f acc[NB];
#pragma unroll
            for (int jb = 0; jb < NB; jb++) acc[jb] = (v4f){0.f, 0.f, 0.f, 0.f};
            float ns = 0.f;
#pragma unroll 2
            for (int s = s0; s < s0 + FLUSH; s += 16) {
                const float4 r4 = *reinterpret_cast<const float4 *>(rrow + s);
                const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
                ns += (rv[0] + rv[1]) + (rv[2] + rv[3]);
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    float xv[NB];
                    load_consecutive<NB>(xrow + (size_t)(s + m) * DP, xv);
#pragma unroll
                    for (int jb = 0; jb < NB; jb++)
                        acc[jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(rv[m], xv[jb], acc[jb], 0, 0, 0);
                }
            }
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
#pragma unroll
                for (int v = 0; v < 4; v++) accd[jb][v] += (double)acc[jb][v];
            nsd += (double)ns;
        }
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[(size_t)(16 * kb + 4 * g + v) * (DP + 1) + NB * c + jb] = accd[jb][v];
        nsd += __shfl_xor(nsd, 16);
        nsd += __shfl_xor(nsd, 32);
        if (g == 0) out[(size_t)(16 * kb + c) * (DP + 1) + DP] = nsd;
    }
}

template <int NB>
static hipError_t means32_t(fem_ctx *c) {
    hipLaunchKernelGGL(k_means32<NB>, dim3(c->m_chunks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad,
                       c->KP, c->mpart);
    return hipGetLastError();
}

hipError_t launch_means32_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return means32_t<1>(c);
        case 2: return means32_t<2>(c);
        case 3: return means32_t<3>(c);
        case 4: return means32_t<4>(c);
        case 5: return means32_t<5>(c);
        case 6: return means32_t<6>(c);
        case 7: return means32_t<7>(c);
        case 8: return means32_t<8>(c);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage (c), FP32 products: T = (x - mu_hi) * sqrt(r) / sqrt(N_k) in FP32
// (v_sqrt_f32), D += T_ib^T T_jb on FP32 MFMA, accumulators flushed into FP64
// every 2048 points (into registers for d <= 64, else into the wave's own
// partial slice in global memory).  Centring with mu_hi = float(mu) alone
// changes D by N_k mu_lo mu_lo^T plus a term proportional to
// sum_i r_ik (x_i - mu_k), both below 1e-9 of D.  Strided feature blocks as
// k_cov64 (a lane's NB features of a point in one load); lane group g takes
// points 4g..4g+3 of each 16, so its four R values are one 16-byte load.  For
// 48 < d <= 64 four register buffers of 16 points: the loads for 64 points
// ahead are issued once the current 16 are formed, before their MFMAs (two
// buffers measured 5% slower there; d <= 48 keeps two, where four measured no
// faster, and d > 64 one buffer and 16 points ahead).  Same grid and XCD remapping
// as k_cov64.
__device__ __host__ constexpr int cov32_parts(int NB) { return NB > 4 ? 2 : 1; }

template <int NB, int PART>
__device__ inline void cov32_body(int cg, int chunk, const float *__restrict__ X, const float *__restrict__ Rt,
                                  int64_t npad, int K, int d, const double *__restrict__ mu,
                                  const double *__restrict__ nk, int64_t chunk_rows, int KP,
                                  double *__restrict__ cpart) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov32_parts(NB);
    constexpr int U0 = upper_part_begin(NB, NPART, PART);
    constexpr int NU = upper_part_size(NB, NPART, PART);
    constexpr int FLUSH = 2048;
    constexpr bool REG = NB <= 4;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    const int64_t i0 = (int64_t)chunk * chunk_rows;
    const int64_t i1 = (i0 + chunk_rows < npad) ? i0 + chunk_rows : npad;
    float mh[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) mh[fb] = (NB * c + fb < d) ? (float)mu[(size_t)k * d + NB * c + fb] : 0.0f;
    const float inv = (float)(1.0 / sqrt(nk[k]));
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
    const float *rrow = Rt + (size_t)k * npad + 4 * g;
    const float *xrow = X + 4 * g * DP + NB * c;
    double accd[REG ? NU : 1][4];
    if (REG) {
#pragma unroll
        for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
            for (int v = 0; v < 4; v++) accd[u][v] = 0.0;
    }
    // H register buffers of 16 points each: the loads run H steps ahead
    constexpr int H = NB == 4 ? 4 : NB < 4 ? 2 : 1;
    float rv[H][4], xv[H][4][NB];
    auto load16 = [&](int h, int64_t i) {
        const float4 r4 = *reinterpret_cast<const float4 *>(rrow + i);
        rv[h][0] = r4.x, rv[h][1] = r4.y, rv[h][2] = r4.z, rv[h][3] = r4.w;
#pragma unroll
        for (int q = 0; q < 4; q++) load_consecutive<NB>(xrow + (size_t)(i + q) * DP, xv[h][q]);
    };
    // chunk bounds are multiples of 4096 points, so every step has a partner
    if (i0 < i1) {
#pragma unroll
        for (int h = 0; h < H; h++) load16(h, i0 + 16 * h);
    }
    bool first = true;
    for (int64_t s0 = i0; s0 < i1; s0 += FLUSH) {
        const int64_t s1 = (s0 + FLUSH < i1) ? s0 + FLUSH : i1;
        v4f acc[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) acc[u] = (v4f){0.f, 0.f, 0.f, 0.f};
        for (int64_t i = s0; i < s1; i += 16 * H) {
#pragma unroll
            for (int h = 0; h < H; h++) {
                float t[4][NB];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float sw = __builtin_amdgcn_sqrtf(rv[h][q]) * inv;
#pragma unroll
                    for (int fb = 0; fb < NB; fb++)
                        if (upper_part_uses(NB, NPART, PART, fb, false)) t[q][fb] = (xv[h][q][fb] - mh[fb]) * sw;
                }
                // unconditional (past the chunk end it re-reads the last 16
                // points, unused), so the buffers need no register copies
                const int64_t ni = i + 16 * (H + h);
                load16(h, ni < i1 ? ni : i1 - 16);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    int u = 0;
#pragma unroll
                    for (int ib = 0; ib < NB; ib++)
#pragma unroll
                        for (int jb = ib; jb < NB; jb++, u++) {
                            if (upper_part_of(u, UB, NPART) != PART) continue;
                            acc[u - U0] =
                                __builtin_amdgcn_mfma_f32_16x16x4f32(t[q][ib], t[q][jb], acc[u - U0], 0, 0, 0);
                        }
                }
            }
        }
        if (REG) {
#pragma unroll
            for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
                for (int v = 0; v < 4; v++) accd[u][v] += (double)acc[u][v];
        } else {
            // each wave owns its partial slice: plain read-add-write, four
            // blocks at a time (the compiler barrier keeps the compiler from
            // loading every block's partials at once, which would take the
            // register file and leave one wave per SIMD)
#pragma unroll
            for (int u = 0; u < NU; u++) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double *o = out + (size_t)(U0 + u) * 256 + (4 * g + v) * 16 + c;
                    *o = (first ? 0.0 : *o) + (double)acc[u][v];
                }
                if ((u & 3) == 3) asm volatile("" ::: "memory");
            }
        }
        first = false;
    }
    if (REG) {
#pragma unroll
        for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[(size_t)(U0 + u) * 256 + (4 * g + v) * 16 + c] = accd[u][v];
    }
}

// For d > 64 the upper blocks of a component are split over two waves
// (cov32_parts; the part varies fastest in the grid), and with the flush
// taken four blocks at a time each wave fits in half the register file
// (capped for two waves per SIMD): 20% faster at d = 80 and 96, 14% at
// d = 112 and 4% at d = 128 than one wave per component.
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_cov32(const float *__restrict__ X, const float *__restrict__ Rt,
                                               int64_t npad, int K, int KP, int d,
                                               const double *__restrict__ mu,
                                               const double *__restrict__ nk, int64_t chunk_rows,
                                               double *__restrict__ cpart) {
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    constexpr int NPART = cov32_parts(NB);
    const int part = logical % NPART, rest = logical / NPART;
    const int cg = rest % KG, chunk = rest / KG;
    if (part == 0) cov32_body<NB, 0>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart);
    else if constexpr (NPART > 1) cov32_body<NB, 1>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart);
}

// The same over the nonzero posteriors only (as k_cov64_sparse, whose exact-
// zero argument holds here with FP32's underflow): per wave and segment of
// kCovSeg32 points (segments strided over all points, as there), lane l reads points 4l..4l+3 of each 256 (one 16-byte
// load), the points with r != 0 are listed in point order in the wave's LDS
// slice, and the products run over the list four entries per MFMA (lane
// group g takes entry 4j + g; entries past the end take r = 0).  The FP32
// accumulators are flushed into FP64 before a segment's list would take them
// past 2048 terms, the dense form's bound.
constexpr int kCovSeg32 = 1024;

template <int NB, int PART>
__device__ inline void cov32_sparse(int cg, int chunk, const float *__restrict__ X, const float *__restrict__ Rt,
                                    int64_t npad, int K, int d, const double *__restrict__ mu,
                                    const double *__restrict__ nk, int64_t chunk_rows, int KP,
                                    double *__restrict__ cpart, int *lidx, float *lrv) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov32_parts(NB);
    constexpr int U0 = upper_part_begin(NB, NPART, PART);
    constexpr int NU = upper_part_size(NB, NPART, PART);
    constexpr int FLUSH = 2048;
    constexpr bool REG = NB <= 4;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    // segments chunk, chunk + chunks, ... (as cov64_sparse)
    const int64_t nseg = npad / kCovSeg32, nch = (npad + chunk_rows - 1) / chunk_rows;
    float mh[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) mh[fb] = (NB * c + fb < d) ? (float)mu[(size_t)k * d + NB * c + fb] : 0.0f;
    const float inv = (float)(1.0 / sqrt(nk[k]));
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
    const float *rrow = Rt + (size_t)k * npad;
    int *wi = lidx + wave * kCovSeg32;
    float *wr = lrv + wave * kCovSeg32;
    double accd[REG ? NU : 1][4];
    if (REG) {
#pragma unroll
        for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
            for (int v = 0; v < 4; v++) accd[u][v] = 0.0;
    }
    bool first = true;
    v4f acc[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) acc[u] = (v4f){0.f, 0.f, 0.f, 0.f};
    int pending = 0;  // terms in acc since the last flush
    auto flush = [&]() {
        if (REG) {
#pragma unroll
            for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
                for (int v = 0; v < 4; v++) accd[u][v] += (double)acc[u][v];
        } else {
#pragma unroll
            for (int u = 0; u < NU; u++) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double *o = out + (size_t)(U0 + u) * 256 + (4 * g + v) * 16 + c;
                    *o = (first ? 0.0 : *o) + (double)acc[u][v];
                }
                if ((u & 3) == 3) asm volatile("" ::: "memory");
            }
        }
        first = false;
#pragma unroll
        for (int u = 0; u < NU; u++) acc[u] = (v4f){0.f, 0.f, 0.f, 0.f};
        pending = 0;
    };

    constexpr int NJ = kCovSeg32 / 256;
    v4f rs[NJ];
    auto loadseg = [&](int64_t s) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
            rs[j] = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(rrow + s + 256 * j + 4 * lane));
    };
    // entries e0 + 4 q + g (q < 2; called with e0 < m): weight and features
    auto fetch = [&](int64_t s, int m, int e0, float (&sw)[2], float (&xv)[2][NB]) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int e = e0 + 4 * q + g;
            const bool ok = e < m;
            const int ec = ok ? e : 0;
            const int off = wi[ec];
            sw[q] = ok ? __builtin_amdgcn_sqrtf(wr[ec]) * inv : 0.0f;
            load_consecutive<NB>(X + (size_t)(s + off) * DP + NB * c, xv[q]);
        }
    };
    auto products = [&](const float (&sw)[2], const float (&xv)[2][NB]) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            float t[NB];
#pragma unroll
            for (int fb = 0; fb < NB; fb++)
                if (upper_part_uses(NB, NPART, PART, fb, false)) t[fb] = (xv[q][fb] - mh[fb]) * sw[q];
            int u = 0;
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = ib; jb < NB; jb++, u++) {
                    if (upper_part_of(u, UB, NPART) != PART) continue;
                    acc[u - U0] = __builtin_amdgcn_mfma_f32_16x16x4f32(t[ib], t[jb], acc[u - U0], 0, 0, 0);
                }
        }
    };

    // npad is a multiple of 4096 points: every segment is whole
    if (chunk < nseg) loadseg((int64_t)chunk * kCovSeg32);
    for (int64_t sg = chunk; sg < nseg; sg += nch) {
        const int64_t s = sg * kCovSeg32;
        int m = 0;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            bool nz[4];
            uint64_t b[4];
            int below = 0;
#pragma unroll
            for (int w = 0; w < 4; w++) {
                nz[w] = rs[j][w] != 0.0f;
                b[w] = __ballot(nz[w]);
                below += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b[w] >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)b[w], 0u));
            }
            int at = m + below;
            const int off = 256 * j + 4 * lane;
#pragma unroll
            for (int w = 0; w < 4; w++) {
                if (nz[w]) {
                    wi[at] = off + w;
                    wr[at] = rs[j][w];
                }
                at += nz[w];
                m += __builtin_popcountll(b[w]);
            }
        }
        if (sg + nch < nseg) loadseg((sg + nch) * kCovSeg32);
        // the list pads to whole groups of 4 terms per MFMA
        if (pending + ((m + 3) & ~3) > FLUSH) flush();
        pending += (m + 3) & ~3;
        float swA[2], swB[2], xvA[2][NB], xvB[2][NB];
        if (m > 0) fetch(s, m, 0, swA, xvA);
        for (int e0 = 0; e0 < m; e0 += 16) {
            if (e0 + 8 < m) fetch(s, m, e0 + 8, swB, xvB);
            products(swA, xvA);
            if (e0 + 16 < m) fetch(s, m, e0 + 16, swA, xvA);
            if (e0 + 8 < m) products(swB, xvB);
        }
    }
    flush();
    if (REG) {
#pragma unroll
        for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[(size_t)(U0 + u) * 256 + (4 * g + v) * 16 + c] = accd[u][v];
    }
}

template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_cov32_sparse(
    const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K, int KP, int d,
    const double *__restrict__ mu, const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    __shared__ int lidx[4 * kCovSeg32];
    __shared__ float lrv[4 * kCovSeg32];
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    constexpr int NPART = cov32_parts(NB);
    const int part = logical % NPART, rest = logical / NPART;
    const int cg = rest % KG, chunk = rest / KG;
    if (part == 0) cov32_sparse<NB, 0>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart, lidx, lrv);
    else if constexpr (NPART > 1)
        cov32_sparse<NB, 1>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart, lidx, lrv);
}

template <int NB>
static hipError_t cov32_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks * cov32_parts(NB);
    if (c->cov_sparse)
        hipLaunchKernelGGL(k_cov32_sparse<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad,
                           c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    else
        hipLaunchKernelGGL(k_cov32<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad, c->K,
                           c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    return hipGetLastError();
}

hipError_t launch_cov32_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return cov32_t<1>(c);
        case 2: return cov32_t<2>(c);
        case 3: return cov32_t<3>(c);
        case 4: return cov32_t<4>(c);
        case 5: return cov32_t<5>(c);
        case 6: return cov32_t<6>(c);
        case 7: return cov32_t<7>(c);
        case 8: return cov32_t<8>(c);
    }
    return hipErrorInvalidValue;
}

__global__ void k_transpose_post32(const float *__restrict__ Rt, int64_t npad, int64_t n, int K,
                                   double *__restrict__ out, int64_t ld, const int *__restrict__ perm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t row = perm ? perm[i] : i;
    for (int k = 0; k < K; k++) out[row * ld + k] = (double)Rt[(size_t)k * npad + i];
}

hipError_t launch_transpose_post32(const fem_ctx *c, double *out_dev, int64_t ld) {
    hipLaunchKernelGGL(k_transpose_post32, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream,
                       c->Rt32, c->npad, c->n, c->K, out_dev, ld, c->sorted ? c->perm : nullptr);
    return hipGetLastError();
}

}  // namespace fem
