// fem_tiles.h -- workgroup-level device functions of the FP64 EM kernels
// (gfx950).  Each function is one unit of work of one workgroup of 256 threads
// (4 waves) and is called uniformly by all its threads; the single-fit kernels
// (fem_kernels.hip) run one unit per workgroup, the batched persistent kernels
// (fem_batch.hip) loop over units of many fits.
//
// MFMA 16x16x4 operand layout (wave64, lane l): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k.  Accumulator element v of lane l holds
// D[4*v + l/16][l%16] for v_mfma_f64_16x16x4f64.  v_mfma_f64_4x4x4f64 (four
// 4x4x4 blocks b): A_b[i][k] from lane i + 4b + 16k, B_b[k][j] from lane
// j + 4b + 16k, D_b[i][j] in lane j + 4b + 16i.  So with the same A in every
// block, B is the 16x16x4 B operand (column c = j + 4b) and the result is
// rows 4v..4v+3 of the 16x16x4 product with a 16-row A in element v of the
// same lanes.  (Both checked on gfx950 with lane-layout probes.)
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace fem {

typedef double v4d __attribute__((ext_vector_type(4)));

// The upper block pairs (ib <= jb, row-major index u) of one component split
// into NPART contiguous parts of equal count, one wave each.
__device__ __host__ constexpr int upper_part_of(int u, int UB, int NPART) { return (u * NPART) / UB; }

__device__ __host__ constexpr int upper_part_begin(int NB, int NPART, int PART) {
    const int UB = NB * (NB + 1) / 2;
    for (int u = 0; u < UB; u++)
        if (upper_part_of(u, UB, NPART) == PART) return u;
    return UB;
}

__device__ __host__ constexpr int upper_part_size(int NB, int NPART, int PART) {
    const int UB = NB * (NB + 1) / 2;
    int s = 0;
    for (int u = 0; u < UB; u++) s += upper_part_of(u, UB, NPART) == PART;
    return s;
}

// Whether feature block fb is an operand of a block pair of part PART: as
// the row block ib (row_only) or as either block.
__device__ __host__ constexpr bool upper_part_uses(int NB, int NPART, int PART, int fb, bool row_only) {
    const int UB = NB * (NB + 1) / 2;
    int u = 0;
    for (int ib = 0; ib < NB; ib++)
        for (int jb = ib; jb < NB; jb++, u++)
            if (upper_part_of(u, UB, NPART) == PART && (ib == fb || (!row_only && jb == fb))) return true;
    return false;
}

__device__ __host__ constexpr int pairs_of(int NB) {
    int p = 0;
    for (int lb = 0; lb < 4 * NB; lb++) p += NB - lb / 4;
    return p;
}

// FP64 E-step operands of U_k^T (lower triangular) for NB 16-feature blocks:
// the (lb, jb) pairs of a 4-feature column strip lb and a 16-row block
// jb > lb/4 are whole 16x16x4 products; in the diagonal block jb = lb/4 only
// the 4-row quads v >= lb%4 touch the triangle, one 4x4x4 product each.
__device__ __host__ constexpr int offd_pairs_of(int NB) { return 2 * NB * (NB - 1); }
__device__ __host__ constexpr int diag_quads_of(int NB) { return 10 * NB; }
// Doubles of one component's FP64 stage block: fragments, quads, -U^T mu.
__device__ __host__ constexpr int stage64_words_of(int NB) {
    return offd_pairs_of(NB) * 64 + diag_quads_of(NB) * 16 + 16 * NB;
}

// Sum over the four 16-lane groups (lanes l, l^16, l^32, l^48) as
// (v_0 + v_1) + (v_2 + v_3) for groups 0..3, in every lane.  The FP32 form
// uses LDS permutes; the FP64 form uses the gfx950 lane-swap instructions:
// permlane16_swap(v, v) returns (v of the even group, v of the odd group) of
// each 32-lane half, permlane32_swap(v, v) returns (v of the lower half, v of
// the upper half), so each step is one swap of each 32-bit half and one add.
// (The 32-bit swap builtin is not used for a single FP32 value: this
// compiler combined its two results into one for a single-register operand.
// The FP32 E-steps sum two or four values at once with group_sums2/4_spread
// below; this form remains for the log-sum-exp sweep.)
__device__ inline float group_sum4(float v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

__device__ inline double group_sum4(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    const auto al = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto ah = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double s = __builtin_bit_cast(double, ((unsigned long long)ah[0] << 32) | al[0]) +
                     __builtin_bit_cast(double, ((unsigned long long)ah[1] << 32) | al[1]);
    const unsigned long long us = __builtin_bit_cast(unsigned long long, s);
    const unsigned slo = (unsigned)us, shi = (unsigned)(us >> 32);
    const auto bl = __builtin_amdgcn_permlane32_swap(slo, slo, false, false);
    const auto bh = __builtin_amdgcn_permlane32_swap(shi, shi, false, false);
    return __builtin_bit_cast(double, ((unsigned long long)bh[0] << 32) | bl[0]) +
           __builtin_bit_cast(double, ((unsigned long long)bh[1] << 32) | bl[1]);
}

__device__ inline double wave_xor_sum_g(double v) { return group_sum4(v); }

// Sums of q[p] over the four 16-lane groups for p = 0..3 at once, with the
// gfx950 lane swaps: lane (g, c) gets the total of q[g] over lanes c, c + 16,
// c + 32, c + 48, added as (g0 + g1) + (g2 + g3) like group_sum4.
// permlane16_swap(x, y) returns x with its odd rows (16 lanes) replaced by y's
// even rows and y with its even rows replaced by x's odd rows;
// permlane32_swap(x, y) returns x with its upper half replaced by y's lower
// half and y with its lower half replaced by x's upper half.
__device__ inline float group_sums4_spread(const float (&q)[4]) {
    auto u = [](float v) { return __builtin_bit_cast(unsigned, v); };
    auto f = [](unsigned v) { return __builtin_bit_cast(float, v); };
    const auto s01 = __builtin_amdgcn_permlane16_swap(u(q[0]), u(q[1]), false, false);
    const auto s23 = __builtin_amdgcn_permlane16_swap(u(q[2]), u(q[3]), false, false);
    // rows of a: q0 (r0 + r1), q1 (r0 + r1), q0 (r2 + r3), q1 (r2 + r3); b likewise
    const float a = f(s01[0]) + f(s01[1]);
    const float b = f(s23[0]) + f(s23[1]);
    const auto t = __builtin_amdgcn_permlane32_swap(u(a), u(b), false, false);
    return f(t[0]) + f(t[1]);
}

// The same for two values: lane (g, c) gets the total of q[g & 1].  The
// second swap pairs a with a copy of itself in another register (a plain
// move: passed the same value twice, the swap builtin's two results were
// combined into one register by this compiler).
__device__ inline float group_sums2_spread(const float (&q)[2]) {
    auto u = [](float v) { return __builtin_bit_cast(unsigned, v); };
    auto f = [](unsigned v) { return __builtin_bit_cast(float, v); };
    const auto s01 = __builtin_amdgcn_permlane16_swap(u(q[0]), u(q[1]), false, false);
    // rows of a: q0 (r0 + r1), q1 (r0 + r1), q0 (r2 + r3), q1 (r2 + r3)
    const float a = f(s01[0]) + f(s01[1]);
    float b;
    asm("v_mov_b32 %0, %1" : "=v"(b) : "v"(a));
    const auto t = __builtin_amdgcn_permlane32_swap(u(a), u(b), false, false);
    return f(t[0]) + f(t[1]);
}

// Fixed-order sum of p[0], p[stride], ..., with 4 interleaved accumulators.
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

// ---------------------------------------------------------------------------
// Stage (d) for component k (EM2/NormalDensity.cpp:43-65; MatrixEM.cpp:83-87).
// Lower-symmetrise, right-looking Cholesky in LDS A[d*d] (the same operation
// sequence per element as column Cholesky-Crout), log-det, the columns of
// L^-1 by forward substitution stored as rows of U = L^-T, then the E-step
// operands as one stage block per component, [16x16x4 fragments of the
// off-diagonal pairs | 4x4x4 fragments of the diagonal quads | -U^T mu (dpad)]
// (stage64_words_of), and, when stage32 is
// given, the FP32 form's per-component stage block
// [fragments (npairs x 64) | mu_hi (dpad) | -U^T mu_lo (dpad)] with
// mu_hi = float(mu), mu_lo = mu - mu_hi; mu_hi is stored lane-group major
// (feature 4 lb + g at g dpad/4 + lb) so that a lane reads its features in
// 16-byte loads.  fail_s: one LDS int.
__device__ inline void factor_unit(int k, int d, int dpad, int NB, int npairs, const double *__restrict__ cov,
                                   const double *__restrict__ w, const double *__restrict__ mu,
                                   double *__restrict__ logdet, int *__restrict__ info, double *__restrict__ coef,
                                   double *__restrict__ U, double *__restrict__ stage64, float *__restrict__ stage32,
                                   double *A, int *fail_s) {
    const int tid = threadIdx.x;
    const double *C = cov + (size_t)k * d * d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        int i = e / d, j = e % d;
        A[e] = (i >= j) ? C[(size_t)i * d + j] : C[(size_t)j * d + i];
    }
    if (tid == 0) *fail_s = 0;
    __syncthreads();
    for (int j = 0; j < d; j++) {
        if (tid == 0) {
            double p = A[j * d + j];
            if (!(p > 0.0)) *fail_s = j + 1;
            else A[j * d + j] = sqrt(p);
        }
        __syncthreads();
        if (*fail_s) break;
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
    const int fail = *fail_s;
    __syncthreads();  // every thread has read fail before the LDS int is reused
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
    // FP32 stage: every (lb, jb >= lb/4) pair as a 16x16x4 fragment, four
    // pairs interleaved per lane per 16 bytes (npairs = 2 NB (NB + 1) is a
    // multiple of 4) so the E-step reads four with one 16-byte LDS load
    if (stage32)
        for (int e = tid; e < npairs * 64; e += blockDim.x) {
            int p = e >> 6, lane = e & 63;
            int lb = 0, rem = p;
            while (rem >= NB - lb / 4) {
                rem -= NB - lb / 4;
                lb++;
            }
            int jb = lb / 4 + rem;
            double v = Uk[(size_t)(4 * lb + (lane >> 4)) * dpad + 16 * jb + (lane & 15)];
            stage32[(size_t)k * (npairs * 64 + 2 * dpad) + (p >> 2) * 256 + lane * 4 + (p & 3)] = (float)v;
        }
    // FP64 stage (stage64_words_of(NB) doubles): the off-diagonal pairs
    // (jb > lb/4) as 16x16x4 fragments, two interleaved per lane per 16 bytes;
    // then the diagonal blocks' quads (lb, v >= lb%4) as 4x4x4 fragments of 16
    // slots s = i + 4 kk holding U^T[16 (lb/4) + 4v + i][4 lb + kk], two
    // interleaved per slot; then -U^T mu
    const int noff = offd_pairs_of(NB), ndiag = diag_quads_of(NB);
    double *sk = stage64 + (size_t)k * stage64_words_of(NB);
    for (int e = tid; e < noff * 64; e += blockDim.x) {
        int p = e >> 6, lane = e & 63;
        int lb = 0, rem = p;
        while (rem >= NB - 1 - lb / 4) {
            rem -= NB - 1 - lb / 4;
            lb++;
        }
        int jb = lb / 4 + 1 + rem;
        sk[(p >> 1) * 128 + lane * 2 + (p & 1)] = Uk[(size_t)(4 * lb + (lane >> 4)) * dpad + 16 * jb + (lane & 15)];
    }
    for (int e = tid; e < ndiag * 16; e += blockDim.x) {
        int q = e >> 4, s = e & 15;
        int lb = 0, rem = q;
        while (rem >= 4 - lb % 4) {
            rem -= 4 - lb % 4;
            lb++;
        }
        const int v = lb % 4 + rem;
        sk[noff * 64 + (q >> 1) * 32 + s * 2 + (q & 1)] =
            Uk[(size_t)(4 * lb + (s >> 2)) * dpad + 16 * (lb / 4) + 4 * v + (s & 3)];
    }
    for (int j = tid; j < dpad; j += blockDim.x) {
        double s = 0.0, slo = 0.0;
        if (j < d)
            for (int f = 0; f <= j; f++) {
                const double m = mu[(size_t)k * d + f];
                s += Uk[(size_t)f * dpad + j] * m;
                slo += Uk[(size_t)f * dpad + j] * (m - (double)(float)m);
            }
        sk[noff * 64 + ndiag * 16 + j] = -s;
        if (stage32) {
            float *st = stage32 + (size_t)k * (npairs * 64 + 2 * dpad) + npairs * 64;
            st[(j & 3) * (dpad >> 2) + (j >> 2)] = (j < d) ? (float)mu[(size_t)k * d + j] : 0.0f;
            st[dpad + j] = (float)(-slo);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Log-sum-exp and posteriors of one point (EM2/MatrixEM.cpp:39-51) whose K
// log-densities sit in the component-major Rt; the lane of group g owns
// components g, g+4, ... (the four groups of a 16-lane column share the
// point).  mx: the row max.  With K <= 64 the lane's 16 entries are read in
// one batch of independent loads, e_k = exp(s_k - mx) is formed once and the
// posteriors are e_k / sum; larger K streams the entries twice,
// r_k = exp(s_k - lse).  Invalid (padding) points get zero posteriors.
// With K <= 64, bit i of `stored` says whether the entry of component
// g + 4i was stored: the E-steps store a log-density only while it is within
// alive_range<T>() of the point's running maximum, and every other entry
// lies at least that far below the final maximum, where exp underflows to
// +0: it is taken as exp(-inf) = +0 without being read, the value its
// exponential has, so the posteriors and the sum are those of reading every
// entry.  The posteriors are stored non-temporally (streamed past L2, which
// then keeps the E-step's stage blocks).  Returns lse = log(sum) + mx for
// every lane of the column.
__device__ inline float exp_of(float v) { return expf(v); }
__device__ inline double exp_of(double v) { return exp(v); }
// Below the maximum by more than this, exp underflows to +0 (FP64: below
// -745.2; FP32: below -103.98, or -87.4 with denormals flushed).
template <typename T>
__device__ __host__ constexpr T alive_range() { return sizeof(T) == 8 ? T(760) : T(110); }
__device__ inline float log_of(float v) { return logf(v); }
__device__ inline double log_of(double v) { return log(v); }

template <typename T>
__device__ __attribute__((always_inline)) inline T lse_posteriors(T *__restrict__ Rt, int64_t npad, int64_t pt, int K, int g, T mx, bool valid,
                                                             unsigned stored = 0xFFFFu) {
    constexpr int KQ = 16;
    if (K <= 4 * KQ) {
        T e[KQ];
#pragma unroll
        for (int i = 0; i < KQ; i++) {
            const int k = g + 4 * i;
            e[i] = (k < K && ((stored >> i) & 1u)) ? Rt[(size_t)k * npad + pt] : T(-INFINITY);
        }
        T se = T(0);
#pragma unroll
        for (int i = 0; i < KQ; i++) {
            e[i] = g + 4 * i < K ? exp_of(e[i] - mx) : T(0);
            se += e[i];
        }
        se = group_sum4(se);
        const T inv = T(1) / se;
#pragma unroll
        for (int i = 0; i < KQ; i++) {
            const int k = g + 4 * i;
            if (k < K) __builtin_nontemporal_store(valid ? e[i] * inv : T(0), &Rt[(size_t)k * npad + pt]);
        }
        return log_of(se) + mx;
    }
    T se = T(0);
    for (int k = g; k < K; k += 4) se += exp_of(Rt[(size_t)k * npad + pt] - mx);
    se = group_sum4(se);
    const T lse = log_of(se) + mx;
    for (int k = g; k < K; k += 4) {
        T *r = Rt + (size_t)k * npad + pt;
        *r = valid ? exp_of(*r - lse) : T(0);
    }
    return lse;
}

// The sweep's stored bits for the lane of group g (components g + 4i, bit i)
// from the 64 bits (bit k: component k) kept by the lane src that stored the
// point's log-densities (the FP32 and BF16x3 E-steps, K <= 64).
__device__ inline unsigned group_bits(uint64_t sb, int g) {
    unsigned st = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) st |= (unsigned)((sb >> (g + 4 * i)) & 1u) << i;
    return st;
}

__device__ inline unsigned sweep_bits(uint64_t stored, int src, int g) {
    const unsigned lo = __shfl((unsigned)stored, src), hi = __shfl((unsigned)(stored >> 32), src);
    return group_bits(((uint64_t)hi << 32) | lo, g);
}

// ---------------------------------------------------------------------------
// Stage (a), FP64, one tile of 64P points (EM2/MatrixEM.cpp:39-87).
// Each wave owns P groups of 16 points.  Lane (g = l/16, c = l%16) holds, for
// point c of each group, features 4*lb + g (lb = 0..4NB-1), loaded as FP32
// and widened to FP64 once (the compiler keeps them widened across the
// component loop): the B operand of every MFMA.  Per component and strip lb
// of four input features, the 16-row blocks jb > lb/4 below the diagonal
// block are whole v_mfma_f64_16x16x4f64 products, and in the diagonal block
// jb = lb/4 the 4-row quads v >= lb%4 (the only ones the triangle reaches)
// are v_mfma_f64_4x4x4f64 products with the quad's 4x4 piece of U_k^T in
// every block, accumulated into element v of the same accumulators: 10 NB
// quads of 512 flops and 2 NB (NB - 1) pairs of 2048 flops, 1.06x the
// N K d^2 flops of the triangle at d = 64 where whole 16x16x4 products of
// every pair touching it are 1.25x.  Component k's stage block (fragments,
// quads, then -m_k = -U_k^T mu_k) is double-buffered in LDS (estep64_lds
// doubles): component k+1's block is copied from global memory straight
// into the idle buffer (global_load_lds_dwordx4) while component k is
// processed, and the one barrier per component publishes it.  Copied
// through registers instead, the compiler sank those loads behind the MFMAs
// and waited on each, which one wave per SIMD (d > 80, where the buffers fill
// the LDS) cannot hide: the direct copy measured 30% faster at d = 96..128
// and 3% at d = 64.  Accumulators start at -m_k so that
// D = U^T (x - mu).  log-densities go to the component-major Rt; the row max is
// tracked as they are produced; a second sweep over the lane's own entries
// forms the log-sum-exp and the posteriors (lse_posteriors).  The tile's LL is
// written to *ll_out by thread 0.  wsum: 4 LDS doubles.
template <int NB>
__device__ __host__ constexpr int estep64_stage() { return stage64_words_of(NB); }
// LDS doubles of the two stage buffers, each padded to whole waves of 64
// 16-byte pieces: a direct copy round writes one piece per lane, 256 per
// round, and in the last round only the waves whose pieces start inside the
// block copy (at d = 96 the two buffers then take 78 KB instead of 80 KB and
// two workgroups fit the 160 KB of LDS)
template <int NB>
__device__ __host__ constexpr int estep64_lds() { return 2 * ((estep64_stage<NB>() / 2 + 63) / 64) * 128; }

// order (K entries, or null for 0..K-1) is the order the components are
// processed in; it changes no result (the row max and the component-ordered
// sums do not depend on it), only which log-densities are within
// alive_range of the running maximum when they are formed, and so stored:
// the context passes the components by descending weight, so most points
// meet their own component early.
//
// Pruning (PRUNE, with ref non-null, K <= 64, NB >= 2; lds_int: 128 ints;
// the instance without it is the plain kernel): ref holds per
// point the component of its largest log-density in the previous E-step (K:
// none).  The components some point of the tile names come first (ascending),
// then the others in `order`.  Per wave and component, the 10 quads of the
// first 16-row block are formed first; their squared norm q0 is a lower bound
// of the point's q, so ub = coef_k - q0 / 2 bounds its log-density from above
// (the rounded sums are monotone: q's in-lane sum starts with q0's terms in
// the same order and adds non-negative ones, and the cross-group sums are the
// same additions).  When ub < running max - alive_range for every point of
// the wave, the unpruned kernel would neither store this log-density nor
// change the maximum, so the rest of the component is skipped and the
// outputs stay bit for bit those of the same order unpruned.  The component
// of largest log-density (the first in processing order) goes back to ref,
// and *nskip counts the skipped (wave, component) pairs.
template <int NB, int P, bool PRUNE = false>
__device__ inline void estep64_unit(int64_t tile, const float *__restrict__ X, int64_t n, int64_t npad, int K,
                                    const double *__restrict__ stage, const double *__restrict__ coef,
                                    double *__restrict__ Rt,
                                    double *__restrict__ ll_out, double *lds, double *wsum,
                                    const int *__restrict__ order = nullptr, int *__restrict__ ref = nullptr,
                                    int *lds_int = nullptr, unsigned long long *nskip = nullptr) {
    constexpr int DP = 16 * NB;
    constexpr int NOFF = offd_pairs_of(NB);
    constexpr int NDIAG = diag_quads_of(NB);
    constexpr int STAGE = estep64_stage<NB>();  // even
    constexpr int STAGE2 = STAGE / 2;
    constexpr int PER = (STAGE2 + 255) / 256;
    constexpr int BUF2 = (STAGE2 + 63) / 64 * 64;  // 16-byte pieces per LDS buffer
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, c = lane & 15;
    const int slot = (lane & 3) + 4 * g;  // this lane's 4x4x4 A slot i + 4 kk
    const int64_t base = tile * (64 * P) + wave * (16 * P);

    // the tile's points, read once: non-temporal, like the posteriors, so
    // L2 keeps the stage blocks every workgroup re-reads (fetches per launch
    // 2.66 -> 1.81 GB at N = 4M, d = K = 64)
    float xr[P][4 * NB];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int lb = 0; lb < 4 * NB; lb++)
            xr[p][lb] = __builtin_nontemporal_load(&X[(base + 16 * p + c) * DP + 4 * lb + g]);

    // stage blocks move in 16-byte pieces; piece indices past the block are
    // clamped to its last piece (same value to the same address)
    const double2 *stage2 = reinterpret_cast<const double2 *>(stage);
    double2 *lds2 = reinterpret_cast<double2 *>(lds);
    const bool prune = PRUNE && NB >= 2 && ref != nullptr && K <= 64;
    if (prune) {
        // the tile's named components first (ascending), then the others
        int *flag = lds_int, *sorder = lds_int + 64;
        if (tid < 64) flag[tid] = 0;
        __syncthreads();
        if (tid < 64 * P) {
            const int r = ref[tile * (64 * P) + tid];
            if (r >= 0 && r < K) flag[r] = 1;
        }
        __syncthreads();
        if (wave == 0) {
            const bool f = lane < K && flag[lane];
            const uint64_t fm = __ballot(f);
            if (f) sorder[__builtin_amdgcn_mbcnt_hi((unsigned)(fm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)fm, 0u))] = lane;
            const int nf = __builtin_popcountll(fm);
            const int kk = lane < K ? (order ? order[lane] : lane) : 0;
            const bool u = lane < K && !flag[kk];
            const uint64_t um = __ballot(u);
            if (u) sorder[nf + __builtin_amdgcn_mbcnt_hi((unsigned)(um >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)um, 0u))] = kk;
        }
        __syncthreads();
    }
    const int *wo = prune ? lds_int + 64 : order;
    auto comp = [&](int it) { return wo ? wo[it] : it; };
    {
        const double2 *first2 = stage2 + (size_t)comp(0) * STAGE2;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = min(tid + 256 * q, STAGE2 - 1);
            lds2[e] = first2[e];
        }
    }
    __syncthreads();

    double mx[P];
    unsigned stored[P];  // bit k/4: log-density of component k stored (K <= 64)
    int best[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        mx[p] = -INFINITY;
        stored[p] = 0u;
        best[p] = K;
    }
    unsigned skipped = 0;

    for (int it = 0; it < K; it++) {
        const int k = comp(it);
        const double *lu = lds + (it & 1) * (2 * BUF2);
        const double *lq = lu + NOFF * 64;
        const double *lm = lq + NDIAG * 16;
        if (it + 1 < K) {
            // the next component's block straight into the idle buffer (each
            // lane one 16-byte piece per round: global_load_lds_dwordx4), so
            // no registers hold it and the copy overlaps this component's MFMAs
            const double2 *next2 = stage2 + (size_t)comp(it + 1) * STAGE2;
            double2 *dst = lds2 + ((it + 1) & 1) * BUF2 + 64 * wave;
#pragma unroll
            for (int q = 0; q < PER; q++)
                if (q + 1 < PER || 256 * q + 64 * wave < STAGE2)
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(next2 + min(tid + 256 * q, STAGE2 - 1)),
                        (__attribute__((address_space(3))) void *)(dst + 256 * q), 16, 0, 0);
        }
        const double ck = coef[k];
        v4d acc[P][NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            v4d init;
            init[0] = lm[16 * jb + 0 + g];
            init[1] = lm[16 * jb + 4 + g];
            init[2] = lm[16 * jb + 8 + g];
            init[3] = lm[16 * jb + 12 + g];
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][jb] = init;
        }
        // quad qd and pair pr of strip lb, in the order the stage block holds
        // them (strip by strip: the strip's quads, then its pairs)
        auto quads = [&](int lb, int qd) {
            double b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
            const int jd = lb / 4;
            double2 q2 = {0.0, 0.0};
#pragma unroll
            for (int v = lb % 4; v < 4; v++, qd++) {
                if ((qd & 1) == 0 || v == lb % 4)
                    q2 = *reinterpret_cast<const double2 *>(lq + (qd >> 1) * 32 + slot * 2);
                const double a = (qd & 1) ? q2.y : q2.x;
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jd][v] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b[p], acc[p][jd][v], 0, 0, 0);
            }
        };
        auto pairs = [&](int lb, int pr) {
            double b[P];
#pragma unroll
            for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
            const int jd = lb / 4;
            double2 a2 = {0.0, 0.0};
#pragma unroll
            for (int jb = jd + 1; jb < NB; jb++, pr++) {
                if ((pr & 1) == 0 || jb == jd + 1)
                    a2 = *reinterpret_cast<const double2 *>(lu + (pr >> 1) * 128 + lane * 2);
                const double a = (pr & 1) ? a2.y : a2.x;
#pragma unroll
                for (int p = 0; p < P; p++)
                    acc[p][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[p], acc[p][jb], 0, 0, 0);
            }
        };
        // quads before strip lb: 4 per block for the blocks before lb / 4,
        // then 4 - v0 for each strip v0 < lb % 4 of its block; pairs before
        // strip lb: NB - 1 - lb' / 4 for each strip lb' < lb
        auto quad0 = [](int lb) { return 10 * (lb / 4) + (lb % 4) * 4 - (lb % 4) * (lb % 4 - 1) / 2; };
        auto pair0 = [](int lb) {
            int r = 0;
            for (int l = 0; l < lb; l++) r += NB - 1 - l / 4;
            return r;
        };
        bool skip = false;
        if constexpr (PRUNE) {
          if (prune) {
            // the first block's quads, then the bound
#pragma unroll
            for (int lb = 0; lb < 4; lb++) quads(lb, quad0(lb));
            bool out = true;
#pragma unroll
            for (int p = 0; p < P; p++) {
                double q0 = 0.0;
#pragma unroll
                for (int v = 0; v < 4; v++) q0 += acc[p][0][v] * acc[p][0][v];
                q0 = wave_xor_sum_g(q0);
                const double ub = -0.5 * q0 + ck;
                out = out && (base + 16 * p + c >= n || ub < mx[p] - alive_range<double>());
            }
            skip = __all(out);
          } else {
#pragma unroll
            for (int lb = 0; lb < 4; lb++) quads(lb, quad0(lb));
          }
        }
        double sv[P];
        if (!skip) {
            if constexpr (PRUNE) {
                // the first block's quads are done
#pragma unroll
                for (int lb = 0; lb < 4; lb++) pairs(lb, pair0(lb));
#pragma unroll
                for (int lb = 4; lb < 4 * NB; lb++) {
                    quads(lb, quad0(lb));
                    pairs(lb, pair0(lb));
                }
            } else {
                // strip by strip, two quads or pairs per 16-byte LDS read
                int pr = 0, qd = 0;
                double2 a2 = {0.0, 0.0}, q2 = {0.0, 0.0};
#pragma unroll
                for (int lb = 0; lb < 4 * NB; lb++) {
                    double b[P];
#pragma unroll
                    for (int p = 0; p < P; p++) b[p] = (double)xr[p][lb];
                    const int jd = lb / 4;
#pragma unroll
                    for (int v = lb % 4; v < 4; v++) {
                        if ((qd & 1) == 0) q2 = *reinterpret_cast<const double2 *>(lq + (qd >> 1) * 32 + slot * 2);
                        const double a = (qd & 1) ? q2.y : q2.x;
                        qd++;
#pragma unroll
                        for (int p = 0; p < P; p++)
                            acc[p][jd][v] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b[p], acc[p][jd][v], 0, 0, 0);
                    }
#pragma unroll
                    for (int jb = jd + 1; jb < NB; jb++) {
                        if ((pr & 1) == 0) a2 = *reinterpret_cast<const double2 *>(lu + (pr >> 1) * 128 + lane * 2);
                        const double a = (pr & 1) ? a2.y : a2.x;
                        pr++;
#pragma unroll
                        for (int p = 0; p < P; p++)
                            acc[p][jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[p], acc[p][jb], 0, 0, 0);
                    }
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
                sv[p] = -0.5 * q + ck;
                if constexpr (PRUNE)
                    if (sv[p] > mx[p]) best[p] = k;
                mx[p] = fmax(mx[p], sv[p]);
            }
        } else {
            skipped++;
        }
        // this wave's pieces of the next block have landed and the barrier
        // publishes all of them; the log-densities are stored after it, so
        // that the wait does not include those stores (a lane reads back only
        // the entries it stored itself)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        // only log-densities within alive_range of the running maximum
        // (NaN included) can have a non-zero posterior (lse_posteriors)
        if (!skip && (k & 3) == g) {
#pragma unroll
            for (int p = 0; p < P; p++)
                if (K > 64 || !(sv[p] < mx[p] - alive_range<double>())) {
                    Rt[(size_t)k * npad + base + 16 * p + c] = sv[p];
                    stored[p] |= 1u << ((k >> 2) & 15);
                }
        }
    }
    if (PRUNE && prune) {
        if (g == 0) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int64_t pt = base + 16 * p + c;
                ref[pt] = pt < n ? min(best[p], K - 1) : K;  // valid points sort first
            }
        }
        if (nskip && lane == 0 && skipped) atomicAdd(nskip, (unsigned long long)skipped);
    }

    double llsum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int64_t pt = base + 16 * p + c;
        const bool valid = pt < n;
        const double lse = lse_posteriors(Rt, npad, pt, K, g, mx[p], valid, stored[p]);
        if (g == 0 && valid) llsum += lse;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) llsum += __shfl_xor(llsum, o);
    if (lane == 0) wsum[wave] = llsum;
    __syncthreads();
    if (tid == 0) *ll_out = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    __syncthreads();
}

// NB consecutive floats p[0..NB-1] in loads of W floats, W the largest power
// of two (at most 4) that divides NB; p must be aligned to W floats.
template <int NB>
__device__ inline void load_consecutive(const float *__restrict__ p, float (&v)[NB]) {
    if constexpr (NB % 4 == 0) {
#pragma unroll
        for (int j = 0; j < NB; j += 4) {
            const float4 q = *reinterpret_cast<const float4 *>(p + j);
            v[j] = q.x, v[j + 1] = q.y, v[j + 2] = q.z, v[j + 3] = q.w;
        }
    } else if constexpr (NB % 2 == 0) {
#pragma unroll
        for (int j = 0; j < NB; j += 2) {
            const float2 q = *reinterpret_cast<const float2 *>(p + j);
            v[j] = q.x, v[j + 1] = q.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NB; j++) v[j] = p[j];
    }
}

// ---------------------------------------------------------------------------
// Stage (b), FP64, one chunk of 4096 points (EM2/MatrixEM.cpp:319-344).  Wave
// w owns component blocks kb = w, w+4, ...; per 16 points each lane reads its
// four R values at once (A[i][k] = R[16kb+i][pt] with pt = 4k + m at MFMA step
// m).  Column j of feature block jb is feature NB j + jb, so lane (j, k) reads
// its NB X values of a point in one load (B[k][j] = X[pt][NB j + jb]); 4 NB
// MFMAs per 16 points, and the A values also sum to N_k.  Partials:
// out[KP][DP+1] of this chunk, in feature order.
constexpr int kMeansChunk = 4096;

template <int NB>
__device__ inline void means64_unit(int chunk, const float *__restrict__ X, const double *__restrict__ Rt,
                                    int64_t npad, int KP, double *__restrict__ out) {
    constexpr int DP = 16 * NB;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int64_t i_begin = (int64_t)chunk * kMeansChunk;
    const int KB = KP / 16;
    for (int kb = wave; kb < KB; kb += 4) {
        // per 16 points, lane group g owns points 4g..4g+3: its R values are
        // one 32-byte read, and MFMA step m takes point 4g + m of each group
        const double *rrow = Rt + (size_t)(16 * kb + c) * npad + i_begin + 4 * g;
        const float *xrow = X + (i_begin + 4 * g) * DP + NB * c;
        v4d acc[NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) acc[jb] = (v4d){0.0, 0.0, 0.0, 0.0};
        double ns = 0.0;
#pragma unroll 2
        for (int s = 0; s < kMeansChunk; s += 16) {
            const double4 r4 = *reinterpret_cast<const double4 *>(rrow + s);
            const double rv[4] = {r4.x, r4.y, r4.z, r4.w};
            ns += (rv[0] + rv[1]) + (rv[2] + rv[3]);
#pragma unroll
            for (int m = 0; m < 4; m++) {
                float xv[NB];
                load_consecutive<NB>(xrow + (size_t)(s + m) * DP, xv);
#pragma unroll
                for (int jb = 0; jb < NB; jb++)
                    acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(rv[m], (double)xv[jb], acc[jb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[(size_t)(16 * kb + 4 * v + g) * (DP + 1) + NB * c + jb] = acc[jb][v];
        ns = wave_xor_sum_g(ns);
        if (g == 0) out[(size_t)(16 * kb + c) * (DP + 1) + DP] = ns;
    }
}

// mu_k (or the raw sums M_k when raw != 0) and N_k from the chunk partials
// for the 64 columns of column tile ct (256 threads per unit).
__device__ inline void means_finalize_unit(int k, int ct, int d, int dpad, int KP, int chunks,
                                           const double *__restrict__ mpart, double *__restrict__ mu,
                                           double *__restrict__ nk, int raw) {
    // 256 threads: wave w sums its quarter of the chunks for features
    // ct*64 + lane (lane 0 also the N_k column), then the quarters are added
    // in a fixed order; the caller's whole block calls this together
    __shared__ double quarter[4][65];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const size_t stride = (size_t)KP * (dpad + 1);
    const int c0 = (chunks * w) / 4, c1 = (chunks * (w + 1)) / 4;
    const double *col = mpart + (size_t)k * (dpad + 1) + (size_t)c0 * stride;
    const int j = ct * 64 + lane;
    quarter[w][lane] = j < d ? chunk_sum4(col + j, stride, c1 - c0) : 0.0;
    if (lane == 0) quarter[w][64] = chunk_sum4(col + dpad, stride, c1 - c0);
    __syncthreads();
    if (w == 0) {
        const double sn = (quarter[0][64] + quarter[1][64]) + (quarter[2][64] + quarter[3][64]);
        if (ct == 0 && lane == 0) nk[k] = sn;
        const double s = (quarter[0][lane] + quarter[1][lane]) + (quarter[2][lane] + quarter[3][lane]);
        if (j < d) mu[(size_t)k * d + j] = raw ? s : s / sn;
    }
    __syncthreads();  // quarter is reused by the caller's next call
}

// ---------------------------------------------------------------------------
// Stage (c), FP64: components 4cg..4cg+3 (one per wave) over one chunk of
// points (EM2/MatrixEM.cpp:53-67,90-99,346-375).  Per step of 4 points each
// lane forms t = x - mu_k and w t with w = r / N_k for every feature block fb
// and issues one MFMA per block pair ib <= jb: D += (w t)_ib^T t_jb, the
// same sum as the reference's T^T T with T = sqrt(w) t, without the square
// root.  Feature blocks are strided: column c of block fb is feature
// NB c + fb, so a lane reads its NB features of a point in one load, and the
// pairs ib <= jb of strided blocks still hold every feature pair of the
// symmetric matrix once (cov_reduce_unit maps them back).  For d <= 64 the
// loads run two steps (16 points) ahead in two register buffers, which
// measured 10% faster at d <= 48 and 1% at d = 64.
template <int NB>
__device__ inline void cov64_whole(int cg, int chunk, const float *__restrict__ X, const double *__restrict__ Rt,
                                   int64_t npad, int K, int KP, int d, const double *__restrict__ mu,
                                   const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    const int64_t i0 = (int64_t)chunk * chunk_rows;
    const int64_t i1 = (i0 + chunk_rows < npad) ? i0 + chunk_rows : npad;
    double muv[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) muv[fb] = (NB * c + fb < d) ? mu[(size_t)k * d + NB * c + fb] : 0.0;
    const double invn = 1.0 / nk[k];
    v4d acc[UB];
#pragma unroll
    for (int u = 0; u < UB; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double *rrow = Rt + (size_t)k * npad + g;
    const float *xrow = X + g * DP + NB * c;
    // t = x - mu_k and w t for the two 4-point groups of one step
    auto form = [&](const double (&rv)[2], const float (&xv)[2][NB], double (&t)[2][NB], double (&wt)[2][NB]) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double wq = rv[q] * invn;
#pragma unroll
            for (int fb = 0; fb < NB; fb++) {
                t[q][fb] = (double)xv[q][fb] - muv[fb];
                wt[q][fb] = t[q][fb] * wq;
            }
        }
    };
    auto products = [&](const double (&t)[2][NB], const double (&wt)[2][NB]) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            int u = 0;
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = ib; jb < NB; jb++) {
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(wt[q][ib], t[q][jb], acc[u], 0, 0, 0);
                    u++;
                }
        }
    };
    auto load8 = [&](double (&rv)[2], float (&xv)[2][NB], int64_t i) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            rv[q] = rrow[i + 4 * q];
            load_consecutive<NB>(xrow + (i + 4 * q) * DP, xv[q]);
        }
    };
    if constexpr (NB <= 4) {
        // two register buffers of 8 points: the loads for the step two
        // ahead are issued once the current step's operands are formed
        double rv[2][2];
        float xv[2][2][NB];
        // chunk bounds are multiples of 4096 points, so every step has a
        // partner
        if (i0 < i1) {
            load8(rv[0], xv[0], i0);
            load8(rv[1], xv[1], i0 + 8);
        }
        for (int64_t i = i0; i < i1; i += 16) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                double t[2][NB], wt[2][NB];
                form(rv[h], xv[h], t, wt);
                // unconditional (past the chunk end it re-reads the last 8
                // points, unused), so the buffers need no register copies
                const int64_t ni = i + 8 * (2 + h);
                load8(rv[h], xv[h], ni < i1 ? ni : i1 - 8);
                products(t, wt);
            }
        }
    } else {
        // larger d: the loads of a step are issued just before its
        // arithmetic (the two-buffer form measured slower at d = 128)
        for (int64_t i = i0; i < i1; i += 8) {
            double rv[2];
            float xv[2][NB];
            load8(rv, xv, i);
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const double wq = rv[q] * invn;
                double t[NB], wt[NB];
#pragma unroll
                for (int fb = 0; fb < NB; fb++) {
                    t[fb] = (double)xv[q][fb] - muv[fb];
                    wt[fb] = t[fb] * wq;
                }
                int u = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ib++)
#pragma unroll
                    for (int jb = ib; jb < NB; jb++) {
                        acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(wt[ib], t[jb], acc[u], 0, 0, 0);
                        u++;
                    }
            }
        }
    }
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
#pragma unroll
    for (int u = 0; u < UB; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)u * 256 + (4 * v + g) * 16 + c] = acc[u][v];
}

// d > 96: the 28 or 36 block pairs' accumulators of a component (224 or 288
// registers) would leave one wave per SIMD, so the pairs are split over two
// waves (cov64_parts); part PART forms t only for the feature blocks its
// pairs use and w t only for its row blocks.
__device__ __host__ constexpr int cov64_parts(int NB) { return NB >= 7 ? 2 : 1; }

template <int NB, int PART>
__device__ inline void cov64_part(int cg, int chunk, const float *__restrict__ X, const double *__restrict__ Rt,
                                  int64_t npad, int K, int KP, int d, const double *__restrict__ mu,
                                  const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov64_parts(NB);
    constexpr int U0 = upper_part_begin(NB, NPART, PART);
    constexpr int NU = upper_part_size(NB, NPART, PART);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    const int64_t i0 = (int64_t)chunk * chunk_rows;
    const int64_t i1 = (i0 + chunk_rows < npad) ? i0 + chunk_rows : npad;
    double muv[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) muv[fb] = (NB * c + fb < d) ? mu[(size_t)k * d + NB * c + fb] : 0.0;
    const double invn = 1.0 / nk[k];
    v4d acc[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double *rrow = Rt + (size_t)k * npad + g;
    const float *xrow = X + g * DP + NB * c;
    for (int64_t i = i0; i < i1; i += 8) {
        double rv[2];
        float xv[2][NB];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            rv[q] = rrow[i + 4 * q];
            load_consecutive<NB>(xrow + (i + 4 * q) * DP, xv[q]);
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double wq = rv[q] * invn;
            double t[NB], wt[NB];
#pragma unroll
            for (int fb = 0; fb < NB; fb++) {
                if (!upper_part_uses(NB, NPART, PART, fb, false)) continue;
                t[fb] = (double)xv[q][fb] - muv[fb];
                if (upper_part_uses(NB, NPART, PART, fb, true)) wt[fb] = t[fb] * wq;
            }
            int u = 0;
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = ib; jb < NB; jb++, u++) {
                    if (upper_part_of(u, UB, NPART) != PART) continue;
                    acc[u - U0] = __builtin_amdgcn_mfma_f64_16x16x4f64(wt[ib], t[jb], acc[u - U0], 0, 0, 0);
                }
        }
    }
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
#pragma unroll
    for (int u = 0; u < NU; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)(U0 + u) * 256 + (4 * v + g) * 16 + c] = acc[u][v];
}

// One workgroup's share of stage (c), FP64: components 4cg..4cg+3 over one
// chunk of points, and for d > 96 one of the two parts of their block pairs.
template <int NB>
__device__ inline void cov64_unit(int cg, int chunk, const float *__restrict__ X, const double *__restrict__ Rt,
                                  int64_t npad, int K, int KP, int d, const double *__restrict__ mu,
                                  const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart,
                                  int part = 0) {
    if constexpr (cov64_parts(NB) > 1) {
        if (part == 0) cov64_part<NB, 0>(cg, chunk, X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart);
        else cov64_part<NB, 1>(cg, chunk, X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart);
    } else {
        cov64_whole<NB>(cg, chunk, X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart);
    }
}

// Sum the chunk partials of block pair u = (ib <= jb) of component k (fixed
// order) and write them to both triangles of cov.  The blocks are strided
// (row r of block ib is feature NB r + ib): an off-diagonal pair holds each of
// its feature pairs once, a diagonal pair holds (a, b) and (b, a), of which
// the entry with row <= column is taken.
__device__ inline void cov_reduce_unit(int k, int u, int d, int NB, int KP, int chunks,
                                       const double *__restrict__ cpart, double *__restrict__ cov) {
    const int tid = threadIdx.x;
    const int UB = NB * (NB + 1) / 2;
    int ib = 0, rem = u;
    while (rem >= NB - ib) {
        rem -= NB - ib;
        ib++;
    }
    const int jb = ib + rem;
    const int r = tid >> 4, col = tid & 15;
    const int a = NB * r + ib, bb = NB * col + jb;
    if (a >= d || bb >= d || (ib == jb && r > col)) return;
    const double s = chunk_sum4(cpart + ((size_t)k * UB + u) * 256 + tid, (size_t)KP * UB * 256, chunks);
    double *C = cov + (size_t)k * d * d;
    C[(size_t)a * d + bb] = s;
    C[(size_t)bb * d + a] = s;
}

// Regularise component k (EM2/EMAlgorithm.cpp:237-244) and set
// w_k = N_k / ntotal.  str: one LDS double.
__device__ inline void cov_finish_unit(int k, int d, const double *__restrict__ nk, double g, double ntotal,
                                       double *__restrict__ cov, double *__restrict__ w, double *str) {
    const int tid = threadIdx.x;
    double *C = cov + (size_t)k * d * d;
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < d; i++) tr += C[(size_t)i * d + i];
        *str = tr;
        w[k] = nk[k] / ntotal;
    }
    __syncthreads();
    const double diag = g * *str / (double)d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        const int a = e / d, bb = e % d;
        C[e] = C[e] * (1.0 - g) + (a == bb ? diag : 0.0);
    }
    __syncthreads();
}

// Deterministic sum of n doubles: fixed strided partials, then a fixed tree
// over 256 threads.  s: 256 LDS doubles.  Result valid in thread 0.
__device__ inline double block_sum(const double *__restrict__ parts, int64_t n, double *s) {
    double a = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) a += parts[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

}  // namespace fem
