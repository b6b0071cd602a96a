// fem_kernels_bf16.hip -- BF16x3 precision form of the E-step and M-step
// contractions on v_mfma_f32_16x16x32_bf16 (gfx950).
//
// Every FP32 operand value v is split as v = hi + lo with hi = bf16(v) and
// lo = bf16(v - hi) (both round-to-nearest-even), so |v - hi - lo| <= 2^-17 |v|.
// A product a*b is formed as hi_a*hi_b + hi_a*lo_b + lo_a*hi_b (three MFMAs
// accumulating in FP32); the dropped lo_a*lo_b term is below 2^-16 |ab|.
// Posteriors are FP32 (shared with the FP32 form), the per-chunk partial sums
// are flushed into FP64 at fixed intervals, and every cross-chunk reduction,
// the Cholesky and the log-determinant stay FP64.
//
// MFMA 16x16x32 operand layout (wave64, lane l, g = l/16, c = l%16):
// A[i][k] for i = c, k = 8g..8g+7 (eight bf16 in four VGPRs, element e in
// the low half of word e/2 when e is even); B[k][j] for j = c, the same k;
// accumulator element v of lane l holds D[4g + v][c].
#include "fem_internal.h"
#include "fem_tiles.h"

#include <cmath>

namespace fem {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

// E-step pairs (ls, jb): 32-feature input strip ls, 16-feature output block jb,
// with jb >= 2 ls (U_k^T is lower triangular).
__device__ __host__ constexpr int pairs_bf16(int NB) {
    int p = 0;
    for (int ls = 0; 2 * ls < NB; ls++) p += NB - 2 * ls;
    return p;
}

// bf16 bits of x, round to nearest even (operands here are finite).
__device__ __host__ inline unsigned int bf16_bits(float x) {
    const unsigned int u = __builtin_bit_cast(unsigned int, x);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

__device__ inline float bf16_value(unsigned int b) { return __builtin_bit_cast(float, b << 16); }

// Split eight floats into packed hi and lo bf16 operands (v_cvt_pk_bf16_f32,
// round to nearest even; the hi halves are widened back by a shift or mask).
__device__ inline void split8(const float (&t)[8], v8bf &hi, v8bf &lo) {
    v4u h, l;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const v2bf hb = __builtin_convertvector((v2f){t[2 * w], t[2 * w + 1]}, v2bf);
        const unsigned int hw = __builtin_bit_cast(unsigned int, hb);
        const float h0 = __builtin_bit_cast(float, hw << 16);
        const float h1 = __builtin_bit_cast(float, hw & 0xFFFF0000u);
        const v2bf lb = __builtin_convertvector((v2f){t[2 * w] - h0, t[2 * w + 1] - h1}, v2bf);
        h[w] = hw;
        l[w] = __builtin_bit_cast(unsigned int, lb);
    }
    hi = __builtin_bit_cast(v8bf, h);
    lo = __builtin_bit_cast(v8bf, l);
}

__device__ inline v4f mfma3(v8bf ah, v8bf al, v8bf bh, v8bf bl, v4f acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}




// ---------------------------------------------------------------------------
// E-step operands of component k, packed as one contiguous stage block of
// estepb_stage_words(NB) words so the E-step copies it with plain strided
// loads:
//   [0, FW)            fragments A[i][kk] = U_k^T[16 jb + i][32 ls + kk]
//                      = U_k[32 ls + kk][16 jb + i] for every pair (ls, jb),
//                      split hi/lo, stored [pair][hi|lo][lane][4 words];
//   [FW, FW + 32 NS)   mu_hi = float(mu_k), zero beyond d;
//   [.., + 16 NB)      -U_k^T mu_lo with mu_lo = mu_k - mu_hi, the seed of
//                      the accumulators: D = U_k^T (x - mu_hi) - U_k^T mu_lo.
template <int NB>
__device__ __host__ constexpr int estepb_stage_words() {
    return pairs_bf16(NB) * 512 + 32 * ((NB + 1) / 2) + 16 * NB;
}

__device__ __host__ inline int stage_words_of(int NB) {
    return pairs_bf16(NB) * 512 + 32 * ((NB + 1) / 2) + 16 * NB;
}

// The hi/lo fragments of U_k^T for every (ls, jb) pair into st[0, 512 pairs),
// by the threads of one workgroup.
__device__ void pack_fragments_bf16(const double *__restrict__ Uk, int dpad, int NB, unsigned int *__restrict__ st) {
    const int np = pairs_bf16(NB);
    for (int e = threadIdx.x; e < np * 64 * 4; e += blockDim.x) {
        const int p = e / 256, lane = (e / 4) % 64, w = e % 4;
        int ls = 0, rem = p;
        while (rem >= NB - 2 * ls) {
            rem -= NB - 2 * ls;
            ls++;
        }
        const int jb = 2 * ls + rem;
        const int g = lane >> 4, c = lane & 15;
        unsigned int hw = 0, lw = 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int f = 32 * ls + 8 * g + 2 * w + h;
            const double v = (f < dpad) ? Uk[(size_t)f * dpad + 16 * jb + c] : 0.0;
            const unsigned int hb = bf16_bits((float)v);
            const unsigned int lb = bf16_bits((float)(v - (double)bf16_value(hb)));
            hw |= hb << (16 * h);
            lw |= lb << (16 * h);
        }
        unsigned int *dst = st + p * 512;
        dst[lane * 4 + w] = hw;
        dst[256 + lane * 4 + w] = lw;
    }
}

__global__ __launch_bounds__(256) void k_prep_bf16(int d, int dpad, int NB, const double *__restrict__ U,
                                                   const double *__restrict__ mu, unsigned int *__restrict__ stage) {
    const int k = blockIdx.x;
    const int np = pairs_bf16(NB), NS = (NB + 1) / 2;
    const int FW = np * 512;
    const double *Uk = U + (size_t)k * dpad * dpad;
    unsigned int *st = stage + (size_t)k * stage_words_of(NB);
    pack_fragments_bf16(Uk, dpad, NB, st);
    for (int j = threadIdx.x; j < 32 * NS; j += blockDim.x)
        st[FW + j] = __builtin_bit_cast(unsigned int, (j < d) ? (float)mu[(size_t)k * d + j] : 0.0f);
    for (int j = threadIdx.x; j < dpad; j += blockDim.x) {
        double s = 0.0;
        if (j < d)
            for (int f = 0; f <= j; f++) {
                const double m = mu[(size_t)k * d + f];
                s += Uk[(size_t)f * dpad + j] * (m - (double)(float)m);
            }
        st[FW + 32 * NS + j] = __builtin_bit_cast(unsigned int, (float)(-s));
    }
}

// ---------------------------------------------------------------------------
// Stage (a), BF16x3.  Workgroup = 4 waves x P groups of 16 points.  Lane
// (g, c) holds, for point c of each group, features 32 ls + 8g .. +7 of every
// strip ls as FP32 (the B operand layout).  Per component: accumulators
// seeded with -U_k^T mu_lo, centre with mu_hi, split, 3 MFMAs per pair; the
// component's stage block is double-buffered in LDS as in the FP32 form, one
// barrier per component.
template <int NB, int P>
__global__ __launch_bounds__(256) void k_estepb(const float *__restrict__ X, int64_t n, int64_t npad, int K,
                                                const unsigned int *__restrict__ stage,
                                                const double *__restrict__ coef, float *__restrict__ Rt,
                                                double *__restrict__ ll_part) {
    constexpr int DP = 16 * NB;
    constexpr int NS = (NB + 1) / 2;
    constexpr int NPAIR = pairs_bf16(NB);
    constexpr int FW = NPAIR * 512;  // fragment words per component
    constexpr int STAGE = estepb_stage_words<NB>();  // a multiple of 4
    constexpr int STAGE4 = STAGE / 4;
    constexpr int PER = (STAGE4 + 255) / 256;
    extern __shared__ __align__(16) unsigned int ldsb[];  // 2 * STAGE words
    __shared__ double wsum[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int64_t base = (int64_t)blockIdx.x * (64 * P) + wave * (16 * P);

    float xr[P][NS][8];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int ls = 0; ls < NS; ls++) {
            const int f0 = 32 * ls + 8 * g;
            if (f0 < DP) {
                const float4 *src = reinterpret_cast<const float4 *>(X + (base + 16 * p + c) * DP + f0);
                const float4 u0 = src[0], u1 = src[1];
                xr[p][ls][0] = u0.x; xr[p][ls][1] = u0.y; xr[p][ls][2] = u0.z; xr[p][ls][3] = u0.w;
                xr[p][ls][4] = u1.x; xr[p][ls][5] = u1.y; xr[p][ls][6] = u1.z; xr[p][ls][7] = u1.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) xr[p][ls][e] = 0.0f;
            }
        }

    // stage blocks move in 16-byte pieces; piece indices past the block are
    // clamped to its last piece (same value to the same address)
    const uint4 *stage4 = reinterpret_cast<const uint4 *>(stage);
    uint4 *lds4 = reinterpret_cast<uint4 *>(ldsb);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int e = min(tid + 256 * q, STAGE4 - 1);
        lds4[e] = stage4[e];
    }
    __syncthreads();

    float mx = -INFINITY;  // of point (g mod P, c)
    uint64_t stored = 0;   // bit k: its log-density stored (K <= 64)

    for (int k = 0; k < K; k++) {
        const unsigned int *lf = ldsb + (k & 1) * STAGE;
        const float *lmh = reinterpret_cast<const float *>(lf + FW);
        const float *lml = lmh + 32 * NS;
        const uint4 *next4 = stage4 + (size_t)(k + 1 < K ? k + 1 : k) * STAGE4;
        uint4 pre[PER];
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
        int pr = 0;
#pragma unroll
        for (int ls = 0; ls < NS; ls++) {
            const float4 m0 = *reinterpret_cast<const float4 *>(lmh + 32 * ls + 8 * g);
            const float4 m1 = *reinterpret_cast<const float4 *>(lmh + 32 * ls + 8 * g + 4);
            const float mh[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            v8bf bh[P], bl[P];
#pragma unroll
            for (int p = 0; p < P; p++) {
                float t[8];
#pragma unroll
                for (int e = 0; e < 8; e++) t[e] = xr[p][ls][e] - mh[e];
                split8(t, bh[p], bl[p]);
            }
#pragma unroll
            for (int jb = 2 * ls; jb < NB; jb++) {
                const v4u *fp = reinterpret_cast<const v4u *>(lf + pr * 512);
                const v8bf ah = __builtin_bit_cast(v8bf, fp[lane]);
                const v8bf al = __builtin_bit_cast(v8bf, fp[64 + lane]);
                pr++;
#pragma unroll
                for (int p = 0; p < P; p++) acc[p][jb] = mfma3(ah, al, bh[p], bl[p], acc[p][jb]);
            }
        }
        // squared norms summed over the lane groups so that lane (g, c)
        // holds the log-density of point (g mod P, c) of its wave
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
            // dead) and ahead of the branch below, so that the loads of the
            // next block stay in this basic block, ahead of the MFMAs
            uint4 *dst = lds4 + ((k + 1) & 1) * STAGE4;
#pragma unroll
            for (int q = 0; q < PER; q++) dst[min(tid + 256 * q, STAGE4 - 1)] = pre[q];
        }
        // stored only while within alive_range of the point's running maximum
        // (fem_tiles.h lse_posteriors)
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

// ---------------------------------------------------------------------------
// Reference-centred form of stage (a), BF16x3.  The form above centres and
// splits every (point, component, feature): three VALU operations per value
// against 3/16 of an MFMA, which leaves it VALU-bound.  Here each point is
// centred once, on a reference component a(i) (the component of its largest
// log-density in the previous E-step of this form), and split once:
//   U_k^T (x - mu_k) = U_k^T (x - mu_a,hi) + s_k[a],
//   s_k[a] = U_k^T (mu_a,hi - mu_k)   (FP64, rounded once to FP32),
// with mu_a,hi = float(mu_a).  The accumulators start at s_k[a(i)], read
// from a per-component table in LDS, and the MFMAs take the point's split
// operands held in registers for all K components.  For k = a this is the
// other form's arithmetic; for k != a the split error is relative to
// |x - mu_a| instead of |x - mu_k|, no larger for a point near its
// reference.  Any reference gives the same densities up to rounding; a
// distant one only makes them less precise.
//
// Stage block of component k (estepr_buffer_pieces(NB, K) 16-byte pieces,
// padded with zeros to whole copy rounds of 256 pieces):
//   [0, FW)                 fragments as in the form above;
//   [FW, FW + K SR)         s_k[a][j] as FP32, row a at FW + a SR, SR = 16 NB + 4
//                           (the 4-word pad spreads the rows' banks), zero for
//                           j >= d.
// After the K blocks: the K x 16 NB table of mu_a,hi (zero beyond d).
__device__ __host__ inline int estepr_row(int NB) { return 16 * NB + 4; }
__device__ __host__ inline int estepr_stage_words(int NB, int K) {
    return pairs_bf16(NB) * 512 + K * estepr_row(NB);
}
// Copy rounds of 256 16-byte pieces (one per thread) per block, rounded up
// to an even count (the kernel is instantiated per count); at most 8: two
// 32 KB LDS buffers, two workgroups per CU.
__device__ __host__ inline int estepr_rounds(int NB, int K) {
    const int r = (estepr_stage_words(NB, K) / 4 + 255) / 256;
    return r + (r & 1);
}
__device__ __host__ inline int estepr_buffer_pieces(int NB, int K) { return 256 * estepr_rounds(NB, K); }

// Grid (K, kPrepParts): workgroup (k, part) writes every kPrepParts-th group
// of 256 seed words of component k's block; part 0 also packs the fragments
// and mu_k,hi.
constexpr int kPrepParts = 8;

__global__ __launch_bounds__(256) void k_prep_ref(int d, int dpad, int NB, int K, const double *__restrict__ U,
                                                  const double *__restrict__ mu, unsigned int *__restrict__ stage) {
    const int k = blockIdx.x, part = blockIdx.y;
    const int FW = pairs_bf16(NB) * 512, SR = estepr_row(NB);
    const int BW = 4 * estepr_buffer_pieces(NB, K);  // words per padded block
    const double *Uk = U + (size_t)k * dpad * dpad;
    unsigned int *st = stage + (size_t)k * BW;
    if (part == 0) pack_fragments_bf16(Uk, dpad, NB, st);
    for (int e = part * 256 + threadIdx.x; e < BW - FW; e += kPrepParts * 256) {
        const int a = e / SR, j = e % SR;
        double s = 0.0;
        if (a < K && j < d)
            for (int f = 0; f <= j; f++)
                s += Uk[(size_t)f * dpad + j] * ((double)(float)mu[(size_t)a * d + f] - mu[(size_t)k * d + f]);
        st[FW + e] = __builtin_bit_cast(unsigned int, (float)s);
    }
    float *muh = reinterpret_cast<float *>(stage + (size_t)K * BW) + (size_t)k * dpad;
    if (part == 0)
        for (int j = threadIdx.x; j < dpad; j += blockDim.x) muh[j] = (j < d) ? (float)mu[(size_t)k * d + j] : 0.0f;
}

// One component of the reference-centred E-step: accumulators from the seed
// rows, 3 MFMAs per (strip, block) pair and point group, squared norms summed
// over the lane groups so that lane (g, c) holds the log-density of point
// (g, c) of its wave, and that point's running max / argmax; component k+1's stage block is copied straight
// from global memory into the other LDS buffer meanwhile (one 16-byte piece
// per thread and round: global_load_lds_dwordx4, no registers held).  The two
// buffers are distinct LDS arrays, so the compiler knows the copy does not
// write what this component reads and waits for it only where the caller
// asks (a single shared array made it wait before the first LDS read).
template <int NB, int R>
__device__ __attribute__((always_inline)) inline void estepr_component(
    int k, int K, const uint4 *__restrict__ cur4, uint4 *__restrict__ nxt4, const uint4 *__restrict__ stage4,
    const double *__restrict__ coef, int tid, int wave, int lane, int g, const int (&ar)[4],
    const v8bf (&bh)[4][(NB + 1) / 2], const v8bf (&bl)[4][(NB + 1) / 2], float &mx, int &best, float &sv) {
    constexpr int NS = (NB + 1) / 2;
    constexpr int FW = pairs_bf16(NB) * 512;
    constexpr int SR = 16 * NB + 4;
    constexpr int P = 4;
    constexpr int BUF4 = 256 * R;
    const float ck = (float)coef[k];
    constexpr int NPAIR = pairs_bf16(NB);
    const float *lsd = reinterpret_cast<const float *>(cur4) + FW;
    // every LDS read of the component first (fragments in pair order, then
    // the seed rows), so that one LDS latency precedes the MFMAs instead of
    // one per pair: 16-cycle MFMAs do not hide a read issued just ahead
    v8bf fh[NPAIR], fl[NPAIR];
#pragma unroll
    for (int pr = 0; pr < NPAIR; pr++) {
        const v4u *fp = reinterpret_cast<const v4u *>(cur4 + pr * 128);
        fh[pr] = __builtin_bit_cast(v8bf, fp[lane]);
        fl[pr] = __builtin_bit_cast(v8bf, fp[64 + lane]);
    }
    v4f acc[P][NB];
#pragma unroll
    for (int jb = 0; jb < NB; jb++)
#pragma unroll
        for (int p = 0; p < P; p++) {
            const float4 s = *reinterpret_cast<const float4 *>(lsd + ar[p] * SR + 16 * jb + 4 * g);
            acc[p][jb] = (v4f){s.x, s.y, s.z, s.w};
        }
    // the copy is issued after the reads: a wait the compiler places before
    // an LDS read then finds no copy in flight
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 < K) {
        const uint4 *next4 = stage4 + (size_t)(k + 1) * BUF4;
#pragma unroll
        for (int q = 0; q < R; q++)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(next4 + 256 * q + tid),
                (__attribute__((address_space(3))) void *)(nxt4 + 256 * q + 64 * wave), 16, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    int pr = 0;
#pragma unroll
    for (int ls = 0; ls < NS; ls++)
#pragma unroll
        for (int jb = 2 * ls; jb < NB; jb++) {
#pragma unroll
            for (int p = 0; p < P; p++) acc[p][jb] = mfma3(fh[pr], fl[pr], bh[p][ls], bl[p][ls], acc[p][jb]);
            pr++;
        }
    float q[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        q[p] = 0.f;
#pragma unroll
        for (int jb = 0; jb < NB; jb++)
#pragma unroll
            for (int v = 0; v < 4; v++) q[p] += acc[p][jb][v] * acc[p][jb][v];
    }
    sv = -0.5f * group_sums4_spread(q) + ck;
    if (sv > mx) best = k;
    mx = fmaxf(mx, sv);
}

// Workgroup = 4 waves x 4 groups of 16 points, lanes as in the form above; R
// copy rounds per stage block (estepr_rounds), two LDS buffers of 4R KB.
// ref[i] is read at the start (clamped into [0, K)) and replaced at the end by
// the first component of largest log-density.  The workgroup's 256 points are
// ranked by (reference, row) and dealt to the lanes in that order, so a
// 16-point group holds few distinct references, consecutive ones: its reads
// of the seed rows are mostly broadcasts over distinct LDS banks instead of
// 16 rows that collide in the banks.  Log-densities, posteriors and
// references stay in point order.  One barrier per component.  Lane (g, c)
// stores the log-densities of point (g, c) of its wave (those within
// alive_range of its running maximum); the log-sum-exp sweep (4 lanes per
// point, components g, g + 4, ...) runs over the points in their own order,
// so 16 lanes cover a whole row segment of Rt, taking each point's maximum
// and stored bits from LDS (written by location before its barrier).
template <int NB, int R>
__global__ __launch_bounds__(256) void k_estepr(const float *__restrict__ X, int64_t n, int64_t npad, int K,
                                                const unsigned int *__restrict__ stage,
                                                const double *__restrict__ coef, int *__restrict__ ref,
                                                float *__restrict__ Rt, double *__restrict__ ll_part) {
    constexpr int DP = 16 * NB;
    constexpr int NS = (NB + 1) / 2;
    constexpr int P = 4;  // one point per thread of the 256-thread workgroup
    constexpr int BUF4 = 256 * R;  // 16-byte pieces per (padded) block
    __shared__ __align__(16) uint4 bufA[BUF4];
    __shared__ __align__(16) uint4 bufB[BUF4];
    __shared__ double wsum[4];
    __shared__ __align__(16) int skey[256];  // ranking keys; then each point's max (as float)
    __shared__ uint64_t sbits[256];          // each point's stored bits
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, c = lane & 15;
    const int64_t blk0 = (int64_t)blockIdx.x * 256;
    const uint4 *stage4 = reinterpret_cast<const uint4 *>(stage);
    const float *muh = reinterpret_cast<const float *>(stage4 + (size_t)K * BUF4);

    // block 0 into buffer A
#pragma unroll
    for (int q = 0; q < R; q++) bufA[256 * q + tid] = stage4[256 * q + tid];

    // rank of each point's key (reference, row) among the workgroup's: a
    // permutation, the same on every run
    {
        const int key = min(max(ref[blk0 + tid], 0), K - 1) * 256 + tid;
        skey[tid] = key;
        __syncthreads();
        int rank = 0;
        for (int i = 0; i < 256; i += 4) {
            const int4 q = *reinterpret_cast<const int4 *>(skey + i);
            rank += (q.x < key) + (q.y < key) + (q.z < key) + (q.w < key);
        }
        __syncthreads();
        skey[rank] = key;
        __syncthreads();
    }

    int ar[P], loc[P];
    v8bf bh[P][NS], bl[P][NS];
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int key = skey[wave * 64 + 16 * p + c];
        loc[p] = key & 255;
        const int64_t pt = blk0 + loc[p];
        const int a = key >> 8;
        ar[p] = a;
#pragma unroll
        for (int ls = 0; ls < NS; ls++) {
            const int f0 = 32 * ls + 8 * g;
            float t[8];
            if (f0 < DP) {
                const float4 *src = reinterpret_cast<const float4 *>(X + pt * DP + f0);
                const float4 *mh = reinterpret_cast<const float4 *>(muh + (size_t)a * DP + f0);
                // the points are read once: non-temporal (L2 keeps the stage blocks)
                const v4f w0 = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(src));
                const v4f w1 = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(src) + 1);
                const float4 u0 = {w0[0], w0[1], w0[2], w0[3]}, u1 = {w1[0], w1[1], w1[2], w1[3]};
                const float4 m0 = mh[0], m1 = mh[1];
                t[0] = u0.x - m0.x; t[1] = u0.y - m0.y; t[2] = u0.z - m0.z; t[3] = u0.w - m0.w;
                t[4] = u1.x - m1.x; t[5] = u1.y - m1.y; t[6] = u1.z - m1.z; t[7] = u1.w - m1.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) t[e] = 0.0f;
            }
            split8(t, bh[p][ls], bl[p][ls]);
        }
    }
    __syncthreads();

    // lane (g, c) follows point (g, c) of its wave: its log-densities, max
    // and argmax
    int myloc = loc[0], best = ar[0];
#pragma unroll
    for (int p = 1; p < P; p++)
        if (p == g) {
            myloc = loc[p];
            best = ar[p];
        }
    float mx = -INFINITY, sv = 0.f;
    float *rt = Rt + blk0 + myloc;
    // stored only while within alive_range of the point's running maximum
    // (fem_tiles.h lse_posteriors); bit k of `stored`: component k stored
    uint64_t stored = 0;
    auto store = [&](int k) {
        if (K > 64 || !(sv < mx - alive_range<float>())) {
            rt[(size_t)k * npad] = sv;
            stored |= 1ull << (k & 63);
        }
    };
    // components in pairs: even ones read A and fill B, odd ones the reverse;
    // at each barrier this wave's copy has landed and the barrier publishes
    // every wave's; the log-densities are stored after the wait
    // (the scheduling barriers keep the component's MFMAs ahead of the wait)
    for (int k = 0; k < K; k += 2) {
        estepr_component<NB, R>(k, K, bufA, bufB, stage4, coef, tid, wave, lane, g, ar, bh, bl, mx, best, sv);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        store(k);
        if (k + 1 < K) {
            estepr_component<NB, R>(k + 1, K, bufB, bufA, stage4, coef, tid, wave, lane, g, ar, bh, bl, mx, best,
                                    sv);
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            store(k + 1);
        }
    }
    // the sweep below runs over the points in their own order (whole rows
    // of Rt per 16 lanes) and reads entries other lanes stored: each point's
    // max and stored bits go to LDS by its location
    reinterpret_cast<float *>(skey)[myloc] = mx;
    sbits[myloc] = stored;
    __syncthreads();

    ref[blk0 + myloc] = best;
    double llsum = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int l = wave * 64 + 16 * p + c;
        const int64_t pt = blk0 + l;
        const bool valid = pt < n;
        const float mxp = reinterpret_cast<const float *>(skey)[l];
        const float lse = lse_posteriors(Rt, npad, pt, K, g, mxp, valid, group_bits(sbits[l], g));
        if (g == 0 && valid) llsum += (double)lse;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) llsum += __shfl_xor(llsum, o);
    if (lane == 0) wsum[wave] = llsum;
    __syncthreads();
    if (tid == 0) ll_part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ __launch_bounds__(256) void k_reduce_sumb(const double *__restrict__ parts, int64_t n,
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

bool estepr_fits(int NB, int K) {
    // two LDS stage buffers of at most 64 KB together: two workgroups per CU
    return NB <= 4 && K >= 1 && estepr_rounds(NB, K) <= 8;
}

size_t estepr_words(int NB, int K) { return (size_t)K * 4 * estepr_buffer_pieces(NB, K) + (size_t)K * 16 * NB; }

// Packs the reference-centred stage when the context has it and its
// references come from an E-step since the data or means were last set;
// otherwise the per-component form's stage.
hipError_t launch_prep_bf16(fem_ctx *c) {
    c->stage_ref = c->stager && c->refc && c->ref_ok && estepr_fits(c->NB, c->K);
    if (c->stage_ref)
        hipLaunchKernelGGL(k_prep_ref, dim3(c->K, kPrepParts), dim3(256), 0, c->stream, c->d, c->dpad, c->NB, c->K, c->U,
                           c->mu, c->stager);
    else
        hipLaunchKernelGGL(k_prep_bf16, dim3(c->K), dim3(256), 0, c->stream, c->d, c->dpad, c->NB, c->U, c->mu,
                           c->stageb);
    return hipGetLastError();
}

template <int NB, int P>
static hipError_t estepb_t(fem_ctx *c) {
    const size_t lds = (size_t)2 * estepb_stage_words<NB>() * sizeof(unsigned int);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_estepb<NB, P>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = (int)(c->npad / (64 * P));
    hipLaunchKernelGGL((k_estepb<NB, P>), dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->n, c->npad, c->K,
                       c->stageb, c->coef, c->Rt32, c->ll_part);
    hipLaunchKernelGGL(k_reduce_sumb, dim3(1), dim3(256), 0, c->stream, c->ll_part, (int64_t)blocks, c->ll_dev);
    return hipGetLastError();
}

// Point groups per wave: 4 for d <= 64 unless estep_groupsb == 2, else 2.
template <int NB, int R>
static hipError_t estepr_t(fem_ctx *c) {
    const int blocks = (int)(c->npad / 256);
    hipLaunchKernelGGL((k_estepr<NB, R>), dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->n, c->npad, c->K,
                       c->stager, c->coef, c->refc, c->Rt32, c->ll_part);
    hipLaunchKernelGGL(k_reduce_sumb, dim3(1), dim3(256), 0, c->stream, c->ll_part, (int64_t)blocks, c->ll_dev);
    return hipGetLastError();
}

template <int NB>
static hipError_t estepr_rounds_t(fem_ctx *c) {
    switch (estepr_rounds(NB, c->K)) {
        case 2: return estepr_t<NB, 2>(c);
        case 4: return estepr_t<NB, 4>(c);
        case 6: return estepr_t<NB, 6>(c);
        case 8: return estepr_t<NB, 8>(c);
    }
    return hipErrorInvalidValue;
}

static hipError_t estepb_ref(fem_ctx *c) {
    switch (c->NB) {
        case 1: return estepr_rounds_t<1>(c);
        case 2: return estepr_rounds_t<2>(c);
        case 3: return estepr_rounds_t<3>(c);
        case 4: return estepr_rounds_t<4>(c);
    }
    return hipErrorInvalidValue;
}

static hipError_t estepb_plain(fem_ctx *c);

// The reference-centred kernel when the last prep packed its stage; else the
// per-component form, after which (where the reference form fits) the
// posterior argmax becomes the references of the next E-step.
hipError_t launch_estepb(fem_ctx *c) {
    if (c->stage_ref) return estepb_ref(c);
    hipError_t e = estepb_plain(c);
    if (e == hipSuccess && c->stager && c->refc && estepr_fits(c->NB, c->K)) {
        e = launch_classes(c, c->refc);
        if (e == hipSuccess) c->ref_ok = true;
    }
    return e;
}

static hipError_t estepb_plain(fem_ctx *c) {
    const bool p4 = c->estep_groupsb == 4;
    switch (c->NB) {
        case 1: return p4 ? estepb_t<1, 4>(c) : estepb_t<1, 2>(c);
        case 2: return p4 ? estepb_t<2, 4>(c) : estepb_t<2, 2>(c);
        case 3: return p4 ? estepb_t<3, 4>(c) : estepb_t<3, 2>(c);
        case 4: return p4 ? estepb_t<4, 4>(c) : estepb_t<4, 2>(c);
        case 5: return estepb_t<5, 2>(c);
        case 6: return estepb_t<6, 2>(c);
        case 7: return estepb_t<7, 2>(c);
        case 8: return estepb_t<8, 2>(c);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage (b), BF16x3: M[k][j] = sum_i R[i][k] X[i][j] per 4096-point chunk.
// A = R^T fragment (components 16 kb + c, points 8g..8g+7 of the step), B = X
// fragment (points 8g.., feature NB c + jb in column c of block jb, so a lane
// reads its NB values of a point in wide loads); FP32 accumulators flushed
// into FP64 every 1024 points.  Output layout as the FP32 form.
template <int NB>
__global__ __launch_bounds__(256) void k_meansb(const float *__restrict__ X, const float *__restrict__ Rt,
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
        const float *rrow = Rt + (size_t)(16 * kb + c) * npad + i_begin + 8 * g;
        const float *xrow = X + (i_begin + 8 * g) * DP + NB * c;
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
            for (int s = s0; s < s0 + FLUSH; s += 32) {
                const float4 r0 = *reinterpret_cast<const float4 *>(rrow + s);
                const float4 r1 = *reinterpret_cast<const float4 *>(rrow + s + 4);
                const float rv[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
                ns += ((rv[0] + rv[1]) + (rv[2] + rv[3])) + ((rv[4] + rv[5]) + (rv[6] + rv[7]));
                v8bf rh, rl;
                split8(rv, rh, rl);
                float xp[8][NB];
#pragma unroll
                for (int e = 0; e < 8; e++) load_consecutive<NB>(xrow + (size_t)(s + e) * DP, xp[e]);
#pragma unroll
                for (int jb = 0; jb < NB; jb++) {
                    float xv[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) xv[e] = xp[e][jb];
                    v8bf xh, xl;
                    split8(xv, xh, xl);
                    acc[jb] = mfma3(rh, rl, xh, xl, acc[jb]);
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
static hipError_t meansb_t(fem_ctx *c) {
    hipLaunchKernelGGL(k_meansb<NB>, dim3(c->m_chunks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad, c->KP,
                       c->mpart);
    return hipGetLastError();
}

hipError_t launch_meansb_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return meansb_t<1>(c);
        case 2: return meansb_t<2>(c);
        case 3: return meansb_t<3>(c);
        case 4: return meansb_t<4>(c);
        case 5: return meansb_t<5>(c);
        case 6: return meansb_t<6>(c);
        case 7: return meansb_t<7>(c);
        case 8: return meansb_t<8>(c);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage (c), BF16x3: T = (x - mu_hi) sqrt(r / N_k) in FP32, split, and
// D += T_ib^T T_jb (upper blocks) with three MFMAs per block per 32 points.
// The same register fragment is the A operand (T_ib^T) and the B operand
// (T_jb).  Centring with mu_hi alone changes D by N_k mu_lo mu_lo^T plus a
// term proportional to sum_i r_ik (x_i - mu_k), both below 1e-9 of D here.
//
// For d > 64 the upper blocks of a component are split into NPART = 2 parts
// (contiguous in row-major block order, equal counts), each its own wave, so
// a wave's accumulators stay in registers; a part forms T only for the
// feature blocks its upper blocks touch.  FP32 accumulators are flushed into
// FP64 every 2048 points: into registers for d <= 48, else into the wave's
// own slice of LDS (written to the partial slice once at the end).  Grid and
// XCD remapping as the FP32 form, with the part as the fastest-varying index;
// strided feature blocks as the FP32 form (a lane's NB features of a point in
// wide loads).  For d <= 64 the 32-point steps are software-pipelined (each
// step's MFMAs beside the next step's split); for d > 64 the loads of the
// next 32 points are issued once the current 32 are split, before their
// MFMAs.
__device__ __host__ constexpr int cov_parts(int NB) { return NB > 4 ? 2 : 1; }

// Part of upper block u (row-major over ib <= jb) among NPART parts, the
// first block of part PART, its block count, and whether it uses feature
// block fb (fem_tiles.h's upper_part_* with NPART = cov_parts(NB)).
__device__ __host__ constexpr int cov_part_of(int u, int UB, int NPART) { return upper_part_of(u, UB, NPART); }

__device__ __host__ constexpr bool cov_needs_fb(int NB, int PART, int fb) {
    return upper_part_uses(NB, cov_parts(NB), PART, fb, false);
}

__device__ __host__ constexpr int cov_part_begin(int NB, int PART) { return upper_part_begin(NB, cov_parts(NB), PART); }

__device__ __host__ constexpr int cov_part_size(int NB, int PART) { return upper_part_size(NB, cov_parts(NB), PART); }

// LDS doubles per wave for the FP64 accumulators (0 when they are registers).
__device__ __host__ constexpr int cov_lds_doubles(int NB) {
    int m = 0;
    for (int p = 0; p < cov_parts(NB); p++) m = cov_part_size(NB, p) > m ? cov_part_size(NB, p) : m;
    return NB <= 3 ? 0 : m * 256;
}

template <int NB, int PART>
__device__ inline void covb_wave(const float *__restrict__ X, const float *__restrict__ rrow, int64_t i0, int64_t i1,
                                 const float (&mh)[NB], float inv, int g, int c, double *__restrict__ out,
                                 double *__restrict__ lacc) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov_parts(NB);
    constexpr int U0 = cov_part_begin(NB, PART);
    constexpr int NU = cov_part_size(NB, PART);
    constexpr int FLUSH = 2048;
    constexpr bool REG = NB <= 3;
    double accd[REG ? NU : 1][4];
    const int lane = 16 * g + c;
    if (REG) {
#pragma unroll
        for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
            for (int v = 0; v < 4; v++) accd[u][v] = 0.0;
    } else {
        // the wave's FP64 accumulators live in its own LDS slice [NU][4][64]
        for (int e = lane; e < NU * 256; e += 64) lacc[e] = 0.0;
    }
    // FP32 block sums into the FP64 accumulators
    auto flush = [&](const v4f (&acc)[NU]) {
        if (REG) {
#pragma unroll
            for (int u = 0; u < (REG ? NU : 1); u++)
#pragma unroll
                for (int v = 0; v < 4; v++) accd[u][v] += (double)acc[u][v];
        } else {
            // four blocks at a time (the compiler barrier keeps it from
            // loading every block's FP64 sums at once)
#pragma unroll
            for (int u = 0; u < NU; u++) {
#pragma unroll
                for (int v = 0; v < 4; v++) lacc[(u * 4 + v) * 64 + lane] += (double)acc[u][v];
                if ((u & 3) == 3) asm volatile("" ::: "memory");
            }
        }
    };
    // chunk bounds are multiples of 4096 points: every step has its partners
    if constexpr (NB <= 4) {
        // Software pipeline over 32-point steps: step s's MFMAs run from the
        // split operands T[s & 1] while the same instruction stream forms
        // and splits step s + 1 (independent registers, so the VALU work
        // issues beside the MFMAs); the point buffer that split frees is
        // refilled with step s + 3, and the posteriors, read from HBM once
        // per component, run RD steps ahead in a ring of RD (four at d = 64,
        // 10% faster there than two; two below, 4-11% faster there than
        // four; the unpipelined loop was 7-11% slower than either).  Loads
        // past the chunk re-read its last step (in bounds, never multiplied).
        constexpr int RD = NB == 4 ? 4 : 2;
        v8bf th[2][NB], tl[2][NB];
        float xv[2][NB][8], rb[RD][8];
        auto loadx = [&](int h, int64_t i) {
            const float *xp = X + (i + 8 * g) * DP + NB * c;
#pragma unroll
            for (int fb = 0; fb < NB; fb++)
                if (cov_needs_fb(NB, PART, fb))
#pragma unroll
                    for (int e = 0; e < 8; e++) xv[h][fb][e] = xp[e * DP + fb];
        };
        auto loadr = [&](int q, int64_t i) {
            const float *rp = rrow + i + 8 * g;
            const float4 r0 = *reinterpret_cast<const float4 *>(rp);
            const float4 r1 = *reinterpret_cast<const float4 *>(rp + 4);
            rb[q][0] = r0.x; rb[q][1] = r0.y; rb[q][2] = r0.z; rb[q][3] = r0.w;
            rb[q][4] = r1.x; rb[q][5] = r1.y; rb[q][6] = r1.z; rb[q][7] = r1.w;
        };
        auto split32 = [&](int h, int q, int o) {
            float sw[8];
#pragma unroll
            for (int e = 0; e < 8; e++) sw[e] = __builtin_amdgcn_sqrtf(rb[q][e]) * inv;
#pragma unroll
            for (int fb = 0; fb < NB; fb++) {
                if (!cov_needs_fb(NB, PART, fb)) continue;
                float t[8];
#pragma unroll
                for (int e = 0; e < 8; e++) t[e] = (xv[h][fb][e] - mh[fb]) * sw[e];
                split8(t, th[o][fb], tl[o][fb]);
            }
        };
        auto products = [&](int o, v4f (&acc)[NU]) {
            int u = 0;
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = ib; jb < NB; jb++, u++) {
                    if (cov_part_of(u, UB, NPART) != PART) continue;
                    acc[u - U0] = mfma3(th[o][ib], tl[o][ib], th[o][jb], tl[o][jb], acc[u - U0]);
                }
        };
        const int64_t last = i1 - 32;
        auto at = [&](int64_t i) { return i < i1 ? i : last; };
        // invariant at step s: T[s & 1] holds step s, point buffer (s + 1) & 1
        // step s + 1, buffer s & 1 step s + 2, ring slot (s + j) % RD the
        // posteriors of step s + j for j = 1..RD
        if (i0 < i1) {
            loadx(0, i0);
            loadx(1, at(i0 + 32));
#pragma unroll
            for (int q = 0; q < RD; q++) loadr(q, at(i0 + 32 * q));
            split32(0, 0, 0);
            loadx(0, at(i0 + 64));
            loadr(0, at(i0 + 32 * RD));
        }
        for (int64_t s0 = i0; s0 < i1; s0 += FLUSH) {
            const int64_t s1 = (s0 + FLUSH < i1) ? s0 + FLUSH : i1;
            v4f acc[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) acc[u] = (v4f){0.f, 0.f, 0.f, 0.f};
            // four steps per trip (FLUSH and the chunk are multiples of 128)
            for (int64_t i = s0; i < s1; i += 128) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int64_t is = i + 32 * j;  // step s
                    products(j & 1, acc);
                    split32((j + 1) & 1, (j + 1) % RD, (j + 1) & 1);
                    loadx((j + 1) & 1, at(is + 96));
                    loadr((j + 1) % RD, at(is + 32 * (RD + 1)));
                }
            }
            flush(acc);
        }
    } else {
        // d > 64 (one wave per SIMD): one 32-point register buffer, the next
        // step's loads issued once the current step is split
        float rv[8], xv[NB][8];
        auto load32 = [&](int64_t i) {
            const float *rp = rrow + i + 8 * g;
            const float4 r0 = *reinterpret_cast<const float4 *>(rp);
            const float4 r1 = *reinterpret_cast<const float4 *>(rp + 4);
            rv[0] = r0.x; rv[1] = r0.y; rv[2] = r0.z; rv[3] = r0.w;
            rv[4] = r1.x; rv[5] = r1.y; rv[6] = r1.z; rv[7] = r1.w;
            const float *xp = X + (i + 8 * g) * DP + NB * c;
#pragma unroll
            for (int fb = 0; fb < NB; fb++)
                if (cov_needs_fb(NB, PART, fb))
#pragma unroll
                    for (int e = 0; e < 8; e++) xv[fb][e] = xp[e * DP + fb];
        };
        if (i0 < i1) load32(i0);
        for (int64_t s0 = i0; s0 < i1; s0 += FLUSH) {
            const int64_t s1 = (s0 + FLUSH < i1) ? s0 + FLUSH : i1;
            v4f acc[NU];
#pragma unroll
            for (int u = 0; u < NU; u++) acc[u] = (v4f){0.f, 0.f, 0.f, 0.f};
            for (int64_t i = s0; i < s1; i += 32) {
                float sw[8];
#pragma unroll
                for (int e = 0; e < 8; e++) sw[e] = __builtin_amdgcn_sqrtf(rv[e]) * inv;
                v8bf th[NB], tl[NB];
#pragma unroll
                for (int fb = 0; fb < NB; fb++) {
                    if (!cov_needs_fb(NB, PART, fb)) continue;
                    float t[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) t[e] = (xv[fb][e] - mh[fb]) * sw[e];
                    split8(t, th[fb], tl[fb]);
                }
                if (i + 32 < i1) load32(i + 32);
                int u = 0;
#pragma unroll
                for (int ib = 0; ib < NB; ib++)
#pragma unroll
                    for (int jb = ib; jb < NB; jb++, u++) {
                        if (cov_part_of(u, UB, NPART) != PART) continue;
                        acc[u - U0] = mfma3(th[ib], tl[ib], th[jb], tl[jb], acc[u - U0]);
                    }
            }
            flush(acc);
        }
    }
#pragma unroll
    for (int u = 0; u < NU; u++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            out[(size_t)(U0 + u) * 256 + (4 * g + v) * 16 + c] = REG ? accd[REG ? u : 0][v] : lacc[(u * 4 + v) * 64 + lane];
}

template <int NB>
__device__ __attribute__((always_inline)) inline void covb_body(const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K,
                                 int KP, int d, const double *__restrict__ mu, const double *__restrict__ nk,
                                 int64_t chunk_rows, double *__restrict__ cpart) {
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov_parts(NB);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    const int part = logical % NPART, rest = logical / NPART;
    const int cg = rest % KG, chunk = rest / KG;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    const int64_t i0 = (int64_t)chunk * chunk_rows;
    const int64_t i1 = (i0 + chunk_rows < npad) ? i0 + chunk_rows : npad;
    float mh[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) mh[fb] = (NB * c + fb < d) ? (float)mu[(size_t)k * d + NB * c + fb] : 0.0f;
    const float inv = (float)(1.0 / sqrt(nk[k]));
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
    const float *rrow = Rt + (size_t)k * npad;
    extern __shared__ double lacc_all[];  // 4 waves x cov_lds_doubles(NB)
    double *lacc = lacc_all + (size_t)wave * cov_lds_doubles(NB);
    if (part == 0) covb_wave<NB, 0>(X, rrow, i0, i1, mh, inv, g, c, out, lacc);
    else if (NPART > 1) covb_wave<NB, (NPART > 1 ? 1 : 0)>(X, rrow, i0, i1, mh, inv, g, c, out, lacc);
}

// d <= 64: registers capped for two waves per SIMD (two 32-point buffers fit);
// wider d keeps the compiler's allocation (one wave per SIMD).
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_covb(
    const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K, int KP, int d,
    const double *__restrict__ mu, const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    covb_body<NB>(X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart);
}

template <int NB>
__global__ __launch_bounds__(256) void k_covb_wide(const float *__restrict__ X, const float *__restrict__ Rt,
                                                   int64_t npad, int K, int KP, int d,
                                                   const double *__restrict__ mu, const double *__restrict__ nk,
                                                   int64_t chunk_rows, double *__restrict__ cpart) {
    covb_body<NB>(X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart);
}

// The same over the nonzero posteriors only (the exact-zero argument of
// k_cov64_sparse): per wave and segment of kCovSegB points (segments
// strided over all points, as there) the nonzero
// posteriors are listed in point order in the wave's LDS slice (lane l
// reads points 4l..4l+3 of each 256), and each 32-entry group of the list is
// one K = 32 step of the dense form (lane group g takes entries 8g..8g+7 of
// the group; entries past the end take r = 0), the next group's points
// loaded during the current group's products.  FP32 accumulators are
// flushed into FP64 (registers for d <= 48, else the wave's partial slice)
// before a segment's list would take them past 2048 terms.
constexpr int kCovSegB = 1024;

template <int NB, int PART>
__device__ inline void covb_sparse(const float *__restrict__ X, const float *__restrict__ rrow, int64_t seg0,
                                   int64_t nseg, int64_t nch, const float (&mh)[NB], float inv, int g, int c,
                                   double *__restrict__ out, int *wi, float *wr) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov_parts(NB);
    constexpr int U0 = cov_part_begin(NB, PART);
    constexpr int NU = cov_part_size(NB, PART);
    constexpr int FLUSH = 2048;
    constexpr bool REG = NB <= 3;
    const int lane = 16 * g + c;
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
    int pending = 0;
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

    constexpr int NJ = kCovSegB / 256;
    v4f rs[NJ];
    auto loadseg = [&](int64_t s) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
            rs[j] = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(rrow + s + 256 * j + 4 * lane));
    };
    // entries e0 + 8 g + e of the list (called with e0 < m): weights and the
    // lane's NB features of each point
    auto fetch = [&](int64_t s, int m, int e0, float (&sw)[8], float (&xv)[NB][8]) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int en = e0 + 8 * g + e;
            const bool ok = en < m;
            const int ec = ok ? en : 0;
            const int off = wi[ec];
            sw[e] = ok ? __builtin_amdgcn_sqrtf(wr[ec]) * inv : 0.0f;
            float xr[NB];
            load_consecutive<NB>(X + (size_t)(s + off) * DP + NB * c, xr);
#pragma unroll
            for (int fb = 0; fb < NB; fb++) xv[fb][e] = xr[fb];
        }
    };
    auto products = [&](const float (&sw)[8], const float (&xv)[NB][8]) {
        v8bf th[NB], tl[NB];
#pragma unroll
        for (int fb = 0; fb < NB; fb++) {
            if (!cov_needs_fb(NB, PART, fb)) continue;
            float t[8];
#pragma unroll
            for (int e = 0; e < 8; e++) t[e] = (xv[fb][e] - mh[fb]) * sw[e];
            split8(t, th[fb], tl[fb]);
        }
        int u = 0;
#pragma unroll
        for (int ib = 0; ib < NB; ib++)
#pragma unroll
            for (int jb = ib; jb < NB; jb++, u++) {
                if (cov_part_of(u, UB, NPART) != PART) continue;
                acc[u - U0] = mfma3(th[ib], tl[ib], th[jb], tl[jb], acc[u - U0]);
            }
    };

    // segments seg0, seg0 + nch, ... (as cov64_sparse); npad is a multiple
    // of 4096 points, so every segment is whole
    if (seg0 < nseg) loadseg(seg0 * kCovSegB);
    for (int64_t sg = seg0; sg < nseg; sg += nch) {
        const int64_t s = sg * kCovSegB;
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
        if (sg + nch < nseg) loadseg((sg + nch) * kCovSegB);
        const int padded = (m + 31) & ~31;
        if (pending + padded > FLUSH) flush();
        pending += padded;
        float swA[8], swB[8], xvA[NB][8], xvB[NB][8];
        if (m > 0) fetch(s, m, 0, swA, xvA);
        for (int e0 = 0; e0 < m; e0 += 64) {
            if (e0 + 32 < m) fetch(s, m, e0 + 32, swB, xvB);
            products(swA, xvA);
            if (e0 + 64 < m) fetch(s, m, e0 + 64, swA, xvA);
            if (e0 + 32 < m) products(swB, xvB);
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
__device__ __attribute__((always_inline)) inline void covb_sparse_body(
    const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K, int KP, int d,
    const double *__restrict__ mu, const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart,
    int *lidx, float *lrv) {
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int NPART = cov_parts(NB);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    const int part = logical % NPART, rest = logical / NPART;
    const int cg = rest % KG, chunk = rest / KG;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    float mh[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) mh[fb] = (NB * c + fb < d) ? (float)mu[(size_t)k * d + NB * c + fb] : 0.0f;
    const float inv = (float)(1.0 / sqrt(nk[k]));
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
    const float *rrow = Rt + (size_t)k * npad;
    int *wi = lidx + wave * kCovSegB;
    float *wr = lrv + wave * kCovSegB;
    const int64_t nseg = npad / kCovSegB, nch = (npad + chunk_rows - 1) / chunk_rows;
    if (part == 0) covb_sparse<NB, 0>(X, rrow, chunk, nseg, nch, mh, inv, g, c, out, wi, wr);
    else if constexpr (NPART > 1) covb_sparse<NB, 1>(X, rrow, chunk, nseg, nch, mh, inv, g, c, out, wi, wr);
}

// registers capped for two waves per SIMD for d <= 64, the compiler's
// allocation above (as k_covb and k_covb_wide)
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_covb_sparse(
    const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K, int KP, int d,
    const double *__restrict__ mu, const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    __shared__ int lidx[4 * kCovSegB];
    __shared__ float lrv[4 * kCovSegB];
    covb_sparse_body<NB>(X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart, lidx, lrv);
}

template <int NB>
__global__ __launch_bounds__(256) void k_covb_sparse_wide(
    const float *__restrict__ X, const float *__restrict__ Rt, int64_t npad, int K, int KP, int d,
    const double *__restrict__ mu, const double *__restrict__ nk, int64_t chunk_rows, double *__restrict__ cpart) {
    __shared__ int lidx[4 * kCovSegB];
    __shared__ float lrv[4 * kCovSegB];
    covb_sparse_body<NB>(X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart, lidx, lrv);
}

template <int NB>
static hipError_t covb_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks * cov_parts(NB);
    if (c->cov_sparse) {
        if constexpr (NB <= 4)
            hipLaunchKernelGGL(k_covb_sparse<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad, c->K,
                               c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
        else
            hipLaunchKernelGGL(k_covb_sparse_wide<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt32, c->npad,
                               c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
        return hipGetLastError();
    }
    const size_t lds = (size_t)4 * cov_lds_doubles(NB) * sizeof(double);
    auto kern = NB <= 4 ? k_covb<NB> : k_covb_wide<NB>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), lds, c->stream, c->Xe(), c->Rt32, c->npad, c->K, c->KP, c->d,
                       c->mu, c->nk, c->c_chunk_rows, c->cpart);
    return hipGetLastError();
}

hipError_t launch_covb_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1: return covb_t<1>(c);
        case 2: return covb_t<2>(c);
        case 3: return covb_t<3>(c);
        case 4: return covb_t<4>(c);
        case 5: return covb_t<5>(c);
        case 6: return covb_t<6>(c);
        case 7: return covb_t<7>(c);
        case 8: return covb_t<8>(c);
    }
    return hipErrorInvalidValue;
}

int estepb_stage(int NB) { return stage_words_of(NB); }

}  // namespace fem
