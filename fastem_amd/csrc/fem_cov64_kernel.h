// fem_cov64_kernel.h -- the FP64 covariance kernel and its launcher, shared by
// fem_cov64.hip (d > 48, AGPR accumulators) and fem_cov64_narrow.hip
// (d <= 48, VGPR accumulators); each file instantiates its own NB.
#ifndef FEM_COV64_KERNEL_H
#define FEM_COV64_KERNEL_H
#include "fem_internal.h"
#include "fem_tiles.h"

namespace fem {

// ---------------------------------------------------------------------------
// Workgroups are remapped so the ones that share a chunk of X run on the same
// XCD (dispatch is round-robin over 8 XCDs).
template <int NB>
__global__ __launch_bounds__(256) void k_cov64(const float *__restrict__ X,
                                               const double *__restrict__ Rt, int64_t npad,
                                               int K, int KP, int d,
                                               const double *__restrict__ mu,
                                               const double *__restrict__ nk, int64_t chunk_rows,
                                               double *__restrict__ cpart) {
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    // the part (d > 96) varies fastest, then the component group
    constexpr int NPART = cov64_parts(NB);
    const int part = logical % NPART, rest = logical / NPART;
    cov64_unit<NB>(rest % KG, rest / KG, X, Rt, npad, K, KP, d, mu, nk, chunk_rows, cpart, part);
}

// ---------------------------------------------------------------------------
// Stage (c), FP64, over the nonzero posteriors only.  A point whose posterior
// of component k is zero adds exactly zero to component k's sum (w t = 0 in
// each of its products), and in a separated mixture most posteriors are zero:
// exp underflows to +0 for log-density gaps beyond about 745.  Per wave
// (component k) and segment of kCovSeg points of its chunk (segments
// chunk, chunk + chunks, ... of all points), the segment's
// posteriors are read once (lane l: points 2l and 2l + 1 of each 128, one
// 16-byte load), the points with r != 0 are listed in the wave's LDS slice
// in point order (ballots and mbcnt), and the products of the dense form run
// over the list in steps of 8 entries (two groups of 4; entries past the end
// take r = 0 and the first listed point, adding exact zeros).  The sum is
// that of the nonzero terms, grouped 4 at a time in list order.  The next
// segment's posteriors are in flight during the current segment's products,
// and the next step's points during the current step's.
constexpr int kCovSeg = 256;

typedef double v2d __attribute__((ext_vector_type(2)));

template <int NB, int NPART, int PART>
__device__ inline void cov64_sparse(int cg, int chunk, const float *__restrict__ X, const double *__restrict__ Rt,
                                    int64_t npad, int K, int d, const double *__restrict__ mu,
                                    const double *__restrict__ nk, int64_t chunk_rows, int KP,
                                    double *__restrict__ cpart, int *lidx, double *lrv) {
    constexpr int DP = 16 * NB;
    constexpr int UB = NB * (NB + 1) / 2;
    constexpr int U0 = upper_part_begin(NB, NPART, PART);
    constexpr int NU = upper_part_size(NB, NPART, PART);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int k = 4 * cg + wave;
    if (k >= K) return;
    // the chunk's segments are strided over the points (segments chunk,
    // chunk + chunks, ...), so that in a sorted slot order, where each
    // component's nonzero posteriors are contiguous, every chunk (and so
    // every workgroup) takes an equal share of them
    const int64_t nseg = npad / kCovSeg, nch = (npad + chunk_rows - 1) / chunk_rows;
    double muv[NB];
#pragma unroll
    for (int fb = 0; fb < NB; fb++) muv[fb] = (NB * c + fb < d) ? mu[(size_t)k * d + NB * c + fb] : 0.0;
    const double invn = 1.0 / nk[k];
    v4d acc[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) acc[u] = (v4d){0.0, 0.0, 0.0, 0.0};
    const double *rrow = Rt + (size_t)k * npad;
    int *wi = lidx + wave * kCovSeg;
    double *wr = lrv + wave * kCovSeg;

    // the posteriors of one segment: lane l holds points 128 j + 2 l, + 1
    constexpr int NJ = kCovSeg / 128;
    v2d rs[NJ];
    auto loadseg = [&](int64_t s) {
#pragma unroll
        for (int j = 0; j < NJ; j++)
            rs[j] = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(rrow + s + 128 * j + 2 * lane));
    };
    // entries e0 + 4 q + g of the list (q = 0, 1): weight and point features;
    // past the end (called only with e0 < m, so entry 0 is listed) entry 0's
    // point with weight 0
    auto fetch = [&](int64_t s, int m, int e0, double (&rv)[2], float (&xv)[2][NB]) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int e = e0 + 4 * q + g;
            const bool ok = e < m;
            const int ec = ok ? e : 0;
            const int off = wi[ec];
            rv[q] = ok ? wr[ec] : 0.0;
            load_consecutive<NB>(X + (size_t)(s + off) * DP + NB * c, xv[q]);
        }
    };
    auto products = [&](const double (&rv)[2], const float (&xv)[2][NB]) {
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
    };

    // npad is a multiple of 4096 points: every segment is whole
    if (chunk < nseg) loadseg((int64_t)chunk * kCovSeg);
    for (int64_t sg = chunk; sg < nseg; sg += nch) {
        const int64_t s = sg * kCovSeg;
        int m = 0;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const bool n0 = rs[j].x != 0.0, n1 = rs[j].y != 0.0;
            const uint64_t b0 = __ballot(n0), b1 = __ballot(n1);
            const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b0, 0u)) +
                              (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b1, 0u));
            const int at = m + below;
            const int off = 128 * j + 2 * lane;
            if (n0) {
                wi[at] = off;
                wr[at] = rs[j].x;
            }
            if (n1) {
                wi[at + n0] = off + 1;
                wr[at + n0] = rs[j].y;
            }
            m += __builtin_popcountll(b0) + __builtin_popcountll(b1);
        }
        if (sg + nch < nseg) loadseg((sg + nch) * kCovSeg);
        double rvA[2], rvB[2];
        float xvA[2][NB], xvB[2][NB];
        if (m > 0) fetch(s, m, 0, rvA, xvA);
        for (int e0 = 0; e0 < m; e0 += 16) {
            if (e0 + 8 < m) fetch(s, m, e0 + 8, rvB, xvB);
            products(rvA, xvA);
            if (e0 + 16 < m) fetch(s, m, e0 + 16, rvA, xvA);
            if (e0 + 8 < m) products(rvB, xvB);
        }
    }
    double *out = cpart + ((size_t)chunk * KP + k) * UB * 256;
#pragma unroll
    for (int u = 0; u < NU; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)(U0 + u) * 256 + (4 * v + g) * 16 + c] = acc[u][v];
}

// The sparse form's kernel: the dense kernel's grid and XCD remapping, and
// 48 KB of LDS for the four waves' lists.
template <int NB>
__global__ __launch_bounds__(256) void k_cov64_sparse(const float *__restrict__ X,
                                                      const double *__restrict__ Rt, int64_t npad,
                                                      int K, int KP, int d,
                                                      const double *__restrict__ mu,
                                                      const double *__restrict__ nk, int64_t chunk_rows,
                                                      double *__restrict__ cpart) {
    __shared__ int lidx[4 * kCovSeg];
    __shared__ double lrv[4 * kCovSeg];
    const int KG = KP / 4;
    const int nblk = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
    constexpr int NPART = cov64_parts(NB);
    const int part = logical % NPART, rest = logical / NPART;
    const int cg = rest % KG, chunk = rest / KG;
    if constexpr (NPART > 1) {
        if (part == 0)
            cov64_sparse<NB, NPART, 0>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart, lidx, lrv);
        else
            cov64_sparse<NB, NPART, 1>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart, lidx, lrv);
    } else {
        cov64_sparse<NB, 1, 0>(cg, chunk, X, Rt, npad, K, d, mu, nk, chunk_rows, KP, cpart, lidx, lrv);
    }
}

template <int NB>
inline hipError_t cov_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks * cov64_parts(NB);
    if (c->cov_sparse)
        hipLaunchKernelGGL(k_cov64_sparse<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt, c->npad,
                           c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    else
        hipLaunchKernelGGL(k_cov64<NB>, dim3(blocks), dim3(256), 0, c->stream, c->Xe(), c->Rt, c->npad,
                           c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    return hipGetLastError();
}

}  // namespace fem
#endif
