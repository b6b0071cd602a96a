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

template <int NB>
inline hipError_t cov_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks * cov64_parts(NB);
    hipLaunchKernelGGL(k_cov64<NB>, dim3(blocks), dim3(256), 0, c->stream, c->X, c->Rt, c->npad,
                       c->K, c->KP, c->d, c->mu, c->nk, c->c_chunk_rows, c->cpart);
    return hipGetLastError();
}

}  // namespace fem
#endif
