// fem_cov64.hip -- stage (c) of the FP64 form: weighted SYRK per component
// and chunk of points on v_mfma_f64_16x16x4f64 (EM2/MatrixEM.cpp:53-67,90-99,
// 346-375); the per-workgroup body is cov64_unit in fem_tiles.h, shared with
// the batched engine.  Kept in its own file because it runs faster with the
// MFMA accumulators in AGPRs, while the E-step and means kernels of
// fem_kernels.hip are built with VGPR accumulators.
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
static hipError_t cov_t(fem_ctx *c) {
    const int blocks = (c->KP / 4) * c->c_chunks * cov64_parts(NB);
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

}  // namespace fem
