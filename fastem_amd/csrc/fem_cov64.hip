// fem_cov64.hip -- stage (c) of the FP64 form: weighted SYRK per component
// and chunk of points on v_mfma_f64_16x16x4f64 (EM2/MatrixEM.cpp:53-67,90-99,
// 346-375); the per-workgroup body is cov64_unit in fem_tiles.h, shared with
// the batched engine.  Kept in its own file because for d > 48 it runs
// faster with the MFMA accumulators in AGPRs (7% at d = 64), while the E-step
// and means kernels of fem_kernels.hip are built with VGPR accumulators; the
// d <= 48 instances, 2-8% faster with VGPR accumulators, are built from
// fem_cov64_narrow.hip.
#include "fem_cov64_kernel.h"

namespace fem {

hipError_t launch_cov64_partials(fem_ctx *c) {
    switch (c->NB) {
        case 1:
        case 2:
        case 3: return launch_cov64_narrow(c);
        case 4: return cov_t<4>(c);
        case 5: return cov_t<5>(c);
        case 6: return cov_t<6>(c);
        case 7: return cov_t<7>(c);
        case 8: return cov_t<8>(c);
    }
    return hipErrorInvalidValue;
}

}  // namespace fem
