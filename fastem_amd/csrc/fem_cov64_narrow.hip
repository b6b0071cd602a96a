// fem_cov64_narrow.hip -- the FP64 covariance kernel (fem_cov64_kernel.h)
// for d <= 48, built with VGPR accumulators (-mllvm -amdgpu-mfma-vgpr-form),
// which measured 2-8% faster there than the AGPR form that fem_cov64.hip
// keeps for d > 48.
#include "fem_cov64_kernel.h"

namespace fem {

hipError_t launch_cov64_narrow(fem_ctx *c) {
    switch (c->NB) {
        case 1: return cov_t<1>(c);
        case 2: return cov_t<2>(c);
        case 3: return cov_t<3>(c);
    }
    return hipErrorInvalidValue;
}

}  // namespace fem
