// fem_host.h -- host-side helpers shared inside libfastem.
#pragma once
#include <cstdint>

namespace fem_host {
// See fem_tools.cpp.  seeding: 0 = Forgy random partition, 1 = k-means++.
int kmeans_mixture(int64_t n, int d, int K, const float *X, double (*rnd)(void *), void *rs,
                   int seeding, int max_sweeps, double *w, double *mu, double *cov);
}  // namespace fem_host
