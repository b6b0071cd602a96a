// CachedEM.h -- the per-point BLAS-2 EM variant (EM2/CachedEM.h:8-49).  Same
// mathematics as MatrixEM; on the GPU it runs the same kernels.
#pragma once
#include "EMAlgorithm.h"

class CachedEM : public EMAlgorithm {
public:
    CachedEM(int ncols, long nrows, int ncl, const char *ReducerName, int anDensBlocks = 1,
             int amStepBlocks = 1);
    const char *GetName() const override { return "CachedEM"; }
    void PrintNumaLocalityInfo() override {}
};
