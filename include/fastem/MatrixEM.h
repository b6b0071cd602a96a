// MatrixEM.h -- the blocked-BLAS EM variants (EM2/MatrixEM.h).  On the GPU the
// two variants differ only in name: both run the fused E-step and the MFMA
// M-step.  The reducer name is validated as in the reference
// (EM2/MatrixEM.cpp:16-23); on the device the partial sums are combined in a
// fixed order in FP64.
#pragma once
#include "BlockManager.h"
#include "EMAlgorithm.h"

class MatrixEMBase : public EMAlgorithm {
protected:
    int nDensBlocks, nMStepBlocks;

public:
    MatrixEMBase(int ncols, long nrows, int ncl, const char *ReducerName, int anDensBlocks = 1,
                 int anMStepBlocks = 1);
};

class MatrixEM : public MatrixEMBase {
public:
    MatrixEM(int ncols, long nrows, int ncl, const char *ReducerName, int anDensBlocks = 1,
             int anMStepBlocks = 1)
        : MatrixEMBase(ncols, nrows, ncl, ReducerName, anDensBlocks, anMStepBlocks) {}
    const char *GetName() const override { return "MatrixEM"; }
};

class SimpleMatrixEM : public MatrixEMBase {
public:
    SimpleMatrixEM(int ncols, long nrows, int ncl, const char *ReducerName, int anDensBlocks = 1,
                   int anMStepBlocks = 1)
        : MatrixEMBase(ncols, nrows, ncl, ReducerName, anDensBlocks, anMStepBlocks) {}
    const char *GetName() const override { return "SimpleMatrixEM"; }
};
