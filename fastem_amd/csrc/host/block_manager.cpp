// block_manager.cpp -- EM2/BlockManager.cpp:10-28.
#include <cstdio>
#include <stdexcept>

#include "../../../include/fastem/BlockManager.h"

void BlockManagerBase::ComputeEMBlockCounts(long nRows, int nCols, int K, double L3,
                                            int &nDensBlocks, int &nMStepBlocks) {
    L3 *= (1024 * 1024);
    nDensBlocks = (int)((double)nRows * (16 * (double)nCols + 16 * (double)K) / L3 * 1.2);
    if (nDensBlocks < 1) nDensBlocks = 1;
    nMStepBlocks = nDensBlocks;
}

void BlockManagerBase::ParseEMBlockParams(const char *BlockParams, long nRows, int nCols, int K,
                                          int &nDensBlocks, int &nMStepBlocks) {
    double L3;
    if (BlockParams == nullptr) {
        nDensBlocks = nMStepBlocks = 1;
        return;
    }
    if (sscanf(BlockParams, "automax:%lf", &L3) == 1) {
        ComputeEMBlockCounts(nRows, nCols, K, L3, nDensBlocks, nMStepBlocks);
    } else if (sscanf(BlockParams, "%d:%d", &nDensBlocks, &nMStepBlocks) != 2)
        throw std::invalid_argument("Bad value of MStepLoopParam");
}
