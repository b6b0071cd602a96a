// BlockManager.h -- row-block parameters of the reference (EM2/BlockManager.h).
// The `-N automax:<L3 MB>` / `-N a:b` spec is parsed with identical results and
// exceptions; on the GPU the counts are reported but rows are tiled by the
// kernels themselves.
#pragma once

class BlockManagerBase {
protected:
    static void ComputeEMBlockCounts(long nRows, int nCols, int K, double L3, int &nDensBlocks,
                                     int &nMStepBlocks);

public:
    static void ParseEMBlockParams(const char *BlockParams, long nRows, int nCols, int K,
                                   int &nDensBlocks, int &nMStepBlocks);
};

// Util/OpenMP.h:37-46
template <typename T>
inline void FindOpenMPItems(T nItems, T &Offset, T &Count, int Size, int Id) {
    Count = nItems / Size;
    int SubRemainder = (int)(nItems % Size);
    if (Id < SubRemainder) {
        Count++;
        Offset = Count * Id;
    } else {
        Offset = Count * Id + SubRemainder;
    }
}
