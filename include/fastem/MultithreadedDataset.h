// MultithreadedDataset.h -- the FP32 .bin dataset (Util/MultithreadedDataset.h:64-135).
// Format: int32 N, int32 d; if N == -1 an int64 N follows; then N*d FP32
// row-major (Util/MultithreadedDataset.cpp:119-134).
#pragma once
#include <vector>

class MultithreadedDataset {
    long nRows = 0, nTotalRows = 0;
    int nCols = 0;
    std::vector<float> Data;
    std::vector<double> Means, StdDevs;
    double TraceCov = 0.0;

public:
    explicit MultithreadedDataset(int nBlocks = 1);
    // Throws FileError (std::runtime_error) on I/O failure.
    void Load(const char *fname);
    static void QueryHeader(const char *fname, long &nRows, int &nCols);
    // Writes the header form LoadHeader reads (8-byte N only when N > INT32_MAX).
    static void Save(const char *fname, const float *X, long nRows, int nCols);
    // In-memory dataset (copies X).
    void SetData(const float *X, long nRows, int nCols);
    void ComputeStatistics();
    long GetRowCount() const { return nRows; }
    long GetTotalRowCount() const { return nTotalRows; }
    int GetColCount() const { return nCols; }
    const float *GetRowNew(long i) const { return Data.data() + (size_t)i * nCols; }
    const float *GetData() const { return Data.data(); }
    double GetCovMatrixTrace() const { return TraceCov; }
    double GetMean(int j) const { return Means[j]; }
    double GetStdDev(int j) const { return StdDevs[j]; }
};
