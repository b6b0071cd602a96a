// dataset.cpp -- the .bin reader/writer and column statistics
// (Util/MultithreadedDataset.cpp:119-134,179-301).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>

#include "../../../include/fastem/MultithreadedDataset.h"

MultithreadedDataset::MultithreadedDataset(int) {}

static void read_header(FILE *f, long &nRows, int &nCols) {
    int32_t small, cols;
    if (fread(&small, 4, 1, f) != 1 || fread(&cols, 4, 1, f) != 1) throw std::runtime_error("read failed");
    if (small == -1) {
        int64_t big;
        if (fread(&big, 8, 1, f) != 1) throw std::runtime_error("read failed");
        nRows = (long)big;
    } else {
        nRows = small;
    }
    nCols = cols;
    if (nRows < 0 || nCols <= 0) throw std::runtime_error("Invalid dataset header");
}

void MultithreadedDataset::QueryHeader(const char *fname, long &nRows, int &nCols) {
    FILE *f = fopen(fname, "rb");
    if (!f) throw std::runtime_error("Cannot open dataset");
    try {
        read_header(f, nRows, nCols);
    } catch (...) {
        fclose(f);
        throw;
    }
    fclose(f);
}

void MultithreadedDataset::Load(const char *fname) {
    FILE *f = fopen(fname, "rb");
    if (!f) throw std::runtime_error("Cannot open dataset");
    long n;
    int d;
    read_header(f, n, d);
    Data.resize((size_t)n * d);
    const size_t got = fread(Data.data(), sizeof(float), Data.size(), f);
    fclose(f);
    if (got != Data.size()) throw std::runtime_error("Invalid number of bytes read");
    nRows = nTotalRows = n;
    nCols = d;
    ComputeStatistics();
}

void MultithreadedDataset::Save(const char *fname, const float *X, long n, int d) {
    FILE *f = fopen(fname, "wb");
    if (!f) throw std::runtime_error("Cannot open dataset for writing");
    if (n > (long)std::numeric_limits<int32_t>::max()) {
        const int32_t m1 = -1, cols = d;
        const int64_t big = n;
        fwrite(&m1, 4, 1, f);
        fwrite(&cols, 4, 1, f);
        fwrite(&big, 8, 1, f);
    } else {
        const int32_t small = (int32_t)n, cols = d;
        fwrite(&small, 4, 1, f);
        fwrite(&cols, 4, 1, f);
    }
    const size_t put = fwrite(X, sizeof(float), (size_t)n * d, f);
    fclose(f);
    if (put != (size_t)n * d) throw std::runtime_error("Short write of dataset");
}

void MultithreadedDataset::SetData(const float *X, long n, int d) {
    Data.assign(X, X + (size_t)n * d);
    nRows = nTotalRows = n;
    nCols = d;
    ComputeStatistics();
}

// Column means, then sum of squared deviations / (N-1); TraceCov is their
// sum.  Fixed 64-way row partition combined in order: deterministic.
void MultithreadedDataset::ComputeStatistics() {
    const int d = nCols;
    const int P = 64;
    Means.assign(d, 0.0);
    StdDevs.assign(d, 0.0);
    std::vector<double> part((size_t)P * d);
    auto pass = [&](bool second) {
        std::fill(part.begin(), part.end(), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int p = 0; p < P; p++) {
            const long q = nRows / P, r = nRows % P;
            const long b = p * q + (p < r ? p : r), e = b + q + (p < r ? 1 : 0);
            double *acc = part.data() + (size_t)p * d;
            for (long i = b; i < e; i++)
                for (int j = 0; j < d; j++) {
                    const double x = Data[(size_t)i * d + j];
                    acc[j] += second ? (x - Means[j]) * (x - Means[j]) : x;
                }
        }
        std::vector<double> tot(d, 0.0);
        for (int p = 0; p < P; p++)
            for (int j = 0; j < d; j++) tot[j] += part[(size_t)p * d + j];
        return tot;
    };
    std::vector<double> s1 = pass(false);
    for (int j = 0; j < d; j++) Means[j] = s1[j] / (double)nTotalRows;
    std::vector<double> s2 = pass(true);
    TraceCov = 0.0;
    for (int j = 0; j < d; j++) {
        const double v = s2[j] / ((double)nTotalRows - 1.0);
        TraceCov += v;
        StdDevs[j] = std::sqrt(v);
    }
}
