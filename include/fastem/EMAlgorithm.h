// EMAlgorithm.h -- the EM driver class (EM2/EMAlgorithm.h:26-81), with the same
// public interface.  The protected hooks that the reference implements with
// OpenMP + BLAS (CopyDataset, InitDensities, PrecomputeDensityMatrix,
// ComputeLogLike, MStep) run on one MI355X through the C ABI in fastem_abi.h;
// the iteration loop, convergence test and exceptions stay on the host.
#pragma once
#include <vector>

#include "../fastem_abi.h"
#include "EMException.h"
#include "GaussianMixture.h"
#include "MultithreadedDataset.h"

class EMAlgorithm {
protected:
    int nCols;
    int K;
    int IterCounter;
    long TotalRowCount;
    long nRows;

    int MaxSweeps;
    OPTFLOAT Eps;
    OPTFLOAT AbortThr;
    OPTFLOAT RegCoef;

    DynamicArray<EigenMatrix> Covs;
    EigenMatrix Means;
    EigenVector Probs;
    // Per-component partial log-likelihoods.  The reference sizes them
    // (EM2/EMAlgorithm.cpp:53,63) but never writes them and ComputePLLs has
    // no caller, so they stay zero; the same holds here.
    DynamicArray<OPTFLOAT> PLLs;
    OPTFLOAT LogLike;

    fem_ctx *Ctx;
    int Precision;
    bool HostParamsNewer;     // host Probs/Means/Covs must be uploaded
    bool DeviceParamsNewer;   // device holds parameters the host has not fetched
    bool DeviceRegularized;   // the last MStep already applied RegularizeCovariances

    void Check(int rc, const char *what) const;
    void FetchParams();

    virtual void MPIReduceLogL(OPTFLOAT &) {}
    OPTFLOAT FindLargestConditionNumber();
    // the same on the host copies of the covariances alone (no fetch): run on
    // a host thread while the device computes the next E-step
    OPTFLOAT LargestHostConditionNumber() const;
    // the parameters fem_get_params_start queued, into Probs/Means/Covs
    // (cov: K d^2 doubles of staging)
    void FinishStagedFetch(double *cov);
    OPTFLOAT FindMinTrace();
    void DecodeMixture(const GaussianMixture &G);
    void ComputePLLs() {}
    virtual void EStep() {}
    virtual void InitDensities();
    virtual void PrecomputeDensityMatrix();
    virtual OPTFLOAT ComputeLogLike() { return LogLike; }
    virtual void MStep();
    virtual void CopyDataset(const MultithreadedDataset &Data);
    void RegularizeCovariances();

public:
    EMAlgorithm(int ncols, long nrows, int ncl);
    virtual ~EMAlgorithm();
    // Arithmetic of the contractions for algorithms created afterwards
    // (FEM_PREC_FP64 by default; the FASTEM_PRECISION environment variable,
    // "fp64" | "fp32" | "bf16x3", sets the initial value).
    static void SetDefaultPrecision(int prec);
    static int GetDefaultPrecision();

    void ComputePosterriorCorrelation(EigenMatrix &Corr);
    // Class of each row = the component of largest posterior from the last
    // E-step, ties to the lower index (EM2/EMAlgorithm.cpp:107-122); sizes
    // Classes to the row count.
    void RetrieveClasses(LargeVector<int> &Classes) const;
    int GetIterCounter() const { return IterCounter; }
    void ClearIterCounter() { IterCounter = 0; }
    int GetK() const { return K; }
    OPTFLOAT GetPLL(int k) const { return PLLs[k]; }
    void SetParameters(int aMaxSweeps, double aEps, double aAbortThr, double aRegCoef = 0.0);
    virtual OPTFLOAT Train(const MultithreadedDataset &Data, GaussianMixture &G, int Verbosity, bool SVD);
    virtual OPTFLOAT TrainNoCopy(GaussianMixture &G, int Verbosity, bool SVD);
    virtual void PrintNumaLocalityInfo();
    virtual const char *GetName() const = 0;
    fem_ctx *GetContext() const { return Ctx; }
};

// EM2/EMAlgorithm.cpp:247-269: "matrixem" (default), "simplematrixem",
// "cachedem"; throws std::invalid_argument("Unknown EM algorithm").
EMAlgorithm *CreateEMAlgorithm(const char *Name, const char *RName, const char *MName, int nCols,
                               long nRows, long nTotalRows, int nCl,
                               const char *MStepLoopParam = nullptr);
