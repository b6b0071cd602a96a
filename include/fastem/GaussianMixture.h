// GaussianMixture.h -- the mixture container and the .cmp file format
// (EM2/GaussianMixture.h:17-58).  .cmp = int32 K, int32 d, then per component
// d*d FP64 covariance (row-major), d FP64 mean, FP64 weight; then one FP64 0.0.
#pragma once
#include <cstdio>

#include "MultithreadedDataset.h"
#include "NormalDensity.h"

class GaussianMixture {
    int nComponents;
    int nCols;
    DynamicArray<NormalDensity *> Densities;
    DynamicArray<OPTFLOAT> Weights;
    void AllocateData(int Comps);
    void FreeData();

public:
    GaussianMixture(int Comp, int Cols);
    explicit GaussianMixture(FILE *pF);
    GaussianMixture(const GaussianMixture &x);
    GaussianMixture &operator=(const GaussianMixture &x);
    virtual ~GaussianMixture();

    int GetComponentCount() const { return nComponents; }
    int GetDimension() const { return nCols; }
    OPTFLOAT GetWeight(int Component) const { return Weights[Component]; }
    void SetWeight(int Component, OPTFLOAT Weight) { Weights[Component] = Weight; }
    const EigenMatrix GetCovariance(int Component) const { return Densities[Component]->GetCovariance(); }
    const EigenRowVector GetMean(int Component) const { return Densities[Component]->GetMean(); }
    const NormalDensity &GetDensity(int Component) const { return *Densities[Component]; }
    void Init(const DynamicArray<OPTFLOAT> &aWeights, const DynamicArray<EigenMatrix> &aCovs,
              const EigenMatrix &aMeans);
    void InitComponent(int Component, OPTFLOAT Weight, const EigenMatrix &Cov, const EigenRowVector &Mean);
    void SplitComponent(int Idx, OPTFLOAT W1, const EigenRowVector &M1, const EigenMatrix &C1,
                        OPTFLOAT W2, const EigenRowVector &M2, const EigenMatrix &C2);
    void MergeComponents(int Idx1, int Idx2, OPTFLOAT W, const EigenRowVector &M, const EigenMatrix &C);
    void NormalizeWeights();

    OPTFLOAT Evaluate(const float *row) const;
    void ComputePosteriors(const float *row, EigenRowVector &Posteriors) const;
    int Classify(const float *row);
    void ClassifyData(const MultithreadedDataset &D, DynamicArray<int> &ClassNums);
    void WriteClasses(const char *fname, const MultithreadedDataset &D);
    OPTFLOAT FindLargestConditionNumber() const;
    // Bhattacharyya distance of component Comp to component CG of G
    // (EM2/GaussianMixture.cpp:81-83).
    OPTFLOAT ComponentBhattDist(BhattDist &Dist, int Comp, const GaussianMixture &G, int CG) const;
    void Dump() const;
    void Write(FILE *pF);
    void Write(const char *fname);
};
