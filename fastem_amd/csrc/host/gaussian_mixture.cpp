// gaussian_mixture.cpp -- EM2/GaussianMixture.cpp:19-244 semantics.
#include <cstdio>

#include "../../../include/fastem/GaussianMixture.h"

void GaussianMixture::AllocateData(int Comps) {
    Densities.SetSize(Comps);
    Weights.SetSize(Comps);
    for (int i = 0; i < Comps; i++) Densities[i] = new NormalDensity(nCols);
}

void GaussianMixture::FreeData() {
    for (long i = 0; i < Densities.GetSize(); i++) delete Densities[i];
    Densities.SetSize(0);
}

GaussianMixture::GaussianMixture(int Comp, int Cols) : nComponents(Comp), nCols(Cols) { AllocateData(Comp); }

GaussianMixture::~GaussianMixture() { FreeData(); }

GaussianMixture::GaussianMixture(const GaussianMixture &x) : nComponents(x.nComponents), nCols(x.nCols) {
    AllocateData(nComponents);
    for (int i = 0; i < nComponents; i++) {
        Densities[i]->Init(x.Densities[i]->GetCovariance(), x.Densities[i]->GetMean());
        Weights[i] = x.Weights[i];
    }
}

GaussianMixture &GaussianMixture::operator=(const GaussianMixture &x) {
    if (this == &x) return *this;
    if (nComponents != x.nComponents || nCols != x.nCols) {
        FreeData();
        nComponents = x.nComponents;
        nCols = x.nCols;
        AllocateData(nComponents);
    }
    for (int i = 0; i < nComponents; i++) {
        Densities[i]->Init(x.Densities[i]->GetCovariance(), x.Densities[i]->GetMean());
        Weights[i] = x.Weights[i];
    }
    return *this;
}

// Reads the .cmp format (EM2/GaussianMixture.cpp:228-244).
GaussianMixture::GaussianMixture(FILE *pF) : nComponents(0), nCols(0) {
    if (fread(&nComponents, sizeof(int), 1, pF) != 1 || fread(&nCols, sizeof(int), 1, pF) != 1 ||
        nComponents <= 0 || nCols <= 0)
        throw std::runtime_error("Invalid Gaussian mixture file");
    EigenMatrix C(nCols, nCols);
    EigenRowVector M(nCols);
    AllocateData(nComponents);
    for (int i = 0; i < nComponents; i++) {
        if (fread(C.data(), sizeof(double), (size_t)nCols * nCols, pF) != (size_t)nCols * nCols ||
            fread(M.data(), sizeof(double), (size_t)nCols, pF) != (size_t)nCols ||
            fread(&Weights[i], sizeof(double), 1, pF) != 1)
            throw std::runtime_error("Truncated Gaussian mixture file");
        Densities[i]->Init(C, M);
    }
    double last;
    (void)!fread(&last, sizeof(double), 1, pF);
}

void GaussianMixture::Init(const DynamicArray<OPTFLOAT> &aWeights, const DynamicArray<EigenMatrix> &aCovs,
                           const EigenMatrix &aMeans) {
    Weights = aWeights;
    for (int i = 0; i < nComponents; i++) Densities[i]->Init(aCovs[i], EigenRowVector(aMeans.row(i), nCols));
}

void GaussianMixture::InitComponent(int Component, OPTFLOAT Weight, const EigenMatrix &Cov,
                                    const EigenRowVector &Mean) {
    Weights[Component] = Weight;
    Densities[Component]->Init(Cov, Mean);
}

void GaussianMixture::SplitComponent(int Idx, OPTFLOAT W1, const EigenRowVector &M1, const EigenMatrix &C1,
                                     OPTFLOAT W2, const EigenRowVector &M2, const EigenMatrix &C2) {
    Weights.Add(0.0);
    Densities.Add(new NormalDensity(nCols));
    nComponents++;
    InitComponent(Idx, W1, C1, M1);
    InitComponent(nComponents - 1, W2, C2, M2);
}

void GaussianMixture::MergeComponents(int Idx1, int Idx2, OPTFLOAT W, const EigenRowVector &M,
                                      const EigenMatrix &C) {
    InitComponent(Idx1, W, C, M);
    delete Densities[Idx2];
    Densities.RemoveAt(Idx2);
    Weights.RemoveAt(Idx2);
    nComponents--;
}

void GaussianMixture::NormalizeWeights() {
    OPTFLOAT Sum = 0.0;
    for (int i = 0; i < nComponents; i++) Sum += Weights[i];
    for (int i = 0; i < nComponents; i++) Weights[i] *= (1 / Sum);
}

OPTFLOAT GaussianMixture::Evaluate(const float *row) const {
    OPTFLOAT Sum = 0;
    for (int i = 0; i < nComponents; i++) Sum += Densities[i]->GetDensity(row) * Weights[i];
    return Sum;
}

void GaussianMixture::ComputePosteriors(const float *row, EigenRowVector &P) const {
    OPTFLOAT Sum = 0;
    for (int i = 0; i < nComponents; i++) {
        P[i] = Densities[i]->GetDensity(row) * Weights[i];
        Sum += P[i];
    }
    for (int i = 0; i < nComponents; i++) P[i] *= (1 / Sum);
}

int GaussianMixture::Classify(const float *row) {
    int Cls = 0;
    OPTFLOAT MaxB = Densities[0]->GetDensity(row) * Weights[0];
    for (int j = 1; j < nComponents; j++) {
        const OPTFLOAT B = Weights[j] * Densities[j]->GetDensity(row);
        if (B > MaxB) {
            MaxB = B;
            Cls = j;
        }
    }
    return Cls;
}

void GaussianMixture::ClassifyData(const MultithreadedDataset &D, DynamicArray<int> &ClassNums) {
    const long n = D.GetRowCount();
    ClassNums.SetSize(n);
    for (long i = 0; i < n; i++) ClassNums[i] = Classify(D.GetRowNew(i));
}

void GaussianMixture::WriteClasses(const char *fname, const MultithreadedDataset &D) {
    const long n = D.GetRowCount();
    DynamicArray<int> ClassNums(n);
    ClassifyData(D, ClassNums);
    FILE *pF = fopen(fname, "wb");
    if (!pF) throw std::runtime_error("Cannot open class file");
    int nr = (int)n;
    fwrite(&nr, sizeof(int), 1, pF);
    for (long i = 0; i < n; i++) {
        int c = ClassNums[i] + 1;
        fwrite(&c, sizeof(int), 1, pF);
    }
    fclose(pF);
}

OPTFLOAT GaussianMixture::FindLargestConditionNumber() const {
    OPTFLOAT m = 1.0;
    for (int i = 0; i < nComponents; i++) {
        const OPTFLOAT c = Densities[i]->ComputeConditionNumber();
        if (c > m) m = c;
    }
    return m;
}

OPTFLOAT GaussianMixture::ComponentBhattDist(BhattDist &Dist, int Comp, const GaussianMixture &G, int CG) const {
    return Dist.ComputeDist(*Densities[Comp], *G.Densities[CG]);
}

void GaussianMixture::Dump() const {
    printf("Number of components: %d\n", nComponents);
    for (int i = 0; i < nComponents; i++) {
        printf("Component %d: \n", i);
        printf("Mixing probability: %g\n", Weights[i]);
        EigenRowVector M = Densities[i]->GetMean();
        EigenMatrix C = Densities[i]->GetCovariance();
        printf("Mean vector: ");
        for (int j = 0; j < nCols; j++) printf("%g ", M[j]);
        printf("\nCovariance matrix\n");
        for (int r = 0; r < nCols; r++) {
            for (int c = 0; c < nCols; c++) printf("%g ", C(r, c));
            printf("\n");
        }
    }
    printf("Condition number of the solution: %g\n", FindLargestConditionNumber());
}

void GaussianMixture::Write(FILE *pF) {
    const double LastLogL = 0.0;
    fwrite(&nComponents, sizeof(int), 1, pF);
    fwrite(&nCols, sizeof(int), 1, pF);
    for (int i = 0; i < nComponents; i++) {
        EigenMatrix C = Densities[i]->GetCovariance();
        EigenRowVector M = Densities[i]->GetMean();
        fwrite(C.data(), sizeof(double), (size_t)nCols * nCols, pF);
        fwrite(M.data(), sizeof(double), (size_t)nCols, pF);
        fwrite(&Weights[i], sizeof(double), 1, pF);
    }
    fwrite(&LastLogL, sizeof(double), 1, pF);
}

void GaussianMixture::Write(const char *fname) {
    FILE *pF = fopen(fname, "wb");
    if (!pF) throw std::runtime_error("Cannot open mixture file for writing");
    Write(pF);
    fclose(pF);
}
