// EMInitializer.h -- initial mixtures (EM2/EMInitializer.h:16-80).
// CreateEMInitializer names: "jain" (default), "kmeans" (Forgy + Lloyd),
// "kmeanspp" (k-means++ + Lloyd), or a path to a .cmp file.
#pragma once
#include "GaussianMixture.h"
#include "MultithreadedDataset.h"

class EMInitializer {
protected:
    int ncols;
    int Components;
    DynamicArray<EigenMatrix> Covs;
    EigenMatrix Means;
    DynamicArray<OPTFLOAT> Weights;
    MultithreadedDataset &Data;

public:
    EMInitializer(MultithreadedDataset &D, int nC);
    virtual void Init(GaussianMixture &G) = 0;
    virtual void PrintInfo() = 0;
    virtual ~EMInitializer() {}
};

class JainEMInitializer : public EMInitializer {
public:
    JainEMInitializer(MultithreadedDataset &D, int nC) : EMInitializer(D, nC) {}
    void PrintInfo() override;
    void Init(GaussianMixture &G) override;
};

class FileEMInitializer : public EMInitializer {
    FILE *pF = nullptr;

public:
    FileEMInitializer(MultithreadedDataset &D, int nC, const char *fname);
    void PrintInfo() override;
    void Init(GaussianMixture &G) override;
    ~FileEMInitializer() override;
};

// Lloyd k-means from a seeding, then the mixture of the clusters
// (EM2/EMInitializer.cpp:103-134).  Seeding: Forgy random partition
// (Clust/KMeansInitializer.cpp:33-49).
class KMeansEMInitializer : public EMInitializer {
protected:
    virtual void Seed(DynamicArray<OPTFLOAT> &vec);

public:
    KMeansEMInitializer(MultithreadedDataset &D, int nC) : EMInitializer(D, nC) {}
    void PrintInfo() override;
    void Init(GaussianMixture &G) override;
};

// k-means++ seeding (Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84).
class KMeansPPEMInitializer : public KMeansEMInitializer {
protected:
    void Seed(DynamicArray<OPTFLOAT> &vec) override;

public:
    KMeansPPEMInitializer(MultithreadedDataset &D, int nC) : KMeansEMInitializer(D, nC) {}
    void PrintInfo() override;
};

EMInitializer *CreateEMInitializer(const char *name, MultithreadedDataset &Data, int nC);
