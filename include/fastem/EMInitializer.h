// EMInitializer.h -- initial mixtures (EM2/EMInitializer.h:16-80).
// CreateEMInitializer names: "jain" (default), "kmeans" (Forgy + Lloyd),
// "kmeansoror" (k-means|| + Lloyd), "kmeanspp" (k-means++ + Lloyd),
// "learned:<file>", "learnedbest:<file>", or a path to a .cmp file.
#pragma once
#include <string>

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
// (EM2/EMInitializer.cpp:82-134), on the GPU.  Seeding: Forgy
// (Clust/KMeansInitializer.cpp:33-49), as the reference's constructor picks
// (EM2/EMInitializer.cpp:86); subclasses pick another through Seeding().
class KMeansEMInitializer : public EMInitializer {
protected:
    // enum fem_seeding value (include/fastem_abi.h)
    virtual int Seeding() const;

public:
    KMeansEMInitializer(MultithreadedDataset &D, int nC) : EMInitializer(D, nC) {}
    void PrintInfo() override;
    void Init(GaussianMixture &G) override;
};

// k-means++ seeding (Clust/PlusPlusInitializer/PlusPlusInitializer.cpp:33-84);
// no reference EM initializer uses it, the name "kmeanspp" is this build's.
class KMeansPPEMInitializer : public KMeansEMInitializer {
protected:
    int Seeding() const override;

public:
    KMeansPPEMInitializer(MultithreadedDataset &D, int nC) : KMeansEMInitializer(D, nC) {}
    void PrintInfo() override;
};

// k-means|| seeding with oversampling 1.4 (EM2/EMInitializer.cpp:158-166,
// Clust/KMeansOrOrInitializer.cpp:143-174), in one-thread draw order.
class KMeansOrOrEMInitializer : public KMeansEMInitializer {
protected:
    int Seeding() const override;

public:
    KMeansOrOrEMInitializer(MultithreadedDataset &D, int nC) : KMeansEMInitializer(D, nC) {}
    void PrintInfo() override;
};

// The learned initializer (include/fastem_setnet.h), no reference
// counterpart: "learned:<file>" proposes with the set network stored in
// <file> (2048-row subsample; k-means++ if a cluster keeps fewer than two
// points); "learnedbest:<file>" runs EM (eps 1e-5, at most 1000 sweeps, no
// regularisation) from that proposal and from k-means++ and returns the
// converged mixture with the higher final log-likelihood.  Draws come from
// the global Rand() stream.
class LearnedEMInitializer : public EMInitializer {
    std::string File;
    bool Judged;

public:
    static constexpr int SubsampleRows = 2048;
    LearnedEMInitializer(MultithreadedDataset &D, int nC, const char *file, bool judged);
    void PrintInfo() override;
    void Init(GaussianMixture &G) override;
};

EMInitializer *CreateEMInitializer(const char *name, MultithreadedDataset &Data, int nC);
