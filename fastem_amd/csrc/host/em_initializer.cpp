// em_initializer.cpp -- EM2/EMInitializer.cpp:20-177 semantics.
#include <cstring>
#include <string>
#include <stdexcept>
#include <vector>

#include "../../../include/fastem/EMException.h"
#include "../../../include/fastem/EMInitializer.h"
#include "../../../include/fastem/Rand.h"
#include "../../../include/fastem_abi.h"

EMInitializer::EMInitializer(MultithreadedDataset &D, int nC)
    : ncols(D.GetColCount()), Components(nC), Means(nC, D.GetColCount()), Data(D) {
    Covs.SetSize(Components);
    Weights.SetSize(Components);
    for (int i = 0; i < Components; i++) Covs[i].resize(ncols, ncols);
}

void JainEMInitializer::PrintInfo() { printf("Jain EM Initializer\n"); }

// EM2/EMInitializer.cpp:43-80: K rows drawn with Rand(), weights 1/K,
// covariances tr(Sigma_data)/d/10 * I.  The reference's duplicate check
// re-draws into a shadowed variable, so the first draw is always kept; that
// behaviour (and its Rand() consumption) is reproduced.
void JainEMInitializer::Init(GaussianMixture &G) {
    const OPTFLOAT Weight = 1.0 / (OPTFLOAT)Components;
    const OPTFLOAT Cov = Data.GetCovMatrixTrace() / (OPTFLOAT)ncols / 10.0;
    std::vector<long> ObjNums(Components);
    for (int i = 0; i < Components; i++) {
        const long Idx = (long)(Rand() * Data.GetRowCount());
        bool found;
        do {
            const long Idx2 = (long)(Rand() * Data.GetRowCount());
            found = false;
            for (int j = 0; j < i; j++)
                if (ObjNums[j] == Idx2) {
                    found = true;
                    break;
                }
        } while (found);
        ObjNums[i] = Idx;
        const float *row = Data.GetRowNew(Idx);
        Weights[i] = Weight;
        for (int j = 0; j < ncols; j++) Means(i, j) = row[j];
        Covs[i] = EigenMatrix::Identity(ncols);
        for (int j = 0; j < ncols; j++) Covs[i](j, j) = Cov;
    }
    G.Init(Weights, Covs, Means);
}

FileEMInitializer::FileEMInitializer(MultithreadedDataset &D, int nC, const char *fname)
    : EMInitializer(D, nC) {
    pF = fopen(fname, "rb");
    if (pF == nullptr) throw std::invalid_argument("Cannot open file with Gaussian mixture");
}

FileEMInitializer::~FileEMInitializer() {
    if (pF) fclose(pF);
}

void FileEMInitializer::PrintInfo() { printf("File EM initializer\n"); }

void FileEMInitializer::Init(GaussianMixture &G) {
    GaussianMixture G1(pF);
    G = G1;
}

static double global_rand(void *) { return Rand(); }

int KMeansEMInitializer::Seeding() const { return FEM_SEED_FORGY; }

void KMeansEMInitializer::PrintInfo() { printf("K-means EM initializer\n"); }

// The seeding, NaiveKMA::RunKMeans(vec, 0, 1e-4, 30), classification and the
// per-cluster mixture (EM2/EMInitializer.cpp:103-134), on the global Rand()
// stream.
void KMeansEMInitializer::Init(GaussianMixture &G) {
    const int seeding = Seeding();
    std::vector<double> w(Components), mu((size_t)Components * ncols),
        cov((size_t)Components * ncols * ncols);
    fem_ctx *c = nullptr;
    if (fem_create(&c, Data.GetRowCount(), ncols, Components, FEM_PREC_FP64, 0) != FEM_OK)
        throw std::runtime_error(std::string("k-means initializer: ") + fem_last_error(nullptr));
    int rc = fem_upload_data(c, Data.GetData(), ncols);
    if (rc == FEM_OK)
        rc = fem_kmeans_init_rng(c, global_rand, nullptr, seeding, 29, w.data(), mu.data(), cov.data());
    const std::string msg = fem_last_error(c);
    fem_destroy(c);
    if (rc != FEM_OK && rc != FEM_ERR_ARG) throw std::runtime_error("k-means initializer: " + msg);
    for (int i = 0; i < Components; i++) {
        Weights[i] = w[i];
        for (int j = 0; j < ncols; j++) Means(i, j) = mu[(size_t)i * ncols + j];
        std::memcpy(Covs[i].data(), cov.data() + (size_t)i * ncols * ncols, sizeof(double) * ncols * ncols);
    }
    G.Init(Weights, Covs, Means);
}

int KMeansPPEMInitializer::Seeding() const { return FEM_SEED_KMEANSPP; }

void KMeansPPEMInitializer::PrintInfo() { printf("K-means++ EM initializer\n"); }

int KMeansOrOrEMInitializer::Seeding() const { return FEM_SEED_KMEANSOROR; }

void KMeansOrOrEMInitializer::PrintInfo() { printf("K-means|| EM initializer\n"); }

EMInitializer *CreateEMInitializer(const char *name, MultithreadedDataset &Data, int nC) {
    if (name == nullptr || !std::strcmp(name, "jain")) return new JainEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeans")) return new KMeansEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeansoror")) return new KMeansOrOrEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeanspp")) return new KMeansPPEMInitializer(Data, nC);
    return new FileEMInitializer(Data, nC, name);
}
