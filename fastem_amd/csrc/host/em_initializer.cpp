// em_initializer.cpp -- EM2/EMInitializer.cpp:20-177 semantics.
#include <cstdio>
#include <cstring>
#include <string>
#include <stdexcept>
#include <vector>

#include "../../../include/fastem/EMException.h"
#include "../../../include/fastem/EMInitializer.h"
#include "../../../include/fastem/Rand.h"
#include "../../../include/fastem_abi.h"
#include "../../../include/fastem_setnet.h"

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

LearnedEMInitializer::LearnedEMInitializer(MultithreadedDataset &D, int nC, const char *file, bool judged)
    : EMInitializer(D, nC), File(file), Judged(judged) {
    FILE *f = std::fopen(file, "rb");
    if (f == nullptr) throw std::invalid_argument("Cannot open file with the set network");
    std::fclose(f);
}

void LearnedEMInitializer::PrintInfo() {
    printf("Learned EM initializer%s (set network %s)\n", Judged ? ", judged against k-means++" : "", File.c_str());
}

void LearnedEMInitializer::Init(GaussianMixture &G) {
    std::vector<double> w(Components), mu((size_t)Components * ncols), cov((size_t)Components * ncols * ncols);
    fem_setnet *net = nullptr;
    if (fem_setnet_load(&net, File.c_str(), 0) != FEM_OK)
        throw std::invalid_argument(std::string("learned initializer: ") + fem_last_error(nullptr));
    fem_ctx *c = nullptr;
    if (fem_create(&c, Data.GetRowCount(), ncols, Components, FEM_PREC_FP64, 0) != FEM_OK) {
        fem_setnet_destroy(net);
        throw std::runtime_error(std::string("learned initializer: ") + fem_last_error(nullptr));
    }
    int used = 0, rc = fem_upload_data(c, Data.GetData(), ncols);
    double ll[2] = {0.0, 0.0};
    if (rc == FEM_OK) {
        if (Judged)
            rc = fem_setnet_judge_rng(net, c, SubsampleRows, global_rand, nullptr, 29, 1000, 1e-5, 0.0, w.data(),
                                      mu.data(), cov.data(), &used, ll);
        else
            rc = fem_setnet_propose_rng(net, c, SubsampleRows, global_rand, nullptr, 29, w.data(), mu.data(),
                                        cov.data(), &used);
    }
    const std::string msg = fem_last_error(c);
    fem_destroy(c);
    fem_setnet_destroy(net);
    if (rc == FEM_ERR_ARG && msg.find("differ") != std::string::npos) throw std::invalid_argument(msg);
    if (rc != FEM_OK && rc != FEM_ERR_ARG) throw std::runtime_error("learned initializer: " + msg);
    if (Judged)
        printf("Learned initializer: final LL %.17g from the network, %.17g from k-means++; kept %s\n", ll[0], ll[1],
               used == 2 ? "k-means++" : "the network's");
    else if (used == 0)
        printf("Learned initializer: a proposed cluster kept fewer than d + 1 points, k-means++ used\n");
    for (int i = 0; i < Components; i++) {
        Weights[i] = w[i];
        for (int j = 0; j < ncols; j++) Means(i, j) = mu[(size_t)i * ncols + j];
        std::memcpy(Covs[i].data(), cov.data() + (size_t)i * ncols * ncols, sizeof(double) * ncols * ncols);
    }
    G.Init(Weights, Covs, Means);
}

EMInitializer *CreateEMInitializer(const char *name, MultithreadedDataset &Data, int nC) {
    if (name == nullptr || !std::strcmp(name, "jain")) return new JainEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeans")) return new KMeansEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeansoror")) return new KMeansOrOrEMInitializer(Data, nC);
    if (!std::strcmp(name, "kmeanspp")) return new KMeansPPEMInitializer(Data, nC);
    if (!std::strncmp(name, "learned:", 8)) return new LearnedEMInitializer(Data, nC, name + 8, false);
    if (!std::strncmp(name, "learnedbest:", 12)) return new LearnedEMInitializer(Data, nC, name + 12, true);
    return new FileEMInitializer(Data, nC, name);
}
