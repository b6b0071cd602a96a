// em_algorithm.cpp -- EMAlgorithm / MatrixEM / SimpleMatrixEM / CachedEM over
// the device engine.  The loop, convergence rule, IterCounter bookkeeping,
// regularisation before the loop, condition-number abort and exception
// messages follow EM2/EMAlgorithm.cpp:139-191; the heavy hooks call the C ABI.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/fastem/BlockManager.h"
#include "../../../include/fastem/CachedEM.h"
#include "../../../include/fastem/MatrixEM.h"

namespace {

int precision_from_env() {
    const char *p = std::getenv("FASTEM_PRECISION");
    if (!p || !std::strcmp(p, "fp64")) return FEM_PREC_FP64;
    if (!std::strcmp(p, "fp32")) return FEM_PREC_FP32;
    if (!std::strcmp(p, "bf16x3")) return FEM_PREC_BF16X3;
    throw std::invalid_argument("FASTEM_PRECISION must be fp64, fp32 or bf16x3");
}

int &default_precision() {
    static int p = precision_from_env();
    return p;
}

double monotonic_seconds() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

}  // namespace

void EMAlgorithm::SetDefaultPrecision(int prec) { default_precision() = prec; }
int EMAlgorithm::GetDefaultPrecision() { return default_precision(); }

EMAlgorithm::EMAlgorithm(int ncols, long nrows, int ncl)
    : nCols(ncols), K(ncl), IterCounter(0), TotalRowCount(nrows), nRows(nrows), MaxSweeps(1000000),
      Eps(1e-5), AbortThr(100000.0), RegCoef(0.0), Means(ncl, ncols), Probs(ncl), PLLs(ncl), LogLike(0.0),
      Ctx(nullptr), Precision(default_precision()), HostParamsNewer(false),
      DeviceParamsNewer(false), DeviceRegularized(false) {
    Covs.SetSize(K);
    for (int i = 0; i < K; i++) Covs[i].resize(nCols, nCols);
    int rc = fem_create(&Ctx, nrows, ncols, ncl, Precision, 0);
    if (rc != FEM_OK) throw std::runtime_error(std::string("fastem device context: ") + fem_last_error(nullptr));
}

EMAlgorithm::~EMAlgorithm() { fem_destroy(Ctx); }

void EMAlgorithm::Check(int rc, const char *what) const {
    if (rc == FEM_OK) return;
    throw std::runtime_error(std::string(what) + ": " + fem_last_error(Ctx));
}

void EMAlgorithm::SetParameters(int aMaxSweeps, double aEps, double aAbortThr, double aRegCoef) {
    MaxSweeps = aMaxSweeps;
    Eps = aEps;
    AbortThr = aAbortThr;
    RegCoef = aRegCoef;
}

void EMAlgorithm::CopyDataset(const MultithreadedDataset &Data) {
    if (Data.GetRowCount() != nRows || Data.GetColCount() != nCols)
        throw std::invalid_argument("dataset shape differs from the EM algorithm's");
    Check(fem_upload_data(Ctx, Data.GetData(), nCols), "fem_upload_data");
}

void EMAlgorithm::DecodeMixture(const GaussianMixture &G) {
    for (int i = 0; i < K; i++) {
        Probs[i] = G.GetWeight(i);
        EigenRowVector m = G.GetMean(i);
        for (int j = 0; j < nCols; j++) Means(i, j) = m[j];
        Covs[i] = G.GetCovariance(i);
    }
    HostParamsNewer = true;
    DeviceParamsNewer = false;
}

// EM2/EMAlgorithm.cpp:237-244.  After a device M-step the regularisation has
// already been applied by the engine with the same formula.
void EMAlgorithm::RegularizeCovariances() {
    if (DeviceRegularized) {
        DeviceRegularized = false;
        return;
    }
    for (int j = 0; j < K; j++) {
        const double Trace = Covs[j].trace();
        const double diag = RegCoef * Trace / (double)nCols;
        for (int a = 0; a < nCols; a++)
            for (int b = 0; b < nCols; b++) Covs[j](a, b) = Covs[j](a, b) * (1.0 - RegCoef) + (a == b ? diag : 0.0);
    }
    HostParamsNewer = true;
}

void EMAlgorithm::FetchParams() {
    if (!DeviceParamsNewer) return;
    std::vector<double> cov((size_t)K * nCols * nCols);
    Check(fem_get_params(Ctx, Probs.data(), Means.data(), cov.data()), "fem_get_params");
    for (int k = 0; k < K; k++)
        std::memcpy(Covs[k].data(), cov.data() + (size_t)k * nCols * nCols, sizeof(double) * nCols * nCols);
    DeviceParamsNewer = false;
}

// EMAlgorithm::InitDensities (EM2/EMAlgorithm.cpp:75-87) as one batched device
// factorisation.
void EMAlgorithm::InitDensities() {
    if (HostParamsNewer) {
        std::vector<double> cov((size_t)K * nCols * nCols);
        for (int k = 0; k < K; k++)
            std::memcpy(cov.data() + (size_t)k * nCols * nCols, Covs[k].data(), sizeof(double) * nCols * nCols);
        Check(fem_set_params(Ctx, Probs.data(), Means.data(), cov.data()), "fem_set_params");
        HostParamsNewer = false;
    }
    const int rc = fem_factor(Ctx, nullptr, nullptr);
    if (rc == FEM_ERR_NOT_PD) throw EMException("NormalDensity::Init failed inside a parallel region");
    Check(rc, "fem_factor");
}

void EMAlgorithm::PrecomputeDensityMatrix() {
    double ll = 0.0;
    Check(fem_estep(Ctx, &ll), "fem_estep");
    LogLike = ll;
    MPIReduceLogL(LogLike);
}

void EMAlgorithm::MStep() {
    Check(fem_mstep(Ctx, nullptr, nullptr, nullptr), "fem_mstep");
    DeviceParamsNewer = true;
    DeviceRegularized = true;
}

OPTFLOAT EMAlgorithm::FindLargestConditionNumber() {
    FetchParams();
    return LargestHostConditionNumber();
}

// The largest condition number of the host copies of the covariances (as
// fetched by FetchParams or FinishStagedFetch); throws when an eigenvalue iteration fails.
OPTFLOAT EMAlgorithm::LargestHostConditionNumber() const {
    std::vector<double> conds(K, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < K; i++) {
        std::vector<double> ev(nCols);
        // a QL iteration that does not converge leaves no usable eigenvalues
        conds[i] = fastem::SymmetricEigenvalues(nCols, Covs[i].data(), ev.data())
                       ? std::fabs(ev[nCols - 1] / ev[0])
                       : std::numeric_limits<double>::quiet_NaN();
    }
    double m = 0.0;
    for (int i = 0; i < K; i++) {
        if (std::isnan(conds[i])) throw EMException("Eigenvalue computation of a covariance matrix failed");
        if (conds[i] > m) m = conds[i];
    }
    return m;
}

void EMAlgorithm::FinishStagedFetch(double *cov) {
    Check(fem_get_params_finish(Ctx, Probs.data(), Means.data(), cov), "fem_get_params_finish");
    for (int k = 0; k < K; k++)
        std::memcpy(Covs[k].data(), cov + (size_t)k * nCols * nCols, sizeof(double) * nCols * nCols);
}

OPTFLOAT EMAlgorithm::FindMinTrace() {
    FetchParams();
    double m = std::numeric_limits<double>::max();
    for (int i = 0; i < K; i++) {
        const double t = Covs[i].trace();
        if (t < m) m = t;
    }
    return m;
}

OPTFLOAT EMAlgorithm::Train(const MultithreadedDataset &Data, GaussianMixture &G, int Verbosity, bool SVD) {
    CopyDataset(Data);
    TotalRowCount = Data.GetTotalRowCount();
    return TrainNoCopy(G, Verbosity, SVD);
}

OPTFLOAT EMAlgorithm::TrainNoCopy(GaussianMixture &G, int Verbosity, bool SVD) {
    Check(fem_set_total_rows(Ctx, TotalRowCount), "fem_set_total_rows");
    Check(fem_set_regularization(Ctx, RegCoef), "fem_set_regularization");
    double t0 = monotonic_seconds();
    DecodeMixture(G);
    OPTFLOAT PrevLogL = -std::numeric_limits<OPTFLOAT>::max(), LogL = PrevLogL;
    OPTFLOAT MaxCond = 0.0;
    DeviceRegularized = false;
    RegularizeCovariances();
    InitDensities();
    if (SVD) {
        MaxCond = FindLargestConditionNumber();
        if (MaxCond > AbortThr) throw EMException("Condition number of a covariance matrix is too high");
    }
    // With SVD the condition numbers of an iteration's new covariances are
    // computed on the host while the device runs the next iteration's
    // E-step; the iteration's check, abort and line follow once both are
    // done.  The output is the reference's: an abort still comes before the
    // iteration's line, and the next E-step's result is then never used.
    // one host thread for the whole run (its OpenMP team persists between
    // iterations, where a thread per iteration started a new team each time)
    // The worker also brings the parameters to the host (queued by
    // fem_get_params_start after the iteration's factorisation) and scans the
    // traces, so the device goes on to the next E-step without a host wait.
    struct CondWorker {
        EMAlgorithm *em;
        std::vector<double> staged;
        std::mutex m;
        std::condition_variable cv;
        bool work = false, done = false, stop = false;
        OPTFLOAT result = 0.0, trace = 0.0;
        std::exception_ptr err;
        std::thread th;
        explicit CondWorker(EMAlgorithm *e) : em(e), staged((size_t)e->K * e->nCols * e->nCols) {
            th = std::thread([this]() {
                std::unique_lock<std::mutex> lk(m);
                for (;;) {
                    cv.wait(lk, [this]() { return work || stop; });
                    if (stop) return;
                    lk.unlock();
                    OPTFLOAT r = 0.0, t = std::numeric_limits<double>::max();
                    std::exception_ptr ex;
                    try {
                        em->FinishStagedFetch(staged.data());
                        for (int i = 0; i < em->K; i++) t = std::min<OPTFLOAT>(t, em->Covs[i].trace());
                        r = em->LargestHostConditionNumber();
                    } catch (...) {
                        ex = std::current_exception();
                    }
                    lk.lock();
                    result = r;
                    trace = t;
                    err = ex;
                    work = false;
                    done = true;
                    cv.notify_all();
                }
            });
        }
        void start() {
            std::lock_guard<std::mutex> lk(m);
            work = true;
            done = false;
            cv.notify_all();
        }
        OPTFLOAT get(OPTFLOAT *min_trace) {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [this]() { return done; });
            done = false;
            if (err) std::rethrow_exception(err);
            *min_trace = trace;
            return result;
        }
        ~CondWorker() {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [this]() { return !work; });
                stop = true;
                cv.notify_all();
            }
            th.join();
        }
    };
    std::unique_ptr<CondWorker> cond;
    if (SVD) cond = std::make_unique<CondWorker>(this);
    struct Line {
        int iter;
        OPTFLOAT LogL, e, MinTrace;
        double ms;
    } pending{};
    bool have_pending = false;
    auto finish = [&]() {  // the check and line of the pending iteration
        MaxCond = cond->get(&pending.MinTrace);
        if (MaxCond > AbortThr) throw EMException("Condition number of a covariance matrix is too high");
        if (Verbosity > 1)
            printf("iter: %d LogL: %10.6f e: %g Cond: %g Trace: %g Time[ms]: %g\n", pending.iter, pending.LogL,
                   pending.e, MaxCond, pending.MinTrace, pending.ms);
        if (Verbosity > 2) printf("iter: %d LogL(17g): %.17g\n", pending.iter, pending.LogL);
        have_pending = false;
    };
    for (IterCounter = 1; IterCounter <= MaxSweeps; IterCounter++) {
        PrecomputeDensityMatrix();
        if (have_pending) finish();
        LogL = ComputeLogLike();
        EStep();
        MStep();
        RegularizeCovariances();
        InitDensities();
        OPTFLOAT e = std::fabs((LogL - PrevLogL) / LogL);
        if (LogL < PrevLogL) e = 0.0;
        const double t1 = monotonic_seconds();
        if (SVD) {
            Check(fem_get_params_start(Ctx), "fem_get_params_start");
            DeviceParamsNewer = false;  // on the host once the worker has run
            pending = {IterCounter, LogL, e, 0.0, 1000.0 * (t1 - t0)};
            have_pending = true;
            cond->start();
            if (e < Eps || IterCounter == MaxSweeps) finish();  // no further E-step to overlap
        } else {
            if (Verbosity > 1)
                printf("iter: %d LogL: %10.6f e: %g Time[ms]: %g\n", IterCounter, LogL, e, 1000.0 * (t1 - t0));
            if (Verbosity > 2) printf("iter: %d LogL(17g): %.17g\n", IterCounter, LogL);
        }
        t0 = t1;
        PrevLogL = LogL;
        if (e < Eps) {
            IterCounter++;
            break;
        }
    }
    FetchParams();
    DynamicArray<OPTFLOAT> Weights(K);
    for (int i = 0; i < K; i++) Weights[i] = Probs[i];
    G.Init(Weights, Covs, Means);
    return LogL;
}

void EMAlgorithm::RetrieveClasses(LargeVector<int> &Classes) const {
    Classes.SetSize(nRows);
    Check(fem_get_classes(Ctx, Classes.GetData()), "fem_get_classes");
}

void EMAlgorithm::ComputePosterriorCorrelation(EigenMatrix &Corr) {
    std::vector<double> R((size_t)nRows * K);
    Check(fem_get_posteriors(Ctx, R.data(), K), "fem_get_posteriors");
    Corr.resize(K, K);
    for (long i = 0; i < nRows; i++)
        for (int a = 0; a < K; a++) {
            const double ra = R[(size_t)i * K + a];
            for (int b = 0; b < K; b++) Corr(a, b) += ra * R[(size_t)i * K + b];
        }
}

void EMAlgorithm::PrintNumaLocalityInfo() {
    printf("NUMA locality of posterior matrix : not applicable (device HBM)\n");
}

static void check_reducer(const char *ReducerName) {
    if (ReducerName == nullptr || !std::strcmp(ReducerName, "naive") || !std::strcmp(ReducerName, "log2")) return;
    throw std::invalid_argument("Unknown OpenMP reducer name");
}

MatrixEMBase::MatrixEMBase(int ncols, long nrows, int ncl, const char *ReducerName, int anDensBlocks,
                           int anMStepBlocks)
    : EMAlgorithm(ncols, nrows, ncl), nDensBlocks(anDensBlocks), nMStepBlocks(anMStepBlocks) {
    check_reducer(ReducerName);
}

CachedEM::CachedEM(int ncols, long nrows, int ncl, const char *ReducerName, int, int)
    : EMAlgorithm(ncols, nrows, ncl) {
    check_reducer(ReducerName);
}

EMAlgorithm *CreateEMAlgorithm(const char *Name, const char *RName, const char *, int nCols, long nRows,
                               long nTotalRows, int nCl, const char *MStepLoopParam) {
    EMAlgorithm *p = nullptr;
    const bool known = Name == nullptr || !std::strcmp(Name, "matrixem") ||
                       !std::strcmp(Name, "simplematrixem") || !std::strcmp(Name, "cachedem");
    if (known) check_reducer(RName);  // before any device allocation
    if (Name == nullptr || !std::strcmp(Name, "matrixem") || !std::strcmp(Name, "simplematrixem")) {
        int P1, P2;
        BlockManagerBase::ParseEMBlockParams(MStepLoopParam, nRows, nCols, nCl, P1, P2);
        printf("Blocks in density computation: %d, in Covariance computation: %d\n", P1, P2);
        if (Name == nullptr || !std::strcmp(Name, "matrixem")) p = new MatrixEM(nCols, nRows, nCl, RName, P1, P2);
        else p = new SimpleMatrixEM(nCols, nRows, nCl, RName, P1, P2);
    } else if (!std::strcmp(Name, "cachedem")) {
        int P1, P2;
        BlockManagerBase::ParseEMBlockParams(MStepLoopParam, nRows, nCols, nCl, P1, P2);
        printf("Blocks in density computation: %d, in Covariance computation: %d\n", P1, P2);
        p = new CachedEM(nCols, nRows, nCl, RName, P1, P2);
    } else {
        throw std::invalid_argument("Unknown EM algorithm");
    }
    (void)nTotalRows;  // TotalRowCount comes from Train(), as in the reference
    return p;
}
