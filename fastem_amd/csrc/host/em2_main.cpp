// em2_main.cpp -- the em2 driver (EM2/main.cpp:54-159, EM2/EMArgs.cpp:17-184)
// with the same getopt string, defaults, validation messages and output lines.
// The EM path runs on the MI355X engine.  Extensions: -o writes the final
// mixture as .cmp (the reference parses -o but never writes); FASTEM_PRECISION
// selects fp64 | fp32 | bf16x3 contractions.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <stdexcept>

#include "../../../include/fastem/BlockManager.h"
#include "../../../include/fastem/EMAlgorithm.h"
#include "../../../include/fastem/EMInitializer.h"
#include "../../../include/fastem/Rand.h"

static char *fname = nullptr, *oname = nullptr, *cname = nullptr, *iname = nullptr, *aname = nullptr;
static char *rname = nullptr, *mname = nullptr, *affname = nullptr, *msteploopparam = nullptr;
static char *numaname = nullptr, *numaparam = nullptr, *init = nullptr;
static int seed = 0, ncl = 1, verbosity = 1, maxiter = 1000000, burnin = 0;
static double regcoeff = 0.0, Eps = 1e-5, athr = 1e5;
static bool svd = true, benchmarkreducer = false;

static void ExitUsage() {
    printf("Arguments: filename - name of the file in .bin format\n");
    printf("-o filename : writes components in a binary format to the file filename\n");
    printf("-O filename : writes classes in a binary format to the file filename\n");
    printf("-r seed : sets the seed of pseudo-random number generator\n");
    printf("-c components : selects the number of Gaussian mixture components\n");
    printf("-I init : jain | kmeans | kmeansoror | kmeanspp | learned:net.bin | learnedbest:net.bin | file.cmp\n");
    printf("-m emiter : sets the maximal number of EM iterations\n");
    printf("-b burnin : sets the number of burn-in EM iterations (default is 0)\n");
    printf("-f eps : sets the Eps EM stopping threshold (default 1e-5)\n");
    printf("-a abort: set EM failure threshold (default is 1.0e5)\n");
    printf("-A algorithm: matrixem | simplematrixem | cachedem\n");
    printf("-R reducer: naive | log2\n");
    printf("-N msteploopparam: automax:<L3 MB> | <nDens>:<nMStep>\n");
    printf("-g coeff use regularization coefficient for covariance matrices, default 0.0\n");
    printf("-S disables svd\n");
    printf("-v verbosity\n");
}

static int ProcessArgs(int argc, char *argv[]) {
    int c;
    while ((c = getopt(argc, argv, "BN:M:o:O:r:c:t:m:f:a:A:Sv:R:T:I:u:U:b:g:")) != -1) {
        switch (c) {
            case 'u': numaname = optarg; break;
            case 'U': numaparam = optarg; break;
            case 'B': benchmarkreducer = true; break;
            case 'A': aname = optarg; break;
            case 'R': rname = optarg; break;
            case 'T': affname = optarg; break;
            case 'M': mname = optarg; break;
            case 'S': svd = false; break;
            case 'o': oname = optarg; break;
            case 'I': iname = optarg; break;
            case 'O': cname = optarg; break;
            case 'r':
                seed = atoi(optarg);
                if (seed < 0) throw std::invalid_argument("invalid value of -r option\n");
                break;
            case 'b':
                burnin = atoi(optarg);
                if (burnin < 0) throw std::invalid_argument("invalid value of -b option\n");
                break;
            case 'N': msteploopparam = optarg; break;
            case 'c':
                ncl = atoi(optarg);
                if (ncl <= 0) throw std::invalid_argument("%invalid value of -c option\n");
                break;
            case 't': init = optarg; break;
            case 'm':
                maxiter = atoi(optarg);
                if (maxiter < 0) throw std::invalid_argument("invalid value of -m option\n");
                break;
            case 'v':
                verbosity = atoi(optarg);
                if (verbosity < 0) throw std::invalid_argument("invalid value of -v option\n");
                break;
            case 'a':
                athr = atof(optarg);
                if (athr < 1.0) throw std::invalid_argument("argument to -a must be greater then 1.0\n");
                break;
            case 'f':
                Eps = atof(optarg);
                if (Eps <= 0.0) throw std::invalid_argument("argument to -f must be greater or than 0\n");
                break;
            case 'g':
                regcoeff = atof(optarg);
                if (regcoeff < 0.0 || regcoeff > 1.0)
                    throw std::invalid_argument("argument to -g must be small positive number\n");
                break;
            default: throw std::invalid_argument("Invalid command line option");
        }
    }
    if (optind < argc) fname = argv[optind];
    else if (!benchmarkreducer) {
        ExitUsage();
        return -1;
    }
    return 1;
}

static double now() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void RunEM() {
    long nRows;
    int nCols;
    MultithreadedDataset::QueryHeader(fname, nRows, nCols);
    int nDensBlocks, nMStepBlocks;
    BlockManagerBase::ParseEMBlockParams(msteploopparam, nRows, nCols, ncl, nDensBlocks, nMStepBlocks);
    const int nBlocks = std::max(nDensBlocks, nMStepBlocks);
    MultithreadedDataset D(nBlocks);
    printf("Loading dataset %s... ", fname);
    double t0 = now();
    D.Load(fname);
    double extime = now() - t0;
    printf("%ld vectors, %d features\n", D.GetRowCount(), D.GetColCount());
    printf("Loading done in %d blocks\n", nBlocks);
    printf("Loading took %g seconds\n", extime);
    printf("Covariance matrix trace: %10.10g\n", D.GetCovMatrixTrace());
    printf("Number of components: %d\n", ncl);
    printf("Regularization factor: %g\n", regcoeff);
    printf("EM convergence threshold: %g\n", Eps);
    printf("EM abort threshold: %g\n", athr);
    GaussianMixture G(ncl, D.GetColCount());
    EMInitializer *pInit = CreateEMInitializer(iname, D, ncl);
    pInit->PrintInfo();
    pInit->Init(G);
    EMAlgorithm *pA = CreateEMAlgorithm(aname, rname, mname, D.GetColCount(), D.GetRowCount(), D.GetRowCount(),
                                        ncl, msteploopparam);
    printf("Using EM algorithm %s\n", pA->GetName());
    double LogL = 0.0;
    try {
        if (burnin > 0) {
            printf("Running %d burn-in iterations\n", burnin);
            pA->SetParameters(burnin, Eps, athr, regcoeff);
            GaussianMixture G2 = G;
            t0 = now();
            pA->Train(D, G2, verbosity, svd);
            extime = now() - t0;
            printf("Burn-in time: %g seconds\n", extime);
            const int IterCount = pA->GetIterCounter() - 1;
            printf("Burn-in iter time: %g miliseconds\n", 1000.0 * extime / (double)IterCount);
        }
        pA->SetParameters(maxiter, Eps, athr, regcoeff);
        t0 = now();
        if (burnin > 0) LogL = pA->TrainNoCopy(G, verbosity, svd);
        else LogL = pA->Train(D, G, verbosity, svd);
        if (verbosity > 3) G.Dump();
        if (oname) G.Write(oname);
    } catch (EMException &E) {
        printf("EM algorithm failed: %s\n", E.what());
    }
    extime = now() - t0;
    printf("EM Execution time: %g seconds\n", extime);
    const int IterCount = pA->GetIterCounter() - 1;
    printf("%d total EM sweeps\n", IterCount);
    printf("Iteration time: %g miliseconds\n", 1000.0 * extime / (double)IterCount);
    printf("Final loglikelihood: %g\n", LogL);
    if (verbosity > 2) printf("Final loglikelihood (17g): %.17g\n", LogL);
    if (verbosity > 0) pA->PrintNumaLocalityInfo();
    delete pInit;
    delete pA;
}

int main(int argc, char *argv[]) {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    char hostname[1024] = {0};
    gethostname(hostname, 1023);
    printf("Running on machine: %s\n", hostname);
    try {
        if (ProcessArgs(argc, argv) < 0) return 0;
        printf("Seeed of rng: %d\n", seed);
        printf("HIP engine: %s\n", fem_version());
        SRand(seed);
        RunEM();
    } catch (std::invalid_argument &E) {
        printf("Invalid program argument: %s\n", E.what());
    }
    return 0;
}
