// eminit_main.cpp -- the eminit tool (EM2/eminit.cpp:82-113): loads a .bin
// dataset, runs an initializer and writes the mixture as .cmp (-o) or dumps
// it.  Same getopt string "o:I:r:c:", default initializer "kmeans"; the k-means
// initializers run on the GPU.  Adds -I kmeanspp.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "../../../include/fastem/EMInitializer.h"
#include "../../../include/fastem/Rand.h"

static int seed = 0, ncl = 1;
static char *iname = (char *)"kmeans";
static char *oname = nullptr, *fname = nullptr;

static void ProcessArgs(int argc, char *argv[]) {
    int c;
    while ((c = getopt(argc, argv, "o:I:r:c:")) != -1) {
        switch (c) {
            case 'o': oname = optarg; break;
            case 'I': iname = optarg; break;
            case 'r':
                seed = atoi(optarg);
                if (seed < 0) throw std::invalid_argument("invalid value of -r option\n");
                break;
            case 'c':
                ncl = atoi(optarg);
                if (ncl <= 0) throw std::invalid_argument("%invalid value of -c option\n");
                break;
            default: throw std::invalid_argument("Invalid command line option");
        }
    }
    if (optind < argc) fname = argv[optind];
    else throw std::invalid_argument("Missing dataset name");
}

int main(int argc, char *argv[]) {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    char hostname[1024] = {0};
    gethostname(hostname, 1023);
    printf("Running on machine: %s\n", hostname);
    try {
        ProcessArgs(argc, argv);
        printf("Seeed of rng: %d\n", seed);
        SRand(seed);
        MultithreadedDataset D;
        printf("Loading dataset %s... ", fname);
        D.Load(fname);
        printf("%ld vectors, %d features\n", D.GetRowCount(), D.GetColCount());
        printf("Covariance matrix trace: %10.10g\n", D.GetCovMatrixTrace());
        printf("Number of components: %d\n", ncl);
        GaussianMixture G(ncl, D.GetColCount());
        EMInitializer *pInit = CreateEMInitializer(iname, D, ncl);
        pInit->PrintInfo();
        pInit->Init(G);
        if (oname == nullptr) G.Dump();
        else G.Write(oname);
        delete pInit;
    } catch (std::invalid_argument &E) {
        printf("Invalid program argument: %s\n", E.what());
    }
    return 0;
}
