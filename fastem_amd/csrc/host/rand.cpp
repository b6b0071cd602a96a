// rand.cpp -- Util/Rand.cpp:7-14,33-36 semantics.
#include <cmath>
#include <random>

#include "../../../include/fastem/Rand.h"

static std::mt19937_64 g_rng;

double Rand() {
    const double r = (double)g_rng() / 18446744073709551616.0;
    return r < 1.0 ? r : std::nextafter(1.0, 0.0);
}

void SRand(unsigned seed) { g_rng.seed(seed); }

int RandInt(unsigned long n) { return (int)(Rand() * n); }
