// Rand.h -- the process-wide Rand() stream (Util/Rand.h:7-17): std::mt19937_64
// with libstdc++'s generate_canonical<double,53> mapping, so seeds give the
// same draws as the reference.
#pragma once

double Rand();
void SRand(unsigned seed);
int RandInt(unsigned long n);
