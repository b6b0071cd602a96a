// linalg.cpp -- host FP64 dense kernels for the parameter-side classes.
#include <cfloat>
#include <cmath>
#include <vector>

#include "../../../include/fastem/LinAlg.h"

namespace fastem {

bool CholeskyLower(int d, const double *C, double *L) {
    for (int i = 0; i < d * d; i++) L[i] = 0.0;
    for (int j = 0; j < d; j++) {
        double s = C[j * d + j];
        for (int k = 0; k < j; k++) s -= L[j * d + k] * L[j * d + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; i++) {
            double t = C[i * d + j];
            for (int k = 0; k < j; k++) t -= L[i * d + k] * L[j * d + k];
            L[i * d + j] = t / ljj;
        }
    }
    return true;
}

// Householder reduction to a symmetric tridiagonal (diag a, off-diagonal e),
// then QL with implicit shifts on the tridiagonal; eigenvalues only.
bool SymmetricEigenvalues(int d, const double *C, double *ev) {
    std::vector<double> A((size_t)d * d), e((size_t)d + 1, 0.0), v(d), p(d);
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) A[(size_t)i * d + j] = (i >= j) ? C[i * d + j] : C[j * d + i];
    for (int k = 0; k + 2 < d; k++) {
        const int m = d - k - 1;
        double nrm = 0.0;
        for (int i = 0; i < m; i++) {
            const double x = A[(size_t)(k + 1 + i) * d + k];
            nrm += x * x;
        }
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) { e[k] = 0.0; continue; }
        const double x0 = A[(size_t)(k + 1) * d + k];
        const double alpha = (x0 > 0.0) ? -nrm : nrm;
        for (int i = 0; i < m; i++) v[i] = A[(size_t)(k + 1 + i) * d + k];
        v[0] -= alpha;
        double vtv = 0.0;
        for (int i = 0; i < m; i++) vtv += v[i] * v[i];
        e[k] = alpha;
        if (vtv == 0.0) continue;
        const double beta = 2.0 / vtv;
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            const double *Ar = A.data() + (size_t)(k + 1 + i) * d + (k + 1);
            for (int j = 0; j < m; j++) s += Ar[j] * v[j];
            p[i] = beta * s;
        }
        double kk = 0.0;
        for (int i = 0; i < m; i++) kk += v[i] * p[i];
        kk *= 0.5 * beta;
        for (int i = 0; i < m; i++) p[i] -= kk * v[i];
        for (int i = 0; i < m; i++) {
            double *Ar = A.data() + (size_t)(k + 1 + i) * d + (k + 1);
            for (int j = 0; j < m; j++) Ar[j] -= v[i] * p[j] + p[i] * v[j];
        }
    }
    for (int i = 0; i < d; i++) ev[i] = A[(size_t)i * d + i];
    if (d >= 2) e[d - 2] = A[(size_t)(d - 1) * d + (d - 2)];
    e[d - 1] = 0.0;
    for (int l = 0; l < d; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < d - 1; m++) {
                const double dd = std::fabs(ev[m]) + std::fabs(ev[m + 1]);
                if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) return false;
            double g = (ev[l + 1] - ev[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = ev[m] - ev[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, pp = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    ev[i + 1] -= pp;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = ev[i + 1] - pp;
                r = (ev[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                ev[i + 1] = g + pp;
                g = c * r - b;
            }
            if (early) continue;
            ev[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    for (int i = 1; i < d; i++) {
        const double t = ev[i];
        int j = i - 1;
        while (j >= 0 && ev[j] > t) { ev[j + 1] = ev[j]; j--; }
        ev[j + 1] = t;
    }
    return true;
}

}  // namespace fastem
