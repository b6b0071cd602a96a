// normal_density.cpp -- EM2/NormalDensity.cpp:43-120 semantics on the host.
#include <cmath>
#include <vector>

#include "../../../include/fastem/EMException.h"
#include "../../../include/fastem/NormalDensity.h"

NormalDensity::NormalDensity(int aDim) : Dim(aDim), C(aDim, aDim), CholInvC(aDim, aDim), M(aDim) {}

void NormalDensity::Init(const EigenMatrix &aC, const EigenRowVector &aM) {
    const int d = Dim;
    M = aM;
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) C(i, j) = (i >= j) ? aC(i, j) : aC(j, i);
    std::vector<double> L((size_t)d * d);
    if (!fastem::CholeskyLower(d, C.data(), L.data()))
        throw EMException("Cholesky decomposition failed");
    double s = 0.0;
    for (int i = 0; i < d; i++) s += std::log(L[(size_t)i * d + i]);
    LogDeterminant = 2.0 * s;
    Determinant = std::exp(LogDeterminant);
    LogCoeff = -0.5 * d * std::log(2.0 * M_PI) - 0.5 * LogDeterminant;
    Coeff = std::exp(LogCoeff);
    // CholInvC = (L^-1)^T: column c of L^-1 by forward substitution is row c.
    CholInvC.setZero();
    for (int c = 0; c < d; c++)
        for (int i = c; i < d; i++) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int q = c; q < i; q++) t -= L[(size_t)i * d + q] * CholInvC(c, q);
            CholInvC(c, i) = t / L[(size_t)i * d + i];
        }
}

OPTFLOAT NormalDensity::ComputeConditionNumber() const {
    std::vector<double> ev(Dim);
    fastem::SymmetricEigenvalues(Dim, C.data(), ev.data());
    return std::fabs(ev[Dim - 1] / ev[0]);
}

OPTFLOAT NormalDensity::ComputeTrace() const { return C.trace(); }

// 0.5 * || U^T (x - mu) ||^2 with U = CholInvC (EM2/NormalDensity.cpp:80-92)
OPTFLOAT NormalDensity::ComputeDot(const OPTFLOAT *row) const {
    const int d = Dim;
    std::vector<double> w(d), y(d, 0.0);
    for (int i = 0; i < d; i++) w[i] = row[i] - M[i];
    for (int l = 0; l < d; l++)
        for (int j = l; j < d; j++) y[j] += CholInvC(l, j) * w[l];
    double s = 0.0;
    for (int j = 0; j < d; j++) s += y[j] * y[j];
    return 0.5 * s;
}

OPTFLOAT NormalDensity::MahalanobisDist(const EigenRowVector &row) const { return 2.0 * ComputeDot(row.data()); }

OPTFLOAT NormalDensity::GetDensity(const OPTFLOAT *row) const { return Coeff * std::exp(-ComputeDot(row)); }

OPTFLOAT NormalDensity::GetDensity(const float *row) const {
    std::vector<double> r(row, row + Dim);
    return Coeff * std::exp(-ComputeDot(r.data()));
}

OPTFLOAT NormalDensity::GetLogDensity(const OPTFLOAT *row) const { return LogCoeff - ComputeDot(row); }
