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

// |lambda_max / lambda_min| (EM2/NormalDensity.cpp:67-71); throws if the
// QL iteration does not converge.
OPTFLOAT NormalDensity::ComputeConditionNumber() const {
    std::vector<double> ev(Dim);
    if (!fastem::SymmetricEigenvalues(Dim, C.data(), ev.data()))
        throw EMException("Eigenvalue computation of a covariance matrix failed");
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

BhattDist::BhattDist(int aDim) : Dens(aDim), Sum(aDim, aDim) {}

// Bhattacharyya distance between two normal densities
// (EM2/NormalDensity.cpp:128-138): with S = (C1 + C2) / 2,
// D = (m2 - m1)^T S^-1 (m2 - m1) / 8 + ln(det S / sqrt(det C1 det C2)) / 2,
// clamped below at 0.
OPTFLOAT BhattDist::ComputeDist(const NormalDensity &Dens1, const NormalDensity &Dens2) {
    const EigenMatrix C1 = Dens1.GetCovariance(), C2 = Dens2.GetCovariance();
    const int d = Dens1.GetDimension();
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) Sum(i, j) = 0.5 * (C1(i, j) + C2(i, j));
    Dens.Init(Sum, Dens1.GetMean());
    const OPTFLOAT D1 = 0.125 * Dens.MahalanobisDist(Dens2.GetMean());
    const OPTFLOAT D2 =
        0.5 * std::log(Dens.GetDeterminant() / std::sqrt(Dens1.GetDeterminant() * Dens2.GetDeterminant()));
    const OPTFLOAT D = D1 + D2;
    return D < 0.0 ? 0.0 : D;
}
