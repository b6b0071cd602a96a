// NormalDensity.h -- one multivariate normal component (EM2/NormalDensity.h:24-57).
// Host FP64; the batched device factorisation used inside EM is fem_factor.
#pragma once
#include "LinAlg.h"

class NormalDensity {
    int Dim;
    EigenMatrix C;
    EigenMatrix CholInvC;
    EigenRowVector M;
    OPTFLOAT Determinant = 0.0;
    OPTFLOAT LogDeterminant = 0.0;
    OPTFLOAT Coeff = 0.0;
    OPTFLOAT LogCoeff = 0.0;

protected:
    OPTFLOAT ComputeDot(const OPTFLOAT *row) const;

public:
    explicit NormalDensity(int aDim);
    virtual ~NormalDensity() {}
    EigenRowVector GetMean() const { return M; }
    EigenMatrix GetCovariance() const { return C; }
    const EigenMatrix &GetInvertedCholCov() const { return CholInvC; }
    OPTFLOAT GetDeterminant() const { return Determinant; }
    OPTFLOAT GetLogDeterminant() const { return LogDeterminant; }
    int GetDimension() const { return Dim; }
    // Declared by the reference (EM2/NormalDensity.h:47) but defined nowhere
    // in it; declared here for source compatibility, with no definition.
    int ScanMatrix();
    // Symmetrises aC from its lower triangle, factorises; throws
    // EMException("Cholesky decomposition failed") (EM2/NormalDensity.cpp:43-65).
    void Init(const EigenMatrix &aC, const EigenRowVector &aM);
    OPTFLOAT GetDensity(const OPTFLOAT *row) const;
    OPTFLOAT GetDensity(const float *row) const;
    OPTFLOAT GetLogDensity(const OPTFLOAT *row) const;
    OPTFLOAT MahalanobisDist(const EigenRowVector &row) const;
    OPTFLOAT ComputeTrace() const;
    OPTFLOAT ComputeConditionNumber() const;
};

// Bhattacharyya distance between two components (EM2/NormalDensity.h:60-69).
class BhattDist {
protected:
    NormalDensity Dens;
    EigenMatrix Sum;

public:
    explicit BhattDist(int aDim);
    OPTFLOAT ComputeDist(const NormalDensity &Dens1, const NormalDensity &Dens2);
};
