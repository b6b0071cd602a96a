// LinAlg.h -- small dense FP64 containers and host linear algebra used by the
// parameter-side classes (NormalDensity, GaussianMixture).  Row-major, like
// the reference's EigenMatrix typedefs (Util/EigenUtil.h:17-20), so the
// reference's calling code that indexes M(i,j), M.row(i) and M.data() keeps
// working.
#pragma once
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace fastem {

class Matrix {
    int r_ = 0, c_ = 0;
    std::vector<double> v_;

public:
    Matrix() = default;
    Matrix(int r, int c) : r_(r), c_(c), v_((size_t)r * c, 0.0) {}
    void resize(int r, int c) {
        r_ = r;
        c_ = c;
        v_.assign((size_t)r * c, 0.0);
    }
    int rows() const { return r_; }
    int cols() const { return c_; }
    double &operator()(int i, int j) { return v_[(size_t)i * c_ + j]; }
    double operator()(int i, int j) const { return v_[(size_t)i * c_ + j]; }
    double *data() { return v_.data(); }
    const double *data() const { return v_.data(); }
    double *row(int i) { return v_.data() + (size_t)i * c_; }
    const double *row(int i) const { return v_.data() + (size_t)i * c_; }
    void setZero() { std::fill(v_.begin(), v_.end(), 0.0); }
    double trace() const {
        double s = 0.0;
        for (int i = 0; i < r_ && i < c_; i++) s += (*this)(i, i);
        return s;
    }
    static Matrix Identity(int n) {
        Matrix m(n, n);
        for (int i = 0; i < n; i++) m(i, i) = 1.0;
        return m;
    }
};

class RowVector {
    std::vector<double> v_;

public:
    RowVector() = default;
    explicit RowVector(int n) : v_((size_t)n, 0.0) {}
    RowVector(const double *p, int n) : v_(p, p + n) {}
    void resize(int n) { v_.assign((size_t)n, 0.0); }
    int size() const { return (int)v_.size(); }
    double &operator[](int i) { return v_[(size_t)i]; }
    double operator[](int i) const { return v_[(size_t)i]; }
    double &operator()(int i) { return v_[(size_t)i]; }
    double operator()(int i) const { return v_[(size_t)i]; }
    double *data() { return v_.data(); }
    const double *data() const { return v_.data(); }
};

// Growable array with the reference's DynamicArray interface (Util/Array.h).
template <typename T>
class DynamicArray {
    std::vector<T> v_;

public:
    DynamicArray() = default;
    explicit DynamicArray(long n) : v_((size_t)n) {}
    void SetSize(long n) { v_.resize((size_t)n); }
    long GetSize() const { return (long)v_.size(); }
    T &operator[](long i) { return v_[(size_t)i]; }
    const T &operator[](long i) const { return v_[(size_t)i]; }
    T *GetData() { return v_.data(); }
    const T *GetData() const { return v_.data(); }
    void Add(const T &x) { v_.push_back(x); }
    void RemoveAt(long i) { v_.erase(v_.begin() + i); }
};

// Zero-initialised vector with the reference's LargeVector interface
// (Util/LargeVector.h:15-95): SetSize discards the old contents and zeroes.
template <typename T>
class LargeVector {
    std::vector<T> v_;

public:
    LargeVector() = default;
    void SetSize(long n) { v_.assign((size_t)n, T(0)); }
    long GetSize() const { return (long)v_.size(); }
    T &operator[](long i) { return v_[(size_t)i]; }
    const T &operator[](long i) const { return v_[(size_t)i]; }
    T *GetData() { return v_.data(); }
    const T *GetData() const { return v_.data(); }
};

// Cholesky-Crout of the lower triangle of C (d x d): L lower.  False if a
// pivot is not positive.
bool CholeskyLower(int d, const double *C, double *L);
// Eigenvalues (ascending) of the lower-symmetrised C: Householder
// tridiagonalisation and implicit QL.  False if QL does not converge.
bool SymmetricEigenvalues(int d, const double *C, double *ev);

}  // namespace fastem

typedef fastem::Matrix EigenMatrix;
typedef fastem::RowVector EigenRowVector;
typedef fastem::RowVector EigenVector;
typedef double OPTFLOAT;
using fastem::DynamicArray;
using fastem::LargeVector;
