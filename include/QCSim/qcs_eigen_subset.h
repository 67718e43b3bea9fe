// qcs_eigen_subset.h - the Eigen types the QCSim statevector API names.
//
// The reference declares its classes with Eigen::VectorXcd / Eigen::MatrixXcd
// defaults (QubitRegister.h:11, SimpleGates.h:12). When Eigen is installed it
// is used as is; otherwise this header provides the subset of those two types
// that the register, gate and subalgorithm interfaces touch, so code written
// against the reference compiles unchanged.
#pragma once

#if __has_include(<Eigen/Dense>) && !defined(QCS_NO_EIGEN)
#include <Eigen/Dense>
#else

#include <cassert>
#include <cmath>
#include <complex>
#include <cstddef>
#include <utility>
#include <vector>

namespace Eigen {

using Index = std::ptrdiff_t;

class VectorXcd;

// (row vector) * (column vector) result of v.adjoint() * w; operator()(0)
// gives the scalar, as Eigen's 1x1 product does.
class ScalarProduct1x1 {
public:
    explicit ScalarProduct1x1(std::complex<double> v) : v_(v) {}
    std::complex<double> operator()(Index) const { return v_; }
    std::complex<double> value() const { return v_; }

private:
    std::complex<double> v_;
};

class VectorXcdAdjoint {
public:
    explicit VectorXcdAdjoint(const VectorXcd& v) : v_(v) {}
    ScalarProduct1x1 operator*(const VectorXcd& w) const;

private:
    const VectorXcd& v_;
};

class VectorXcd {
public:
    using Scalar = std::complex<double>;

    VectorXcd() = default;
    explicit VectorXcd(Index n) : d_(size_t(n)) {}

    static VectorXcd Zero(Index n) { return VectorXcd(n); }
    static VectorXcd Constant(Index n, const Scalar& v) {
        VectorXcd r(n);
        r.setConstant(v);
        return r;
    }

    Index size() const { return Index(d_.size()); }
    Index rows() const { return size(); }
    Index cols() const { return 1; }
    void resize(Index n) { d_.resize(size_t(n)); }
    void setZero() { std::fill(d_.begin(), d_.end(), Scalar(0)); }
    void setConstant(const Scalar& v) { std::fill(d_.begin(), d_.end(), v); }
    Scalar* data() { return d_.data(); }
    const Scalar* data() const { return d_.data(); }

    Scalar& operator()(Index i) { return d_[size_t(i)]; }
    const Scalar& operator()(Index i) const { return d_[size_t(i)]; }
    Scalar& operator[](Index i) { return d_[size_t(i)]; }
    const Scalar& operator[](Index i) const { return d_[size_t(i)]; }

    void swap(VectorXcd& o) { d_.swap(o.d_); }

    double squaredNorm() const {
        double s = 0;
        for (const auto& x : d_) s += std::norm(x);
        return s;
    }
    double norm() const { return std::sqrt(squaredNorm()); }

    VectorXcd& operator*=(double s) {
        for (auto& x : d_) x *= s;
        return *this;
    }
    VectorXcd& operator*=(const Scalar& s) {
        for (auto& x : d_) x *= s;
        return *this;
    }
    VectorXcd& operator/=(const Scalar& s) {
        for (auto& x : d_) x /= s;
        return *this;
    }

    VectorXcdAdjoint adjoint() const { return VectorXcdAdjoint(*this); }

private:
    std::vector<Scalar> d_;
};

inline ScalarProduct1x1 VectorXcdAdjoint::operator*(const VectorXcd& w) const {
    assert(v_.size() == w.size());
    std::complex<double> acc = 0;
    for (Index i = 0; i < w.size(); ++i) acc += std::conj(v_(i)) * w(i);
    return ScalarProduct1x1(acc);
}

class MatrixXcd;

class MatrixXcdBlock {
public:
    MatrixXcdBlock(MatrixXcd& m, Index i, Index j, Index r, Index c) : m_(m), i_(i), j_(j), r_(r), c_(c) {}
    MatrixXcdBlock& operator=(const MatrixXcd& src);

private:
    MatrixXcd& m_;
    Index i_, j_, r_, c_;
};

// Column-major dense complex matrix (Eigen's default storage order).
class MatrixXcd {
public:
    using Scalar = std::complex<double>;

    MatrixXcd() = default;
    MatrixXcd(Index r, Index c) : r_(r), c_(c), d_(size_t(r * c)) {}

    static MatrixXcd Zero(Index r, Index c) { return MatrixXcd(r, c); }
    static MatrixXcd Identity(Index r, Index c) {
        MatrixXcd m(r, c);
        for (Index i = 0; i < r && i < c; ++i) m(i, i) = 1.0;
        return m;
    }

    Index rows() const { return r_; }
    Index cols() const { return c_; }
    void resize(Index r, Index c) {
        r_ = r;
        c_ = c;
        d_.assign(size_t(r * c), Scalar(0));
    }
    void setZero() { std::fill(d_.begin(), d_.end(), Scalar(0)); }

    Scalar& operator()(Index i, Index j) { return d_[size_t(j * r_ + i)]; }
    const Scalar& operator()(Index i, Index j) const { return d_[size_t(j * r_ + i)]; }

    MatrixXcdBlock block(Index i, Index j, Index r, Index c) { return MatrixXcdBlock(*this, i, j, r, c); }

    MatrixXcd adjoint() const {
        MatrixXcd a(c_, r_);
        for (Index i = 0; i < r_; ++i)
            for (Index j = 0; j < c_; ++j) a(j, i) = std::conj((*this)(i, j));
        return a;
    }

    MatrixXcd operator*(const MatrixXcd& o) const {
        assert(c_ == o.r_);
        MatrixXcd p(r_, o.c_);
        for (Index j = 0; j < o.c_; ++j)
            for (Index k = 0; k < c_; ++k) {
                const Scalar b = o(k, j);
                for (Index i = 0; i < r_; ++i) p(i, j) += (*this)(i, k) * b;
            }
        return p;
    }

    VectorXcd operator*(const VectorXcd& v) const {
        assert(c_ == v.size());
        VectorXcd out = VectorXcd::Zero(r_);
        for (Index k = 0; k < c_; ++k)
            for (Index i = 0; i < r_; ++i) out(i) += (*this)(i, k) * v(k);
        return out;
    }

    bool isApprox(const MatrixXcd& o, double prec = 1e-12) const {
        if (r_ != o.r_ || c_ != o.c_) return false;
        double diff = 0, a = 0, b = 0;
        for (size_t i = 0; i < d_.size(); ++i) {
            diff += std::norm(d_[i] - o.d_[i]);
            a += std::norm(d_[i]);
            b += std::norm(o.d_[i]);
        }
        return std::sqrt(diff) <= prec * std::sqrt(std::min(a, b));
    }

private:
    Index r_ = 0, c_ = 0;
    std::vector<Scalar> d_;
};

inline MatrixXcdBlock& MatrixXcdBlock::operator=(const MatrixXcd& src) {
    assert(src.rows() == r_ && src.cols() == c_);
    for (Index a = 0; a < r_; ++a)
        for (Index b = 0; b < c_; ++b) m_(i_ + a, j_ + b) = src(a, b);
    return *this;
}

}  // namespace Eigen

#endif
