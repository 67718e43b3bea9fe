// SimpleGates.h - gate base classes and one-qubit gates of the QCSim API
// (reference: QCSim/SimpleGates.h:12-919). Same class names, constructors,
// trait virtuals and matrices; qubit 0 is the least significant bit and a
// k-qubit matrix reads bit `qubit` as local bit 0, `controllingQubit1` as
// bit 1 and `controllingQubit2` as bit 2.
#pragma once

#include <cassert>
#include <cmath>
#include <complex>
#include <cstddef>

#include "qcs_eigen_subset.h"

namespace QC {
namespace Gates {

// Trait virtuals drive the register's kernel choice (SimpleGates.h:27-60).
template <class MatrixClass = Eigen::MatrixXcd> class QuantumGate {
public:
    virtual ~QuantumGate() = default;
    virtual MatrixClass getOperatorMatrix(size_t nrQubits, size_t qubit = 0, size_t controllingQubit1 = 0,
                                          size_t controllingQubit2 = 0) const = 0;
    virtual size_t getQubitsNumber() const { return 0; }
    virtual bool isControlled() const { return false; }
    virtual bool isControlQubit(size_t /*qubit*/) const { return false; }
    virtual bool isDiagonal() const { return false; }
    virtual bool isAntidiagonal() const { return false; }
    virtual bool isSwapGate() const { return false; }
    virtual bool IsISwapGate() const { return false; }
    virtual bool IsISwapDagGate() const { return false; }
};

namespace detail {
// Dense n-qubit operator of a k-qubit matrix (the Kronecker embedding of
// SimpleGates.h:211-232, 311-333, 411-434).
template <class MatrixClass>
MatrixClass embed(const MatrixClass& op, size_t nrQubits, const size_t* qubits, size_t k) {
    const size_t N = size_t(1) << nrQubits;
    MatrixClass m = MatrixClass::Zero(N, N);
    size_t mask = 0;
    for (size_t i = 0; i < k; ++i) mask |= size_t(1) << qubits[i];
    auto local = [&](size_t s) {
        size_t l = 0;
        for (size_t i = 0; i < k; ++i) l |= ((s >> qubits[i]) & 1) << i;
        return l;
    };
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j)
            if ((i | mask) == (j | mask)) m(i, j) = op(local(i), local(j));
    return m;
}
}  // namespace detail

template <class MatrixClass = Eigen::MatrixXcd> class QuantumGateWithOp : public QuantumGate<MatrixClass> {
public:
    using BaseClass = QuantumGate<MatrixClass>;

    QuantumGateWithOp(const QuantumGateWithOp& other) : operatorMat(other.operatorMat) {}
    QuantumGateWithOp(QuantumGateWithOp&& other) : operatorMat(std::move(other.operatorMat)) {}
    QuantumGateWithOp(const MatrixClass& U) : operatorMat(U) {}
    QuantumGateWithOp(MatrixClass&& U) : operatorMat(std::move(U)) {}

    QuantumGateWithOp& operator=(const QuantumGateWithOp& other) {
        operatorMat = other.operatorMat;
        return *this;
    }
    QuantumGateWithOp& operator=(QuantumGateWithOp&& other) {
        operatorMat = std::move(other.operatorMat);
        return *this;
    }
    QuantumGateWithOp& operator=(const MatrixClass& U) {
        operatorMat = U;
        return *this;
    }

    const MatrixClass& getRawOperatorMatrix() const { return operatorMat; }
    size_t getQubitsNumber() const override { return size_t(std::log2(double(operatorMat.rows()))); }
    void setOperator(const MatrixClass& U) {
        assert(U.rows() == U.cols());
        operatorMat = U;
    }

protected:
    MatrixClass operatorMat;
};

template <class MatrixClass = Eigen::MatrixXcd> class SingleQubitGate : public QuantumGateWithOp<MatrixClass> {
public:
    using BaseClass = QuantumGateWithOp<MatrixClass>;
    SingleQubitGate() noexcept : BaseClass(MatrixClass::Zero(2, 2)) {}
    SingleQubitGate(const MatrixClass& U) noexcept : BaseClass(U) { assert(U.rows() == 2 && U.cols() == 2); }
    size_t getQubitsNumber() const override { return 1; }
    MatrixClass getOperatorMatrix(size_t nrQubits, size_t qubit = 0, size_t = 0, size_t = 0) const override {
        const size_t q[1] = {qubit};
        return detail::embed(BaseClass::operatorMat, nrQubits, q, 1);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class TwoQubitsGate : public QuantumGateWithOp<MatrixClass> {
public:
    using BaseClass = QuantumGateWithOp<MatrixClass>;
    TwoQubitsGate() noexcept : BaseClass(MatrixClass::Identity(4, 4)) {}
    TwoQubitsGate(const MatrixClass& U) noexcept : BaseClass(U) { assert(U.rows() == 4 && U.cols() == 4); }
    size_t getQubitsNumber() const override { return 2; }
    MatrixClass getOperatorMatrix(size_t nrQubits, size_t qubit = 0, size_t controllingQubit1 = 0,
                                  size_t = 0) const override {
        const size_t q[2] = {qubit, controllingQubit1};
        return detail::embed(BaseClass::operatorMat, nrQubits, q, 2);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class ThreeQubitsGate : public QuantumGateWithOp<MatrixClass> {
public:
    using BaseClass = QuantumGateWithOp<MatrixClass>;
    ThreeQubitsGate() noexcept : BaseClass(MatrixClass::Identity(8, 8)) {}
    ThreeQubitsGate(const MatrixClass& U) noexcept : BaseClass(U) { assert(U.rows() == 8 && U.cols() == 8); }
    size_t getQubitsNumber() const override { return 3; }
    MatrixClass getOperatorMatrix(size_t nrQubits, size_t qubit = 0, size_t controllingQubit1 = 0,
                                  size_t controllingQubit2 = 0) const override {
        const size_t q[3] = {qubit, controllingQubit1, controllingQubit2};
        return detail::embed(BaseClass::operatorMat, nrQubits, q, 3);
    }
};

// Recorded gate: matrix + qubits, no traits (SimpleGates.h:439-578).
template <class MatrixClass = Eigen::MatrixXcd> class AppliedGate : public QuantumGateWithOp<MatrixClass> {
public:
    using BaseClass = QuantumGateWithOp<MatrixClass>;
    AppliedGate() noexcept : BaseClass(MatrixClass::Zero(1, 1)), q1(0), q2(0), q3(0) {}
    AppliedGate(const MatrixClass& op, size_t q1 = 0, size_t q2 = 0, size_t q3 = 0) noexcept
        : BaseClass(op), q1(q1), q2(q2), q3(q3) {}
    AppliedGate(const AppliedGate& o) noexcept : BaseClass(o.operatorMat), q1(o.q1), q2(o.q2), q3(o.q3) {}
    AppliedGate& operator=(const AppliedGate& o) {
        BaseClass::operator=(o.operatorMat);
        q1 = o.q1;
        q2 = o.q2;
        q3 = o.q3;
        return *this;
    }
    size_t getQubit1() const { return q1; }
    size_t getQubit2() const { return q2; }
    size_t getQubit3() const { return q3; }
    void setQubit1(size_t q) { q1 = q; }
    void setQubit2(size_t q) { q2 = q; }
    void setQubit3(size_t q) { q3 = q; }
    MatrixClass getOperatorMatrix(size_t nrQubits, size_t qubit = 0, size_t c1 = 0, size_t c2 = 0) const override {
        const size_t q[3] = {qubit, c1, c2};
        return detail::embed(BaseClass::operatorMat, nrQubits, q, BaseClass::getQubitsNumber());
    }

private:
    size_t q1, q2, q3;
};

// ---- one-qubit gates (SimpleGates.h:582-919) ---------------------------------
#define QCS_GATE_PREAMBLE(Base)               \
public:                                       \
    using BaseClass = Base<MatrixClass>;      \
    using OpClass = QuantumGateWithOp<MatrixClass>;

template <class MatrixClass = Eigen::MatrixXcd> class HadamardGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    HadamardGate() {
        const double s = 1. / std::sqrt(2.);
        OpClass::operatorMat(0, 0) = s;
        OpClass::operatorMat(0, 1) = s;
        OpClass::operatorMat(1, 0) = s;
        OpClass::operatorMat(1, 1) = -s;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class HyGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    HyGate() {
        const double s = 1. / std::sqrt(2.);
        OpClass::operatorMat(0, 0) = s;
        OpClass::operatorMat(0, 1) = std::complex<double>(0, -s);
        OpClass::operatorMat(1, 0) = std::complex<double>(0, s);
        OpClass::operatorMat(1, 1) = -s;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class DiagonalSingleQubitGate : public SingleQubitGate<MatrixClass> {
public:
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class SGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    SGate() {
        OpClass::operatorMat(0, 0) = 1;
        OpClass::operatorMat(1, 1) = std::complex<double>(0, 1);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class SDGGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    SDGGate() {
        OpClass::operatorMat(0, 0) = 1;
        OpClass::operatorMat(1, 1) = std::complex<double>(0, -1);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class TGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    TGate() {
        OpClass::operatorMat(0, 0) = 1;
        OpClass::operatorMat(1, 1) = std::polar<double>(1, M_PI / 4.);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class TDGGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    TDGGate() {
        OpClass::operatorMat(0, 0) = 1;
        OpClass::operatorMat(1, 1) = std::polar<double>(1, -M_PI / 4.);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class PhaseShiftGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    PhaseShiftGate(double theta = 0) {
        OpClass::operatorMat(0, 0) = 1;
        SetPhaseShift(theta);
    }
    void SetPhaseShift(double theta) { OpClass::operatorMat(1, 1) = std::polar(1., theta); }
};

template <class MatrixClass = Eigen::MatrixXcd> class PauliXGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    PauliXGate() {
        OpClass::operatorMat(0, 1) = 1;
        OpClass::operatorMat(1, 0) = 1;
    }
    bool isAntidiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class PauliYGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    PauliYGate() {
        OpClass::operatorMat(0, 1) = std::complex<double>(0, -1);
        OpClass::operatorMat(1, 0) = std::complex<double>(0, 1);
    }
    bool isAntidiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class PauliZGate : public DiagonalSingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(DiagonalSingleQubitGate)
    PauliZGate() {
        OpClass::operatorMat(0, 0) = 1;
        OpClass::operatorMat(1, 1) = -1;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class SquareRootNOTGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    SquareRootNOTGate() {
        OpClass::operatorMat(0, 0) = OpClass::operatorMat(1, 1) = std::complex<double>(0.5, 0.5);
        OpClass::operatorMat(0, 1) = OpClass::operatorMat(1, 0) = std::complex<double>(0.5, -0.5);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class SquareRootNOTDagGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    SquareRootNOTDagGate() {
        OpClass::operatorMat(0, 0) = OpClass::operatorMat(1, 1) = std::complex<double>(0.5, -0.5);
        OpClass::operatorMat(0, 1) = OpClass::operatorMat(1, 0) = std::complex<double>(0.5, 0.5);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class SplitterGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    SplitterGate() {
        const double s = 1. / std::sqrt(2.);
        OpClass::operatorMat(0, 0) = OpClass::operatorMat(1, 1) = s;
        OpClass::operatorMat(0, 1) = OpClass::operatorMat(1, 0) = std::complex<double>(0, s);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class RotationGate : public SingleQubitGate<MatrixClass> {
public:
    virtual void SetTheta(double theta) = 0;
};

template <class MatrixClass = Eigen::MatrixXcd> class RxGate : public RotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(RotationGate)
    RxGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        OpClass::operatorMat(0, 0) = OpClass::operatorMat(1, 1) = std::complex<double>(c, 0);
        OpClass::operatorMat(0, 1) = OpClass::operatorMat(1, 0) = std::complex<double>(0, -s);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class RyGate : public RotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(RotationGate)
    RyGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        OpClass::operatorMat(0, 0) = OpClass::operatorMat(1, 1) = c;
        OpClass::operatorMat(0, 1) = -s;
        OpClass::operatorMat(1, 0) = s;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class RzGate : public RotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(RotationGate)
    RzGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        OpClass::operatorMat(0, 0) = std::polar(1., -theta * 0.5);
        OpClass::operatorMat(1, 1) = std::polar(1., theta * 0.5);
    }
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class UGate : public SingleQubitGate<MatrixClass> {
    QCS_GATE_PREAMBLE(SingleQubitGate)
    UGate(double theta = 0, double phi = 0, double lambda = 0, double gamma = 0) { SetParams(theta, phi, lambda, gamma); }
    virtual void SetParams(double theta = 0, double phi = 0, double lambda = 0, double gamma = 0) {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        OpClass::operatorMat(0, 0) = std::polar(1., gamma) * c;
        OpClass::operatorMat(0, 1) = -std::polar(1., gamma + lambda) * s;
        OpClass::operatorMat(1, 0) = std::polar(1., gamma + phi) * s;
        OpClass::operatorMat(1, 1) = std::polar(1., gamma + phi + lambda) * c;
    }
};

}  // namespace Gates
}  // namespace QC
