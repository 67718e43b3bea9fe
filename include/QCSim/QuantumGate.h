// QuantumGate.h - two- and three-qubit gates of the QCSim API
// (reference: QCSim/QuantumGate.h:10-469). Same classes, traits and matrices.
#pragma once

#include "SimpleGates.h"

namespace QC {
namespace Gates {

template <class MatrixClass = Eigen::MatrixXcd> class SwapGate : public TwoQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsGate)
    SwapGate() {
        OpClass::operatorMat(1, 1) = OpClass::operatorMat(2, 2) = 0;
        OpClass::operatorMat(1, 2) = OpClass::operatorMat(2, 1) = 1;
    }
    bool isSwapGate() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class iSwapGate : public TwoQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsGate)
    iSwapGate() {
        OpClass::operatorMat(1, 1) = OpClass::operatorMat(2, 2) = 0;
        OpClass::operatorMat(1, 2) = OpClass::operatorMat(2, 1) = std::complex<double>(0, 1);
    }
    bool IsISwapGate() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class iSwapDagGate : public TwoQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsGate)
    iSwapDagGate() {
        OpClass::operatorMat(1, 1) = OpClass::operatorMat(2, 2) = 0;
        OpClass::operatorMat(1, 2) = OpClass::operatorMat(2, 1) = std::complex<double>(0, -1);
    }
    bool IsISwapDagGate() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class DecrementGate : public TwoQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsGate)
    DecrementGate() {
        for (int i = 0; i < 4; ++i) OpClass::operatorMat(i, i) = 0;
        OpClass::operatorMat(3, 0) = 1;
        OpClass::operatorMat(0, 1) = 1;
        OpClass::operatorMat(1, 2) = 1;
        OpClass::operatorMat(2, 3) = 1;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class TwoQubitsControlledGate : public TwoQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsGate)
    TwoQubitsControlledGate() {}
    bool isControlled() const override { return true; }
    bool isControlQubit(size_t qubit) const override { return qubit == 0; }
    void SetOperation(const MatrixClass& U) {
        assert(U.rows() == 2 && U.cols() == 2);
        OpClass::operatorMat.block(2, 2, 2, 2) = U;
    }

protected:
    void setBlock(std::complex<double> a, std::complex<double> b, std::complex<double> c, std::complex<double> d) {
        OpClass::operatorMat(2, 2) = a;
        OpClass::operatorMat(2, 3) = b;
        OpClass::operatorMat(3, 2) = c;
        OpClass::operatorMat(3, 3) = d;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class CNOTGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    CNOTGate() { this->setBlock(0, 1, 1, 0); }
    bool isAntidiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledYGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledYGate() { this->setBlock(0, std::complex<double>(0, -1), std::complex<double>(0, 1), 0); }
    bool isAntidiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledZGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledZGate() { OpClass::operatorMat(3, 3) = -1; }
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ControlledHadamardGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledHadamardGate() {
        const double s = 1. / std::sqrt(2.);
        this->setBlock(s, s, s, -s);
    }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ControlledSquareRootNOTGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledSquareRootNOTGate() {
        const std::complex<double> a(0.5, 0.5), b(0.5, -0.5);
        this->setBlock(a, b, b, a);
    }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ControlledSquareRootNOTDagGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledSquareRootNOTDagGate() {
        const std::complex<double> a(0.5, -0.5), b(0.5, 0.5);
        this->setBlock(a, b, b, a);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledPhaseGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledPhaseGate() { OpClass::operatorMat(3, 3) = std::complex<double>(0, 1); }
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ControlledPhaseShiftGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledPhaseShiftGate(double theta = 0) { SetPhaseShift(theta); }
    void SetPhaseShift(double theta) { OpClass::operatorMat(3, 3) = std::polar(1., theta); }
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledUGate : public TwoQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(TwoQubitsControlledGate)
    ControlledUGate(double theta = 0, double phi = 0, double lambda = 0, double gamma = 0) {
        SetParams(theta, phi, lambda, gamma);
    }
    virtual void SetParams(double theta = 0, double phi = 0, double lambda = 0, double gamma = 0) {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        this->setBlock(std::polar(1., gamma) * c, -std::polar(1., gamma + lambda) * s, std::polar(1., gamma + phi) * s,
                       std::polar(1., gamma + phi + lambda) * c);
    }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ControlledRotationGate : public TwoQubitsControlledGate<MatrixClass> {
public:
    virtual void SetTheta(double theta) = 0;
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledRxGate : public ControlledRotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ControlledRotationGate)
    ControlledRxGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        this->setBlock(c, std::complex<double>(0, -s), std::complex<double>(0, -s), c);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledRyGate : public ControlledRotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ControlledRotationGate)
    ControlledRyGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        const double c = std::cos(theta * 0.5), s = std::sin(theta * 0.5);
        this->setBlock(c, -s, s, c);
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class ControlledRzGate : public ControlledRotationGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ControlledRotationGate)
    ControlledRzGate(double theta = 0) { SetTheta(theta); }
    void SetTheta(double theta) override {
        OpClass::operatorMat(2, 2) = std::polar(1., -theta * 0.5);
        OpClass::operatorMat(3, 3) = std::polar(1., theta * 0.5);
    }
    bool isDiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd>
class ThreeQubitsControlledGate : public ThreeQubitsGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ThreeQubitsGate)
    ThreeQubitsControlledGate() {}
    bool isControlled() const override { return true; }
    bool isControlQubit(size_t qubit) const override { return qubit == 0; }
    void SetOperation(const MatrixClass& U) {
        assert(U.rows() == 4 && U.cols() == 4);
        OpClass::operatorMat.block(4, 4, 4, 4) = U;
    }
};

template <class MatrixClass = Eigen::MatrixXcd> class ToffoliGate : public ThreeQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ThreeQubitsControlledGate)
    ToffoliGate() {
        OpClass::operatorMat(6, 6) = OpClass::operatorMat(7, 7) = 0;
        OpClass::operatorMat(6, 7) = OpClass::operatorMat(7, 6) = 1;
    }
    bool isControlQubit(size_t qubit) const override { return qubit <= 1; }
    bool isAntidiagonal() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class FredkinGate : public ThreeQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ThreeQubitsControlledGate)
    FredkinGate() {
        OpClass::operatorMat(5, 5) = OpClass::operatorMat(6, 6) = 0;
        OpClass::operatorMat(5, 6) = OpClass::operatorMat(6, 5) = 1;
    }
    bool isSwapGate() const override { return true; }
};

template <class MatrixClass = Eigen::MatrixXcd> class CCZGate : public ThreeQubitsControlledGate<MatrixClass> {
    QCS_GATE_PREAMBLE(ThreeQubitsControlledGate)
    CCZGate() { OpClass::operatorMat(7, 7) = -1; }
    bool isControlQubit(size_t qubit) const override { return qubit <= 1; }
    bool isDiagonal() const override { return true; }
};

}  // namespace Gates
}  // namespace QC
