// QuantumAlgorithm.h - the algorithm / subalgorithm framework of the QCSim
// API (reference: QCSim/QuantumAlgorithm.h:7-257), unchanged in shape: a
// QuantumAlgorithm owns a register and proxies it; subalgorithms act on a
// register passed to Execute().
#pragma once

#include <algorithm>
#include <cassert>
#include <climits>
#include <vector>

#include "QubitRegisterDebug.h"

namespace QC {

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd> class QuantumAlgorithm {
public:
    using RegisterClass = QubitRegisterDebug<VectorClass, MatrixClass>;
    using GateClass = Gates::QuantumGateWithOp<MatrixClass>;

    QuantumAlgorithm(size_t N = 3, unsigned int addseed = 0) : reg(N, addseed) { assert(N > 0); }
    virtual ~QuantumAlgorithm() {}

    virtual size_t Execute() = 0;

    size_t getNrQubits() const { return reg.getNrQubits(); }
    size_t getNrBasisStates() const { return reg.getNrBasisStates(); }
    std::complex<double> getBasisStateAmplitude(size_t State) const { return reg.getBasisStateAmplitude(State); }
    double getBasisStateProbability(size_t State) const { return reg.getBasisStateProbability(State); }
    void setToBasisState(size_t State) { reg.setToBasisState(State); }
    void setToQubitState(size_t q) { reg.setToQubitState(q); }
    void setToCatState() { reg.setToCatState(); }
    void setToEqualSuperposition() { reg.setToEqualSuperposition(); }
    void setRawAmplitude(size_t State, std::complex<double> val) { reg.setRawAmplitude(State, val); }
    void Clear() { reg.Clear(); }
    void Normalize() { reg.Normalize(); }
    void AdjustPhaseAndNormalize() { reg.AdjustPhaseAndNormalize(); }
    static size_t getQubitState(size_t q) { return 1ULL << q; }
    const RegisterClass& getRegister() const { return reg; }
    const VectorClass& getRegisterStorage() const { return reg.getRegisterStorage(); }
    void setRegisterStorage(const VectorClass& vals) { reg.setRegisterStorage(vals); }

    void ApplyGate(const GateClass& gate, size_t qubit, size_t controllingQubit = 0, size_t controllingQubit2 = 0) {
        reg.ApplyGate(gate, qubit, controllingQubit, controllingQubit2);
    }
    void ApplyGate(const Gates::AppliedGate<MatrixClass>& gate) { reg.ApplyGate(gate); }
    void ApplyGates(const std::vector<Gates::AppliedGate<MatrixClass>>& gates) { reg.ApplyGates(gates); }
    void ApplyOperatorMatrix(const MatrixClass& m) { reg.ApplyOperatorMatrix(m); }

    size_t Measure() { return reg.MeasureAll(); }
    size_t Measure(size_t qubit) { return reg.MeasureQubit(qubit); }
    size_t Measure(size_t firstQubit, size_t secondQubit) { return reg.Measure(firstQubit, secondQubit); }
    std::map<size_t, size_t> RepeatedMeasure(size_t nrTimes = 1000) { return reg.RepeatedMeasure(nrTimes); }
    std::map<size_t, size_t> RepeatedMeasure(size_t firstQubit, size_t secondQubit, size_t nrTimes = 1000) {
        return reg.RepeatedMeasure(firstQubit, secondQubit, nrTimes);
    }
    double GetQubitProbability(size_t qubit) const { return reg.GetQubitProbability(qubit); }

    // |<reg|state>|^2 against a known state (QuantumAlgorithm.h:143-146)
    double stateFidelity(const VectorClass& state) const { return reg.stateFidelity(state); }

    // Host-side twin of the register's AdjustPhaseAndNormalize for a vector
    // the caller holds (QuantumAlgorithm.h:148-162): every entry is divided
    // by entry 0 (by 1 when |entry 0| <= 1e-5), then the vector is scaled
    // to unit norm.
    static void AdjustPhaseAndNormalize(VectorClass& regVals) {
        const size_t count = static_cast<size_t>(regVals.size());
        if (count == 0) return;
        const std::complex<double> first = regVals[0];
        const std::complex<double> div = std::abs(first) > 1E-5 ? first : std::complex<double>(1, 0);
        double sum = 0;
        for (size_t i = 0; i < count; ++i) {
            regVals[i] /= div;
            sum += std::norm(regVals[i]);
        }
        const double scale = 1. / std::sqrt(sum);
        for (size_t i = 0; i < count; ++i) regVals[i] *= scale;
    }

    // text dumps of the register (QuantumAlgorithm.h:164-172)
    bool writeToFile(const std::string& name, bool amplitude = true, bool append = false) const {
        return reg.writeToFile(name, amplitude, append);
    }
    void displayRegister() const { reg.displayRegister(); }

    void ComputeStart() { reg.ComputeStart(); }
    void ComputeEnd() { reg.ComputeEnd(); }
    void ComputeClear() { reg.ComputeClear(); }
    void Compute() { reg.Compute(); }
    void Uncompute() { reg.Uncompute(); }

protected:
    RegisterClass reg;
};

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd> class QuantumSubAlgorithm {
public:
    using RegisterClass = QubitRegister<VectorClass, MatrixClass>;
    QuantumSubAlgorithm() {}
    virtual ~QuantumSubAlgorithm() {}
    virtual size_t Execute(RegisterClass& reg) = 0;
};

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QuantumSubAlgorithmOnSubregister : public QuantumSubAlgorithm<VectorClass, MatrixClass> {
public:
    using BaseClass = QuantumSubAlgorithm<VectorClass, MatrixClass>;
    using RegisterClass = QubitRegister<VectorClass, MatrixClass>;

    // the end qubit is clamped to the register (QuantumAlgorithm.h:228-229)
    QuantumSubAlgorithmOnSubregister(size_t N, size_t startQubit = 0, size_t endQubit = INT_MAX)
        : sQubit(startQubit), eQubit(std::max(startQubit, std::min(N - 1, endQubit))) {}

    size_t getStartQubit() const { return sQubit; }
    size_t getEndQubit() const { return eQubit; }

protected:
    size_t sQubit;
    size_t eQubit;
};

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QuantumSubAlgorithmOnSubregisterWithAncilla : public QuantumSubAlgorithmOnSubregister<VectorClass, MatrixClass> {
public:
    using BaseClass = QuantumSubAlgorithmOnSubregister<VectorClass, MatrixClass>;
    QuantumSubAlgorithmOnSubregisterWithAncilla(size_t N, size_t startQubit = 0, size_t endQubit = INT_MAX,
                                                size_t startAncila = INT_MAX)
        : BaseClass(N, startQubit, endQubit), sAncilla(startAncila) {}
    size_t getStartAncilla() const { return sAncilla; }

protected:
    size_t sAncilla;
};

}  // namespace QC
