// QuantumFourierTransform.h - QC::SubAlgo::QuantumFourierTransform
// (reference: QCSim/QuantumFourierTransform.h:12-92). QFT/IQFT keep their
// signatures and results (the reference's H / controlled-phase / SWAP
// sequence); they run as a few fused passes on the GPU, or gate by gate while
// the register records gates. `hadamard` and `cPhaseShift` stay public for
// callers that build on them (e.g. a Draper adder).
#pragma once

#include "QubitsSwapper.h"

namespace QC {
namespace SubAlgo {

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QuantumFourierTransform : public QubitsSwapper<VectorClass, MatrixClass> {
public:
    using BaseClass = QubitsSwapper<VectorClass, MatrixClass>;
    using RegisterClass = QubitRegister<VectorClass, MatrixClass>;

    QuantumFourierTransform(size_t N, size_t startQubit = 0, size_t endQubit = INT_MAX)
        : BaseClass(N, startQubit, endQubit) {}

    size_t Execute(RegisterClass& reg) override {
        QFT(reg);
        return reg.MeasureAll();
    }

    void QFT(RegisterClass& reg, bool doSwap = true) {
        reg.ApplyQFT(this->getStartQubit(), this->getEndQubit(), doSwap, false);
    }

    void IQFT(RegisterClass& reg, bool doSwap = true) {
        reg.ApplyQFT(this->getStartQubit(), this->getEndQubit(), doSwap, true);
    }

    Gates::HadamardGate<MatrixClass> hadamard;
    Gates::ControlledPhaseShiftGate<MatrixClass> cPhaseShift;
};

}  // namespace SubAlgo
}  // namespace QC
