// QubitsSwapper.h - QC::SubAlgo::QubitsSwapper (reference:
// QCSim/QubitsSwapper.h:12-44): reverses the qubit order of [sq, eq] with
// SWAP gates.
#pragma once

#include "QuantumAlgorithm.h"
#include "QuantumGate.h"
#include "QubitRegister.h"

namespace QC {
namespace SubAlgo {

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QubitsSwapper : public QuantumSubAlgorithmOnSubregister<VectorClass, MatrixClass> {
public:
    using BaseClass = QuantumSubAlgorithmOnSubregister<VectorClass, MatrixClass>;
    using RegisterClass = QubitRegister<VectorClass, MatrixClass>;

    QubitsSwapper(size_t N, size_t startQubit = 0, size_t endQubit = INT_MAX) : BaseClass(N, startQubit, endQubit) {}

    void Swap(RegisterClass& reg) const {
        for (size_t a = BaseClass::getStartQubit(), b = BaseClass::getEndQubit(); a < b; ++a, --b)
            reg.ApplyGate(swapOp, a, b);
    }

    size_t Execute(RegisterClass& reg) override {
        Swap(reg);
        return reg.MeasureAll();
    }

protected:
    Gates::SwapGate<MatrixClass> swapOp;
};

}  // namespace SubAlgo
}  // namespace QC
