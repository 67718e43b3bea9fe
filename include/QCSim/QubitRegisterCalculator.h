// QubitRegisterCalculator.h - the calculator base of QC::QubitRegister
// (reference: QCSim/QubitRegisterCalculator.h:32-1295).
//
// In the reference this class holds the static per-state CPU loops. Here the
// loops are gfx950 kernels behind the C-ABI (include/qcsim_amd.h), so only
// the public knobs remain: SetMultithreading/GetMultithreading keep their
// signatures as a flag with no effect (the device always runs in parallel),
// GetNumberOfThreads reports the host's hardware threads, and the OpenMP
// thresholds stay declared for code that reads them.
#pragma once

#include <thread>

#include "QuantumGate.h"

namespace QC {

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd> class QubitRegisterCalculator {
public:
    using GateClass = Gates::QuantumGateWithOp<MatrixClass>;

    QubitRegisterCalculator() = default;
    virtual ~QubitRegisterCalculator() = default;

    static int GetNumberOfThreads() {
        const unsigned t = std::thread::hardware_concurrency();
        return int(t ? t : 1);
    }

    void SetMultithreading(bool enable = true) { enableMultithreading = enable; }
    bool GetMultithreading() const { return enableMultithreading; }

    constexpr static size_t OneQubitOmpLimit = 8192;
    constexpr static size_t TwoQubitOmpLimit = OneQubitOmpLimit;
    constexpr static size_t ThreeQubitOmpLimit = OneQubitOmpLimit;

private:
    bool enableMultithreading = true;
};

}  // namespace QC
