// QubitRegisterDebug.h - QC::QubitRegisterDebug (reference:
// QCSim/QubitRegisterDebug.h:20-90): the register plus text dumps of the
// state, which QuantumAlgorithm owns.
#pragma once

#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "QubitRegister.h"

namespace QC {

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QubitRegisterDebug : public QubitRegister<VectorClass, MatrixClass> {
public:
    using BaseClass = QubitRegister<VectorClass, MatrixClass>;

    QubitRegisterDebug(size_t N = 3, unsigned int addseed = 0) : BaseClass(N, addseed) {}

    // one "index re im" line per basis state
    bool writeToFile(const std::string& name, bool amplitude = true, bool append = false) const {
        std::ofstream f(name, append ? std::ios::app : std::ios::out);
        if (!f) return false;
        const auto& v = BaseClass::getRegisterStorage();
        f << std::setprecision(17);
        for (size_t i = 0; i < BaseClass::getNrBasisStates(); ++i) {
            if (amplitude) f << i << " " << v(i).real() << " " << v(i).imag() << "\n";
            else f << i << " " << std::norm(v(i)) << "\n";
        }
        return bool(f);
    }

    void displayRegister() const {
        const auto& v = BaseClass::getRegisterStorage();
        for (size_t i = 0; i < BaseClass::getNrBasisStates(); ++i) std::cout << i << " " << v(i) << "\n";
    }
};

}  // namespace QC
