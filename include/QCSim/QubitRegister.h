// QubitRegister.h - QC::QubitRegister over the MI355X engine
// (reference: QCSim/QubitRegister.h:11-728).
//
// Same constructors, methods, defaults and semantics; the amplitudes live in
// HBM behind a qcs_register handle (include/qcsim_amd.h) and every operation
// runs on the GPU. Differences a caller can observe are listed where they
// occur:
//   - getRegisterStorage() returns a host mirror that is refreshed lazily
//     (one download after any mutation); the reference is valid until the
//     next mutating call, as in the reference;
//   - setRegisterStorageFastNoNormalize(v) and the (N, v) constructor keep the
//     reference's swap semantics: v receives the previous contents (the
//     constructor leaves v empty, like the reference's default-built storage);
//   - there is no results buffer: kernels update the state in place.
// Random draws use std::mt19937_64 + uniform_real_distribution seeded as the
// reference does (time + addseed), and every measurement consumes exactly
// the draws the reference consumes, with prob = 1 - u.
#pragma once

#include <array>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../qcsim_amd.h"
#include "QuantumGate.h"
#include "QubitRegisterCalculator.h"

namespace QC {

namespace detail {
inline void qcs_check(int code) {
    if (code == QCS_OK) return;
    const std::string msg = qcs_last_error();
    if (code == QCS_EINVAL) throw std::invalid_argument(msg);
    if (code == QCS_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

template <class GateT> uint32_t gate_traits(const GateT& g) {
    uint32_t t = 0;
    if (g.isControlled()) t |= QCS_TRAIT_CONTROLLED;
    if (g.isControlQubit(1)) t |= QCS_TRAIT_CTRL_Q1;
    if (g.isDiagonal()) t |= QCS_TRAIT_DIAGONAL;
    if (g.isAntidiagonal()) t |= QCS_TRAIT_ANTIDIAGONAL;
    if (g.isSwapGate()) t |= QCS_TRAIT_SWAP;
    if (g.IsISwapGate()) t |= QCS_TRAIT_ISWAP;
    if (g.IsISwapDagGate()) t |= QCS_TRAIT_ISWAPDAG;
    return t;
}

// A list of single-qubit gates whose product is coef * P(x, z) with P a
// Pauli string (qubit q: X for bit q of x only, Z for z only, Y = iXZ for
// both): per qubit the 2x2 product (later gates on the left) must be a
// scalar multiple of I, X, Y or Z, exactly.
template <class AppliedGateT>
bool gates_as_pauli(const std::vector<AppliedGateT>& gates, size_t nqubits, std::complex<double>& coef,
                    uint64_t& x, uint64_t& z) {
    using cd = std::complex<double>;
    std::map<size_t, std::array<cd, 4>> per;
    for (const auto& g : gates) {
        if (g.getQubitsNumber() != 1) return false;
        const size_t q = g.getQubit1();
        if (q >= nqubits || q >= 64) return false;
        const auto& m = g.getRawOperatorMatrix();
        const std::array<cd, 4> a{cd(m(0, 0)), cd(m(0, 1)), cd(m(1, 0)), cd(m(1, 1))};
        auto it = per.find(q);
        if (it == per.end()) {
            per.emplace(q, a);
        } else {
            const std::array<cd, 4> b = it->second;
            it->second = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                          a[2] * b[1] + a[3] * b[3]};
        }
    }
    coef = 1;
    x = z = 0;
    for (const auto& kv : per) {
        const std::array<cd, 4>& m = kv.second;
        const uint64_t bit = uint64_t(1) << kv.first;
        if (m[1] == cd(0) && m[2] == cd(0)) {
            if (m[0] == m[3]) coef *= m[0];
            else if (m[0] == -m[3]) {
                coef *= m[0];
                z |= bit;
            } else
                return false;
        } else if (m[0] == cd(0) && m[3] == cd(0)) {
            if (m[1] == m[2]) {
                coef *= m[1];
                x |= bit;
            } else if (m[2] == -m[1]) {  // c Y has entry (1, 0) = i c
                coef *= cd(0, -1) * m[2];
                x |= bit;
                z |= bit;
            } else
                return false;
        } else {
            return false;
        }
    }
    return true;
}

// Row-major interleaved (re, im) copy of a matrix.
template <class MatrixClass> std::vector<double> flatten(const MatrixClass& m) {
    const size_t r = size_t(m.rows()), c = size_t(m.cols());
    std::vector<double> out(2 * r * c);
    for (size_t i = 0; i < r; ++i)
        for (size_t j = 0; j < c; ++j) {
            const std::complex<double> z = m(i, j);
            out[2 * (i * c + j)] = z.real();
            out[2 * (i * c + j) + 1] = z.imag();
        }
    return out;
}
}  // namespace detail

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QubitRegister : public QubitRegisterCalculator<VectorClass, MatrixClass> {
public:
    using GateClass = Gates::QuantumGateWithOp<MatrixClass>;
    using BaseClass = QubitRegisterCalculator<VectorClass, MatrixClass>;

    QubitRegister(size_t N = 3, unsigned int addseed = 0)
        : NrQubits(N), NrBasisStates(1ULL << N), uniformZeroOne(0, 1), recordGates(false) {
        assert(N > 0);
        detail::qcs_check(qcs_create(uint32_t(N), QCS_COMPLEX128, &h_));
        seed(addseed);
    }

    QubitRegister(size_t N, VectorClass& v, unsigned int addseed = 0)
        : NrQubits(N), NrBasisStates(1ULL << N), uniformZeroOne(0, 1), recordGates(false) {
        assert(N > 0);
        detail::qcs_check(qcs_create(uint32_t(N), QCS_COMPLEX128, &h_));
        upload(v);
        v = VectorClass();  // the reference swaps v with its (empty) storage
        seed(addseed);
    }

    QubitRegister(const QubitRegister& o)
        : BaseClass(o), NrQubits(o.NrQubits), NrBasisStates(o.NrBasisStates), uniformZeroOne(0, 1),
          computeGates(o.computeGates), recordGates(o.recordGates) {
        detail::qcs_check(qcs_create(uint32_t(NrQubits), QCS_COMPLEX128, &h_));
        detail::qcs_check(qcs_copy_state(h_, o.h_));
        if (o.saved_) {
            detail::qcs_check(qcs_create(uint32_t(NrQubits), QCS_COMPLEX128, &saved_));
            detail::qcs_check(qcs_copy_state(saved_, o.saved_));
        }
        rng = o.rng;
    }

    QubitRegister& operator=(const QubitRegister&) = delete;

    virtual ~QubitRegister() {
        progFwd_.reset();  // programs first: they sync the register's stream
        progAdj_.reset();
        if (shadow_) qcs_destroy(shadow_);
        if (saved_) qcs_destroy(saved_);
        if (h_) qcs_destroy(h_);
    }

    size_t getNrQubits() const { return NrQubits; }
    size_t getNrBasisStates() const { return NrBasisStates; }

    std::complex<double> getBasisStateAmplitude(size_t State) const {
        if (State >= NrBasisStates) return 0;
        double z[2];
        detail::qcs_check(qcs_get_amplitude(h_, State, z));
        return {z[0], z[1]};
    }

    double getBasisStateProbability(size_t State) const {
        if (State >= NrBasisStates) return 0;
        return std::norm(getBasisStateAmplitude(State));
    }

    void setToBasisState(size_t State) {
        if (State >= NrBasisStates) return;
        detail::qcs_check(qcs_set_basis_state(h_, State));
        dirty_ = true;
    }

    void setToQubitState(size_t q) {
        if (q >= NrQubits) return;
        setToBasisState(1ULL << q);
    }

    void setToCatState() {
        Clear();
        const double s = 1. / std::sqrt(2.);
        detail::qcs_check(qcs_set_amplitude(h_, 0, s, 0));
        detail::qcs_check(qcs_set_amplitude(h_, NrBasisStates - 1, s, 0));
    }

    void Reset() { setToBasisState(0); }

    void setToEqualSuperposition() {
        detail::qcs_check(qcs_fill(h_, 1. / std::sqrt(double(NrBasisStates)), 0));
        dirty_ = true;
    }

    void setRawAmplitude(size_t State, std::complex<double> val) {
        if (State >= NrBasisStates) return;
        detail::qcs_check(qcs_set_amplitude(h_, State, val.real(), val.imag()));
        dirty_ = true;
    }

    void Clear() {
        detail::qcs_check(qcs_fill(h_, 0, 0));
        dirty_ = true;
    }

    void Normalize() {
        detail::qcs_check(qcs_normalize(h_));
        dirty_ = true;
    }

    // QubitRegister.h:133-167
    void AdjustPhaseAndNormalize() {
        const size_t probe[3] = {0, NrBasisStates >> 1, NrBasisStates - 1};
        for (size_t idx : probe) {
            const std::complex<double> v0 = getBasisStateAmplitude(idx);
            if (std::abs(v0) >= 1E-5) {
                const std::complex<double> inv = 1. / v0;
                detail::qcs_check(qcs_scale(h_, inv.real(), inv.imag()));
                break;
            }
        }
        Normalize();
    }

    size_t MeasureAll() {
        uint64_t out = 0;
        detail::qcs_check(qcs_measure_all(h_, uniformZeroOne(rng), &out));
        dirty_ = true;
        return size_t(out);
    }

    size_t MeasureQubit(size_t qubit) { return Measure(qubit, qubit); }

    size_t Measure(size_t firstQubit, size_t secondQubit) {
        uint64_t out = 0;
        detail::qcs_check(qcs_measure(h_, uint32_t(firstQubit), uint32_t(secondQubit), uniformZeroOne(rng), &out));
        dirty_ = true;
        return size_t(out);
    }

    std::map<size_t, size_t> RepeatedMeasure(size_t nrTimes = 1000) { return sampleMap<std::map<size_t, size_t>>(true, 0, 0, nrTimes); }
    std::unordered_map<size_t, size_t> RepeatedMeasureUnordered(size_t nrTimes = 1000) {
        return sampleMap<std::unordered_map<size_t, size_t>>(true, 0, 0, nrTimes);
    }
    std::map<size_t, size_t> RepeatedMeasure(size_t firstQubit, size_t secondQubit, size_t nrTimes = 1000) {
        return sampleMap<std::map<size_t, size_t>>(false, firstQubit, secondQubit, nrTimes);
    }
    std::unordered_map<size_t, size_t> RepeatedMeasureUnordered(size_t firstQubit, size_t secondQubit,
                                                                size_t nrTimes = 1000) {
        return sampleMap<std::unordered_map<size_t, size_t>>(false, firstQubit, secondQubit, nrTimes);
    }

    // controllingQubit1 is for 2-qubit gates, controllingQubit2 for 3-qubit gates
    void ApplyGate(const GateClass& gate, size_t qubit, size_t controllingQubit1 = 0, size_t controllingQubit2 = 0) {
        const size_t gateQubits = gate.getQubitsNumber();
        CheckQubits(gate, qubit, controllingQubit1, controllingQubit2, gateQubits);
        assert(gateQubits > 0 && gateQubits <= 3);
        const std::vector<double> m = detail::flatten(gate.getRawOperatorMatrix());
        detail::qcs_check(qcs_apply_gate(h_, uint32_t(gateQubits), detail::gate_traits(gate), uint32_t(qubit),
                                         uint32_t(controllingQubit1), uint32_t(controllingQubit2), m.data()));
        dirty_ = true;
        if (recordGates) {
            computeGates.emplace_back(
                Gates::AppliedGate<MatrixClass>(gate.getRawOperatorMatrix(), qubit, controllingQubit1, controllingQubit2));
            ++computeGen_;
        }
    }

    void ApplyGate(const Gates::AppliedGate<MatrixClass>& gate) {
        ApplyGate(gate, gate.getQubit1(), gate.getQubit2(), gate.getQubit3());
    }

    void ApplyGates(const std::vector<Gates::AppliedGate<MatrixClass>>& gates) {
        for (const auto& gate : gates) ApplyGate(gate);
    }

    // Dense 2^n x 2^n operator on the GPU: a diagonal operator in one pass
    // over the state, any other as an FP64 matrix-vector product with the
    // matrix streamed in panels of rows (registers of up to 21 qubits).
    void ApplyOperatorMatrix(const MatrixClass& m) {
        bool diag = true;
        for (size_t i = 0; i < NrBasisStates && diag; ++i)
            for (size_t j = 0; j < NrBasisStates; ++j)
                if (i != j && m(i, j) != std::complex<double>(0)) {
                    diag = false;
                    break;
                }
        if (diag) {
            std::vector<double> d(2 * NrBasisStates);
            for (size_t i = 0; i < NrBasisStates; ++i) {
                d[2 * i] = m(i, i).real();
                d[2 * i + 1] = m(i, i).imag();
            }
            detail::qcs_check(qcs_apply_diagonal(h_, d.data(), NrBasisStates));
        } else {
            const std::vector<double> rows = detail::flatten(m);
            detail::qcs_check(qcs_apply_matrix(h_, rows.data(), NrBasisStates));
        }
        dirty_ = true;
        if (recordGates) {
            computeGates.emplace_back(Gates::AppliedGate<MatrixClass>(m));
            ++computeGen_;
        }
    }

    const VectorClass& getRegisterStorage() const {
        if (dirty_) {
            mirror_.resize(Eigen::Index(NrBasisStates));
            detail::qcs_check(qcs_get_state(h_, reinterpret_cast<double*>(mirror_.data()), NrBasisStates));
            dirty_ = false;
        }
        return mirror_;
    }

    // The same, with the capitalised spelling some callers use (the reference
    // spells it getRegisterStorage, QubitRegister.h:507).
    const VectorClass& GetRegisterStorage() const { return getRegisterStorage(); }

    void setRegisterStorage(const VectorClass& vals) {
        if (size_t(vals.size()) != NrBasisStates) return;
        upload(vals);
        Normalize();
    }

    // warning: vals must be normalised and have the right size (no check).
    // The reference swaps vals with its storage (:521-524). Here vals is
    // uploaded; it receives the previous contents when the host mirror holds
    // them already (no mutation since the last getRegisterStorage), and is
    // left empty otherwise, so no state-sized download is ever made.
    void setRegisterStorageFastNoNormalize(VectorClass& vals) {
        const bool mirrored = !dirty_;
        upload(vals);
        if (mirrored) {
            mirror_.swap(vals);  // mirror = the uploaded state, vals = the old one
            dirty_ = false;
        } else {
            vals = VectorClass();
        }
    }

    // |<reg|state>|^2 (:527-534): the vector streams through the device in
    // chunks, no second register is made.
    double stateFidelity(const VectorClass& state) const {
        if (size_t(state.size()) != NrBasisStates) return 0;
        double z[2];
        detail::qcs_check(qcs_inner_host(h_, reinterpret_cast<const double*>(state.data()), NrBasisStates, z));
        return std::norm(std::complex<double>(z[0], z[1]));
    }

    void ComputeStart() {
        recordGates = true;
        computeGates.clear();
        ++computeGen_;
    }
    void ComputeEnd() { recordGates = false; }
    void ComputeClear() {
        computeGates.clear();
        ++computeGen_;
    }

    // Compute and Uncompute replay the recorded gates; a recording of at
    // least 8 gates of <= 3 qubits runs as one compiled graph launch
    // (qcs_program_*), compiled on first use and again after the recording
    // changes. The gates replay as the reference's AppliedGate (generic
    // kernels, no traits).
    void Compute() {
        if (runProgram(false)) return;
        const bool recordSave = recordGates;
        recordGates = false;
        for (const auto& gate : computeGates) {
            if (gate.getQubitsNumber() > 3) ApplyOperatorMatrix(gate.getRawOperatorMatrix());
            else ApplyGate(gate);
        }
        recordGates = recordSave;
    }

    void Uncompute() {
        if (runProgram(true)) return;
        const bool recordSave = recordGates;
        recordGates = false;
        for (auto it = computeGates.crbegin(); it != computeGates.crend(); ++it) {
            if (it->getQubitsNumber() > 3) ApplyOperatorMatrix(it->getRawOperatorMatrix().adjoint());
            else {
                Gates::AppliedGate<MatrixClass> gate(it->getRawOperatorMatrix().adjoint(), it->getQubit1(),
                                                     it->getQubit2(), it->getQubit3());
                ApplyGate(gate);
            }
        }
        recordGates = recordSave;
    }

    double GetQubitProbability(size_t qubit) const {
        double p = 0;
        detail::qcs_check(qcs_qubit_probability(h_, uint32_t(qubit), &p));
        return p;
    }

    void SaveState() {
        if (!saved_) detail::qcs_check(qcs_create(uint32_t(NrQubits), QCS_COMPLEX128, &saved_));
        detail::qcs_check(qcs_copy_state(saved_, h_));
    }

    void RestoreState() {
        if (!saved_) return;
        detail::qcs_check(qcs_copy_state(h_, saved_));
        dirty_ = true;
    }

    void RestoreStateDestructive() {
        if (!saved_) return;
        detail::qcs_check(qcs_swap_state(h_, saved_));
        qcs_destroy(saved_);
        saved_ = nullptr;
        dirty_ = true;
    }

    size_t MeasureNoCollapse() {
        uint64_t out = 0;
        detail::qcs_check(qcs_measure_no_collapse(h_, 1, 0, 0, uniformZeroOne(rng), &out));
        return size_t(out);
    }

    // <psi|G|psi> for the product G of the gates (:646-660). A list of
    // single-qubit gates that are multiples of Pauli matrices (the special
    // case the reference notes at :649) takes one read-only FP64 pass; any
    // other list runs on a copy of the state held in a scratch register that
    // is made on first use and kept.
    std::complex<double> ExpectationValue(const std::vector<Gates::AppliedGate<MatrixClass>>& gates) {
        if (gates.empty()) return 1.;
        std::complex<double> coef;
        uint64_t x = 0, z = 0;
        if (detail::gates_as_pauli(gates, NrQubits, coef, x, z)) {
            double v = 0;
            detail::qcs_check(qcs_pauli_expectation(h_, 1, &x, &z, &v));
            return coef * v;
        }
        if (!shadow_) detail::qcs_check(qcs_create(uint32_t(NrQubits), QCS_COMPLEX128, &shadow_));
        detail::qcs_check(qcs_copy_state(shadow_, h_));
        ApplyGates(gates);
        double zz[2];
        detail::qcs_check(qcs_inner(shadow_, h_, zz));
        detail::qcs_check(qcs_swap_state(h_, shadow_));  // restore
        dirty_ = true;
        return {zz[0], zz[1]};
    }

    // ---- engine extensions: Pauli strings ----------------------------------
    // <psi|P|psi> for P = i^{|x & z|} X^x Z^z (qubit q: X for bit q of x only,
    // Z for z only, Y for both), FP64; many strings in one call share passes.
    std::vector<double> PauliExpectation(const std::vector<uint64_t>& xmask, const std::vector<uint64_t>& zmask) const {
        if (xmask.size() != zmask.size()) throw std::invalid_argument("mask vectors differ in length");
        std::vector<double> out(xmask.size());
        detail::qcs_check(
            qcs_pauli_expectation(h_, uint32_t(xmask.size()), xmask.data(), zmask.data(), out.data()));
        return out;
    }
    double PauliExpectation(uint64_t xmask, uint64_t zmask) const {
        double v = 0;
        detail::qcs_check(qcs_pauli_expectation(h_, 1, &xmask, &zmask, &v));
        return v;
    }

    std::unique_ptr<QubitRegister<VectorClass, MatrixClass>> Clone() const {
        auto qr = std::make_unique<QubitRegister<VectorClass, MatrixClass>>(NrQubits);
        detail::qcs_check(qcs_copy_state(qr->h_, h_));
        if (saved_) {
            qr->SaveState();
            detail::qcs_check(qcs_copy_state(qr->saved_, saved_));
        }
        qr->computeGates = computeGates;
        qr->recordGates = recordGates;
        qr->dirty_ = true;
        return qr;
    }

    // ---- engine extensions (no counterpart in the reference) -----------------
    // QFT / IQFT on [sq, eq] in fused passes; while gates are being recorded it
    // falls back to the reference's gate-by-gate sequence so Uncompute works.
    void ApplyQFT(size_t sq, size_t eq, bool doSwap, bool inverse) {
        if (recordGates) {
            qftSequence(sq, eq, doSwap, inverse);
            return;
        }
        detail::qcs_check(qcs_qft(h_, uint32_t(sq), uint32_t(std::min<size_t>(eq, UINT32_MAX)), doSwap, inverse));
        dirty_ = true;
    }

    qcs_register* engineHandle() const { return h_; }

protected:
    inline void CheckQubits(const GateClass& /*gate*/, size_t qubit, size_t controllingQubit1, size_t controllingQubit2,
                            size_t gateQubits) const {
        if (NrQubits == 0) throw std::invalid_argument("Qubit number is zero");
        else if (NrQubits <= qubit) throw std::invalid_argument("Qubit number is too high");
        else if (gateQubits == 2) {
            if (NrQubits <= controllingQubit1) throw std::invalid_argument("Controlling qubit number is too high");
            else if (qubit == controllingQubit1) throw std::invalid_argument("Qubit and controlling qubit are the same");
        } else if (gateQubits == 3) {
            if (NrQubits <= controllingQubit1 || NrQubits <= controllingQubit2)
                throw std::invalid_argument("Controlling qubit number is too high");
            else if (qubit == controllingQubit1 || qubit == controllingQubit2 || controllingQubit1 == controllingQubit2)
                throw std::invalid_argument("Qubits must be different");
        }
    }

    size_t MeasureNoCollapse(size_t qubit) { return MeasureNoCollapse(qubit, qubit); }

    size_t MeasureNoCollapse(size_t firstQubit, size_t secondQubit) {
        uint64_t out = 0;
        detail::qcs_check(qcs_measure_no_collapse(h_, 0, uint32_t(firstQubit), uint32_t(secondQubit),
                                                  uniformZeroOne(rng), &out));
        return size_t(out);
    }

    size_t NrQubits;
    size_t NrBasisStates;
    std::mt19937_64 rng;
    std::uniform_real_distribution<double> uniformZeroOne;
    std::vector<Gates::AppliedGate<MatrixClass>> computeGates;
    bool recordGates;

private:
    struct ProgramDeleter {
        void operator()(qcs_program* p) const { qcs_program_destroy(p); }
    };
    using ProgramPtr = std::unique_ptr<qcs_program, ProgramDeleter>;

    bool runProgram(bool adjoint) {
        if (computeGates.size() < 8) return false;
        for (const auto& g : computeGates)
            if (g.getQubitsNumber() > 3) return false;
        ProgramPtr& prog = adjoint ? progAdj_ : progFwd_;
        uint64_t& stamp = adjoint ? progAdjGen_ : progFwdGen_;
        if (!prog || stamp != computeGen_ || progSize_[adjoint] != computeGates.size()) {
            std::vector<qcs_gate> gates(computeGates.size());
            for (size_t i = 0; i < computeGates.size(); ++i) {
                const auto& g = computeGates[i];
                qcs_gate& c = gates[i];
                std::memset(&c, 0, sizeof c);
                c.k = uint32_t(g.getQubitsNumber());
                c.qubits[0] = uint32_t(g.getQubit1());
                c.qubits[1] = uint32_t(g.getQubit2());
                c.qubits[2] = uint32_t(g.getQubit3());
                const std::vector<double> m = detail::flatten(g.getRawOperatorMatrix());
                std::memcpy(c.mat, m.data(), m.size() * sizeof(double));
            }
            qcs_program* p = nullptr;
            detail::qcs_check(qcs_program_create(h_, gates.data(), uint32_t(gates.size()), 0, adjoint ? 1 : 0, &p));
            prog.reset(p);
            stamp = computeGen_;
            progSize_[adjoint] = computeGates.size();
        }
        detail::qcs_check(qcs_program_run(h_, prog.get()));
        dirty_ = true;
        return true;
    }

    void seed(unsigned int addseed) {
        if (addseed == 0) {
            std::random_device rd;
            addseed = rd();
        }
        const uint64_t timeSeed =
            uint64_t(std::chrono::high_resolution_clock::now().time_since_epoch().count()) + addseed;
        std::seed_seq s{uint32_t(timeSeed & 0xffffffff), uint32_t(timeSeed >> 32)};
        rng.seed(s);
    }

    void upload(const VectorClass& v) {
        detail::qcs_check(qcs_set_state(h_, reinterpret_cast<const double*>(v.data()), NrBasisStates));
        dirty_ = true;
    }

    template <class MapT> MapT sampleMap(bool all, size_t first, size_t second, size_t nrTimes) {
        MapT counts;
        if (nrTimes == 0) return counts;
        std::vector<double> u(nrTimes);
        for (auto& x : u) x = uniformZeroOne(rng);
        std::vector<uint64_t> out(nrTimes);
        detail::qcs_check(qcs_sample(h_, all ? 1 : 0, uint32_t(first), uint32_t(second), nrTimes, u.data(), out.data()));
        for (uint64_t v : out) ++counts[size_t(v)];
        return counts;
    }

    void qftSequence(size_t sq, size_t eq, bool doSwap, bool inverse) {
        Gates::HadamardGate<MatrixClass> had;
        Gates::ControlledPhaseShiftGate<MatrixClass> cps;
        Gates::SwapGate<MatrixClass> sw;
        eq = std::max(sq, std::min(NrQubits - 1, eq));
        auto swaps = [&]() {
            for (size_t a = sq, b = eq; a < b; ++a, --b) ApplyGate(sw, a, b);
        };
        const int s = int(sq), e = int(eq);
        if (!inverse) {
            ApplyGate(had, eq);
            for (int cur = e; cur > s; --cur) {
                double phase = M_PI_2;
                for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                    cps.SetPhaseShift(phase);
                    ApplyGate(cps, size_t(cur), size_t(ctrl));
                    phase *= 0.5;
                }
                ApplyGate(had, size_t(cur - 1));
            }
            if (doSwap) swaps();
        } else {
            if (doSwap) swaps();
            for (int cur = s + 1; cur <= e; ++cur) {
                ApplyGate(had, size_t(cur - 1));
                double phase = -M_PI_2;
                for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                    cps.SetPhaseShift(phase);
                    ApplyGate(cps, size_t(cur), size_t(ctrl));
                    phase *= 0.5;
                }
            }
            ApplyGate(had, eq);
        }
    }

    qcs_register* h_ = nullptr;
    qcs_register* saved_ = nullptr;
    qcs_register* shadow_ = nullptr;  // ExpectationValue's copy of the state
    mutable VectorClass mirror_;
    mutable bool dirty_ = true;
    uint64_t computeGen_ = 0;  // bumped whenever the recording changes
    ProgramPtr progFwd_, progAdj_;
    uint64_t progFwdGen_ = 0, progAdjGen_ = 0;
    size_t progSize_[2] = {0, 0};
};

}  // namespace QC
