// facade_tests.cpp - programs written against the reference's C++ API
// (QC::QubitRegister, QC::Gates, QC::SubAlgo::QuantumFourierTransform,
// QC::QuantumAlgorithm), compiled against include/QCSim/ and run on the GPU.
// Each test re-expresses a check the reference's own test suite makes; the
// reference file is named beside it. Exit status = number of failed tests.
#include <cmath>
#include <complex>
#include <cstdio>
#include <atomic>
#include <functional>
#include <future>
#include <iostream>
#include <random>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "IsingModel.h"
#include "QuantumAlgorithm.h"
#include "QuantumFourierTransform.h"
#include "QuantumGate.h"
#include "QubitRegister.h"

using cd = std::complex<double>;

static bool approx(cd a, cd b, double e = 1e-10) { return std::abs(a.real() - b.real()) < e && std::abs(a.imag() - b.imag()) < e; }

static void setRegister(QC::QubitRegister<>& reg) {
    // 1/2|00> - i/2|01> + 1/sqrt(2)|11>  (MeasurementsTests.cpp:34-43)
    reg.setToBasisState(0);
    reg.setRawAmplitude(0, 0.5);
    reg.setRawAmplitude(1, cd(0, -0.5));
    reg.setRawAmplitude(3, 1. / std::sqrt(2.));
    reg.Normalize();
}

// MeasurementsTests.cpp:45-167: post-measurement amplitudes for both qubits.
static bool measurementCollapse() {
    QC::QubitRegister reg(2);
    std::map<size_t, size_t> counts0, counts1;
    for (int i = 0; i < 2000; ++i) {
        setRegister(reg);
        const size_t s = reg.MeasureQubit(0);
        ++counts0[s];
        const Eigen::VectorXcd a = reg.getRegisterStorage();
        if (s == 1) {
            if (!approx(a(0), 0) || !approx(a(1), cd(0, -1. / std::sqrt(3.))) || !approx(a(2), 0) ||
                !approx(a(3), std::sqrt(2.) / std::sqrt(3.)))
                return false;
        } else if (!approx(a(0), 1) || !approx(a(1), 0) || !approx(a(2), 0) || !approx(a(3), 0))
            return false;
        setRegister(reg);
        const size_t t = reg.MeasureQubit(1);
        ++counts1[t];
        const Eigen::VectorXcd b = reg.getRegisterStorage();
        if (t == 1) {
            if (!approx(b(0), 0) || !approx(b(1), 0) || !approx(b(2), 0) || !approx(b(3), 1)) return false;
        } else if (!approx(b(0), 0.5 * std::sqrt(2.)) || !approx(b(1), cd(0, -0.5 * std::sqrt(2.))) ||
                   !approx(b(2), 0) || !approx(b(3), 0))
            return false;
    }
    // P(q0 = 1) = 3/4, P(q1 = 1) = 1/2 (MeasurementsTests.cpp:98, 161)
    return std::abs(counts0[1] / 2000. - 0.75) < 0.05 && std::abs(counts1[1] / 2000. - 0.5) < 0.05;
}

// Tests.cpp:665-723
static bool expectationValues() {
    QC::QubitRegister<> reg(3);
    if (!approx(reg.ExpectationValue({}), 1.0, 1e-6)) return false;
    QC::Gates::PauliZGate<> z;
    std::vector<QC::Gates::AppliedGate<>> z0{QC::Gates::AppliedGate<>(z.getRawOperatorMatrix(), 0)};
    if (!approx(reg.ExpectationValue(z0), 1., 1e-6)) return false;
    QC::Gates::PauliXGate<> x;
    reg.ApplyGate(x, 0);
    if (!approx(reg.ExpectationValue(z0), -1., 1e-6)) return false;
    QC::Gates::HadamardGate<> h;
    reg.ApplyGate(h, 0);
    if (!approx(reg.ExpectationValue(z0), 0., 1e-6)) return false;
    reg.setToBasisState(0);
    reg.ApplyGate(h, 1);
    std::vector<QC::Gates::AppliedGate<>> z1{QC::Gates::AppliedGate<>(z.getRawOperatorMatrix(), 1)};
    return approx(reg.ExpectationValue(z1), 0., 1e-6);
}

// GatesTests.cpp:11-225: every gate matrix is unitary.
static bool gatesUnitary() {
    auto unitary = [](const Eigen::MatrixXcd& m) {
        const Eigen::MatrixXcd r = m.adjoint() * m;
        for (Eigen::Index i = 0; i < r.rows(); ++i)
            for (Eigen::Index j = 0; j < r.cols(); ++j)
                if (!approx(r(i, j), i == j ? 1. : 0.)) return false;
        return true;
    };
    using namespace QC::Gates;
    std::vector<Eigen::MatrixXcd> ms = {
        HadamardGate<>().getRawOperatorMatrix(), HyGate<>().getRawOperatorMatrix(), SGate<>().getRawOperatorMatrix(),
        SDGGate<>().getRawOperatorMatrix(), TGate<>().getRawOperatorMatrix(), TDGGate<>().getRawOperatorMatrix(),
        PhaseShiftGate<>(0.3).getRawOperatorMatrix(), PauliXGate<>().getRawOperatorMatrix(),
        PauliYGate<>().getRawOperatorMatrix(), PauliZGate<>().getRawOperatorMatrix(),
        SquareRootNOTGate<>().getRawOperatorMatrix(), SquareRootNOTDagGate<>().getRawOperatorMatrix(),
        SplitterGate<>().getRawOperatorMatrix(), RxGate<>(0.7).getRawOperatorMatrix(), RyGate<>(0.7).getRawOperatorMatrix(),
        RzGate<>(0.7).getRawOperatorMatrix(), UGate<>(0.1, 0.2, 0.3, 0.4).getRawOperatorMatrix(),
        SwapGate<>().getRawOperatorMatrix(), iSwapGate<>().getRawOperatorMatrix(), iSwapDagGate<>().getRawOperatorMatrix(),
        DecrementGate<>().getRawOperatorMatrix(), CNOTGate<>().getRawOperatorMatrix(),
        ControlledYGate<>().getRawOperatorMatrix(), ControlledZGate<>().getRawOperatorMatrix(),
        ControlledHadamardGate<>().getRawOperatorMatrix(), ControlledSquareRootNOTGate<>().getRawOperatorMatrix(),
        ControlledSquareRootNOTDagGate<>().getRawOperatorMatrix(), ControlledPhaseGate<>().getRawOperatorMatrix(),
        ControlledPhaseShiftGate<>(0.4).getRawOperatorMatrix(), ControlledUGate<>(0.1, 0.2, 0.3, 0.4).getRawOperatorMatrix(),
        ControlledRxGate<>(0.5).getRawOperatorMatrix(), ControlledRyGate<>(0.5).getRawOperatorMatrix(),
        ControlledRzGate<>(0.5).getRawOperatorMatrix(), ToffoliGate<>().getRawOperatorMatrix(),
        FredkinGate<>().getRawOperatorMatrix(), CCZGate<>().getRawOperatorMatrix()};
    for (const auto& m : ms)
        if (!unitary(m)) return false;
    return true;
}

// Bell and GHZ states, CNOT / Toffoli truth tables (closed forms).
static bool entanglementAndTruthTables() {
    QC::QubitRegister<> reg(3);
    QC::Gates::HadamardGate<> h;
    QC::Gates::CNOTGate<> cnot;
    reg.ApplyGate(h, 0);
    reg.ApplyGate(cnot, 1, 0);
    reg.ApplyGate(cnot, 2, 1);
    const auto& v = reg.getRegisterStorage();
    const double s = 1. / std::sqrt(2.);
    for (size_t i = 0; i < 8; ++i)
        if (!approx(v(i), (i == 0 || i == 7) ? s : 0.)) return false;
    QC::Gates::ToffoliGate<> tof;
    for (size_t b = 0; b < 8; ++b) {
        reg.setToBasisState(b);
        reg.ApplyGate(tof, 0, 1, 2);
        const size_t want = ((b & 6) == 6) ? (b ^ 1) : b;
        if (!approx(reg.getBasisStateAmplitude(want), 1.)) return false;
    }
    return true;
}

// QFT = sqrt(N) * inverse DFT, IQFT undoes it; the gate-recorded QFT
// (while ComputeStart is active) uncomputes (QuantumFourierTransform.h:35-87).
static bool qftClosedForm() {
    const size_t n = 6, N = 64;
    Eigen::VectorXcd x(N);
    for (size_t i = 0; i < N; ++i) x(i) = cd(std::sin(0.3 * i + 1), std::cos(0.7 * i * i));
    QC::QubitRegister<> reg(n);
    reg.setRegisterStorage(x);
    const Eigen::VectorXcd in = reg.getRegisterStorage();  // normalised
    QC::SubAlgo::QuantumFourierTransform<> qft(n);
    qft.QFT(reg);
    const auto& out = reg.getRegisterStorage();
    for (size_t k = 0; k < N; ++k) {
        cd acc = 0;
        for (size_t j = 0; j < N; ++j) acc += in(j) * std::polar(1., 2 * M_PI * double(j * k % N) / N);
        if (!approx(out(k), acc / std::sqrt(double(N)), 1e-12)) return false;
    }
    qft.IQFT(reg);
    for (size_t k = 0; k < N; ++k)
        if (!approx(reg.getRegisterStorage()(k), in(k), 1e-12)) return false;
    reg.ComputeStart();
    qft.QFT(reg, false);
    reg.ComputeEnd();
    reg.Uncompute();
    for (size_t k = 0; k < N; ++k)
        if (!approx(reg.getRegisterStorage()(k), in(k), 1e-12)) return false;
    return true;
}

// QFT adder (the Draper construction exercised by AdderTests.cpp:213-327):
// |a>|b> -> |a>|a+b mod 2^m> through QFT on b, controlled phases from a's
// bits, IQFT. The outcome is deterministic.
static bool qftAdder() {
    const size_t m = 4;  // a on qubits 0..3, b on 4..7
    QC::Gates::ControlledPhaseShiftGate<> cps;
    QC::Gates::PauliXGate<> x;
    for (size_t a = 0; a < 16; a += 3)
        for (size_t b = 0; b < 16; b += 5) {
            QC::QubitRegister<> reg(2 * m);
            for (size_t i = 0; i < m; ++i) {
                if (a >> i & 1) reg.ApplyGate(x, i);
                if (b >> i & 1) reg.ApplyGate(x, m + i);
            }
            QC::SubAlgo::QuantumFourierTransform<> qft(2 * m, m, 2 * m - 1);
            qft.QFT(reg);  // b -> (1/sqrt N) sum_k e^{2 pi i b k / N} |k>
            for (size_t j = 0; j < m; ++j)      // qubit j of k
                for (size_t l = 0; l + j < m; ++l) {  // bit l of a
                    cps.SetPhaseShift(2 * M_PI * double(1ULL << (l + j)) / double(1ULL << m));
                    reg.ApplyGate(cps, m + j, l);
                }
            qft.IQFT(reg);
            const size_t got = reg.MeasureAll();
            if (got != (a | (((a + b) % 16) << m))) return false;
        }
    return true;
}

// Bernstein-Vazirani as a QuantumAlgorithm (the deterministic-answer tests of
// Tests.cpp:333-617): H on all, phase oracle (-1)^{s.x}, H on all -> s.
class BV : public QC::QuantumAlgorithm<> {
public:
    BV(size_t n, size_t s) : QC::QuantumAlgorithm<>(n), s_(s) {}
    size_t Execute() override {
        QC::Gates::HadamardGate<> h;
        QC::Gates::PauliZGate<> z;
        for (size_t q = 0; q < getNrQubits(); ++q) ApplyGate(h, q);
        for (size_t q = 0; q < getNrQubits(); ++q)
            if (s_ >> q & 1) ApplyGate(z, q);
        for (size_t q = 0; q < getNrQubits(); ++q) ApplyGate(h, q);
        return Measure();
    }

private:
    size_t s_;
};

static bool bernsteinVazirani() {
    for (size_t s : {size_t(0), size_t(5), size_t(1000), size_t(4095)}) {
        BV bv(12, s);
        if (bv.Execute() != s) return false;
    }
    return true;
}

// CheckQubits messages (QubitRegister.h:677-690).
static bool invalidArguments() {
    QC::QubitRegister<> reg(3);
    auto expect = [](const std::function<void()>& f, const std::string& msg) {
        try {
            f();
        } catch (const std::invalid_argument& e) {
            return msg == e.what();
        }
        return false;
    };
    QC::Gates::HadamardGate<> h;
    QC::Gates::CNOTGate<> cnot;
    QC::Gates::ToffoliGate<> tof;
    return expect([&] { reg.ApplyGate(h, 3); }, "Qubit number is too high") &&
           expect([&] { reg.ApplyGate(cnot, 0, 3); }, "Controlling qubit number is too high") &&
           expect([&] { reg.ApplyGate(cnot, 1, 1); }, "Qubit and controlling qubit are the same") &&
           expect([&] { reg.ApplyGate(tof, 0, 1, 1); }, "Qubits must be different");
}

// SaveState / RestoreState, Clone, swap semantics, setRegisterStorage.
static bool stateManagement() {
    QC::QubitRegister<> reg(4);
    QC::Gates::HadamardGate<> h;
    reg.ApplyGate(h, 2);
    reg.SaveState();
    reg.ApplyGate(h, 0);
    reg.RestoreState();
    if (!approx(reg.getBasisStateAmplitude(4), 1. / std::sqrt(2.))) return false;
    if (&reg.GetRegisterStorage() != &reg.getRegisterStorage()) return false;  // the alias
    auto c = reg.Clone();
    if (!approx(c->getBasisStateAmplitude(4), 1. / std::sqrt(2.))) return false;
    Eigen::VectorXcd v = Eigen::VectorXcd::Zero(16);
    v(3) = 1;
    (void)reg.getRegisterStorage();            // the host mirror holds the state now
    reg.setRegisterStorageFastNoNormalize(v);  // so v receives the old state
    if (!approx(reg.getBasisStateAmplitude(3), 1.) || v.size() != 16 || !approx(v(4), 1. / std::sqrt(2.)))
        return false;
    if (!approx(reg.getRegisterStorage()(3), 1.)) return false;  // mirror = the uploaded state
    {
        // after a mutation the mirror is stale: the vector comes back empty
        // instead of costing a download of the whole state
        QC::QubitRegister<> t(4);
        t.ApplyGate(h, 1);
        Eigen::VectorXcd u = Eigen::VectorXcd::Zero(16);
        u(5) = 1;
        t.setRegisterStorageFastNoNormalize(u);
        if (u.size() != 0 || !approx(t.getBasisStateAmplitude(5), 1.)) return false;
    }
    Eigen::VectorXcd w = Eigen::VectorXcd::Zero(16);
    w(7) = 2;
    QC::QubitRegister<> r2(4, w);  // the (N, v) ctor takes v's contents
    if (!approx(r2.getBasisStateAmplitude(7), 2.) || w.size() != 0) return false;
    reg.setRegisterStorage(v);  // normalises
    if (std::abs(reg.stateFidelity(v) - 1.) > 1e-12) return false;
    reg.setToEqualSuperposition();
    return std::abs(reg.GetQubitProbability(3) - 0.5) < 1e-12;
}

// RepeatedMeasure statistics and the diagonal ApplyOperatorMatrix path.
static bool samplingAndOperators() {
    QC::QubitRegister<> reg(2);
    QC::Gates::HadamardGate<> h;
    reg.ApplyGate(h, 0);
    auto counts = reg.RepeatedMeasure(20000);
    if (counts.size() != 2 || std::abs(double(counts[1]) / 20000 - 0.5) > 0.02) return false;
    auto r = reg.RepeatedMeasure(0, 0, 10);
    size_t tot = 0;
    for (auto& kv : r) tot += kv.second;
    if (tot != 10) return false;
    Eigen::MatrixXcd d = Eigen::MatrixXcd::Identity(4, 4);
    d(1, 1) = cd(0, 1);
    reg.ApplyOperatorMatrix(d);
    if (!approx(reg.getBasisStateAmplitude(1), cd(0, 1. / std::sqrt(2.)))) return false;
    Eigen::MatrixXcd p = Eigen::MatrixXcd::Zero(4, 4);  // permutation 0<->1
    p(0, 1) = p(1, 0) = p(2, 2) = p(3, 3) = 1;
    reg.ApplyOperatorMatrix(p);
    return approx(reg.getBasisStateAmplitude(0), cd(0, 1. / std::sqrt(2.)));
}

// One-qubit matrix on qubit q of a host vector (test helper).
static void hostApply1(Eigen::VectorXcd& v, const Eigen::MatrixXcd& m, size_t q) {
    const size_t bit = size_t(1) << q;
    for (size_t s = 0; s < size_t(v.size()); ++s) {
        if (s & bit) continue;
        const cd a = v(s), b = v(s | bit);
        v(s) = m(0, 0) * a + m(0, 1) * b;
        v(s | bit) = m(1, 0) * a + m(1, 1) * b;
    }
}

// ApplyOperatorMatrix with a dense operator (QubitRegister.h:499-505) on the
// device, and ExpectationValue's Pauli-product path (:646-660) against
// host arithmetic and against the copy path.
static bool operatorsAndPauli() {
    const size_t n = 6, N = size_t(1) << n;
    QC::QubitRegister<> reg(n);
    QC::Gates::RyGate<> ry;
    QC::Gates::CNOTGate<> cx;
    for (size_t q = 0; q < n; ++q) {
        ry.SetTheta(0.3 + 0.4 * double(q));
        reg.ApplyGate(ry, q);
    }
    for (size_t q = 0; q + 1 < n; ++q) reg.ApplyGate(cx, q + 1, q);
    const Eigen::VectorXcd v = reg.getRegisterStorage();
    Eigen::MatrixXcd m = Eigen::MatrixXcd::Zero(N, N);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) m(i, j) = cd(std::sin(0.37 * double(i * N + j)), std::cos(0.11 * double(i + 3 * j)));
    Eigen::VectorXcd want = Eigen::VectorXcd::Zero(N);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) want(i) += m(i, j) * v(j);
    reg.ApplyOperatorMatrix(m);
    const Eigen::VectorXcd got = reg.getRegisterStorage();
    for (size_t i = 0; i < N; ++i)
        if (!approx(got(i), want(i), 1e-11)) return false;
    reg.setRegisterStorage(v);

    QC::Gates::PauliXGate<> X;
    QC::Gates::PauliYGate<> Y;
    QC::Gates::PauliZGate<> Z;
    std::vector<QC::Gates::AppliedGate<>> gates{
        QC::Gates::AppliedGate<>(X.getRawOperatorMatrix(), 0), QC::Gates::AppliedGate<>(Z.getRawOperatorMatrix(), 2),
        QC::Gates::AppliedGate<>(Y.getRawOperatorMatrix(), 3), QC::Gates::AppliedGate<>(Z.getRawOperatorMatrix(), 0)};
    Eigen::VectorXcd w = v;
    for (const auto& g : gates) hostApply1(w, g.getRawOperatorMatrix(), g.getQubit1());
    cd ev = 0;
    for (size_t s = 0; s < N; ++s) ev += std::conj(v(s)) * w(s);
    const cd fast = reg.ExpectationValue(gates);
    if (!approx(fast, ev, 1e-12)) return false;
    // Z X on qubit 0 = i Y: the masks carry Y on qubits 0 and 3, Z on 2
    const double p = reg.PauliExpectation(0b1001, 0b1101);
    if (!approx(fast, cd(0, 1) * p, 1e-12)) return false;
    // a Hadamard in the list takes the copy path; the state is restored
    QC::Gates::HadamardGate<> H;
    gates.emplace_back(H.getRawOperatorMatrix(), 5);
    hostApply1(w, H.getRawOperatorMatrix(), 5);
    ev = 0;
    for (size_t s = 0; s < N; ++s) ev += std::conj(v(s)) * w(s);
    if (!approx(reg.ExpectationValue(gates), ev, 1e-12)) return false;
    const Eigen::VectorXcd after = reg.getRegisterStorage();
    for (size_t s = 0; s < N; ++s)
        if (!approx(after(s), v(s), 1e-14)) return false;
    return true;
}

// Compute / Uncompute (QubitRegister.h:536-590) repeated, with the recording
// changed in between: the replay (one compiled graph launch for >= 8 gates)
// follows the current recording.
static bool computeUncomputeReplay() {
    const size_t n = 10;
    QC::QubitRegister<> reg(n);
    QC::Gates::HadamardGate<> h;
    QC::Gates::CNOTGate<> cx;
    QC::Gates::RyGate<> ry;
    ry.SetTheta(0.3);
    const Eigen::VectorXcd in = reg.getRegisterStorage();
    auto same = [&](const Eigen::VectorXcd& a, const Eigen::VectorXcd& b) {
        for (size_t k = 0; k < (size_t(1) << n); ++k)
            if (!approx(a(k), b(k), 1e-12)) return false;
        return true;
    };
    reg.ComputeStart();
    for (size_t q = 0; q < n; ++q) reg.ApplyGate(h, q);
    for (size_t q = 0; q + 1 < n; ++q) reg.ApplyGate(cx, q + 1, q);
    reg.ComputeEnd();
    const Eigen::VectorXcd once = reg.getRegisterStorage();
    for (int rep = 0; rep < 3; ++rep) {
        reg.Uncompute();
        if (!same(reg.getRegisterStorage(), in)) return false;
        reg.Compute();
        if (!same(reg.getRegisterStorage(), once)) return false;
    }
    reg.Uncompute();
    reg.ComputeStart();
    for (size_t q = 0; q < n; ++q) reg.ApplyGate(ry, q);
    reg.ComputeEnd();
    reg.Uncompute();
    if (!same(reg.getRegisterStorage(), in)) return false;
    reg.Compute();
    const double c = std::cos(0.15), s = std::sin(0.15);
    return approx(reg.getBasisStateAmplitude(0), std::pow(c, double(n)), 1e-12) &&
           approx(reg.getBasisStateAmplitude((size_t(1) << n) - 1), std::pow(s, double(n)), 1e-12);
}

// IsingTests.cpp:6-42 (IsingModelTest): 3 spins with fields, p = 3 and the
// extra mixing; the state QAOA picks must be the minimum-energy state.
static bool isingModelQaoa() {
    QC::QubitRegister<> reg(3);
    Models::QAOAIsingSubalgorithm<> ising;
    ising.Set({{0, 1, 1.5}, {1, 2, 2}, {0, 2, -3}}, {1.25, 3.25, -2.5});
    ising.SetP(3);
    ising.SetGammaStart(1);
    ising.SetBetaStart(3);
    ising.SetGammaStart(0.3, 1);
    ising.SetBetaStart(1, 1);
    ising.SetMixing(true);
    const size_t state = ising.Execute(reg);
    return state == ising.GetMinEnergyState() && ising.GetExactEvaluations() > 0;
}

// IsingTests.cpp:45-93 (MaxCutTest): 6 nodes, 8 edges with J = -1, p = 2
// from gamma = (6, 4.2), beta = (0.7, 0.5); the state QAOA picks must be one
// of the minimum-energy states. The optimiser runs on the batched engine.
static bool maxCutQaoa() {
    QC::QubitRegister<> reg(6);
    Models::QAOAIsingSubalgorithm<> maxcut;
    maxcut.Set({{0, 1, -1}, {0, 2, -1}, {0, 5, -1}, {1, 2, -1}, {1, 3, -1}, {2, 3, -1}, {2, 4, -1}, {3, 5, -1}},
               {0, 0, 0, 0, 0, 0});
    const auto minStates = maxcut.GetMinEnergyStates();
    maxcut.SetP(2);
    maxcut.SetGammaStart(6);
    maxcut.SetBetaStart(0.7);
    maxcut.SetGammaStart(4.2, 1);
    maxcut.SetBetaStart(0.5, 1);
    const size_t state = maxcut.Execute(reg);
    return minStates.count(state) == 1 && maxcut.GetExactEvaluations() % 9 == 0;
}

// A QuantumAlgorithm subclass used the way SimulationTests.cpp:66-79 uses
// one: an exact vector from a host operator, the static
// AdjustPhaseAndNormalize on it, the member one on the register, then
// stateFidelity; plus writeToFile / displayRegister (QuantumAlgorithm.h:143-172).
class RotationAlgorithm : public QC::QuantumAlgorithm<> {
public:
    explicit RotationAlgorithm(size_t n) : QC::QuantumAlgorithm<>(n) {}
    size_t Execute() override {
        QC::Gates::RyGate<> ry;
        QC::Gates::RzGate<> rz;
        QC::Gates::CNOTGate<> cx;
        for (size_t q = 0; q < getNrQubits(); ++q) {
            ry.SetTheta(0.4 + 0.3 * double(q));
            ApplyGate(ry, q);
            rz.SetTheta(0.7 - 0.2 * double(q));
            ApplyGate(rz, q);
        }
        for (size_t q = 0; q + 1 < getNrQubits(); ++q) ApplyGate(cx, q + 1, q);
        return 0;
    }
};

static bool algorithmInterface() {
    const size_t n = 3, N = size_t(1) << n;
    RotationAlgorithm sim(n);
    // exact result: the same product built as a dense operator on the host
    auto kron1 = [&](const Eigen::MatrixXcd& g, size_t q) {
        Eigen::MatrixXcd m = Eigen::MatrixXcd::Zero(N, N);
        for (size_t r = 0; r < N; ++r)
            for (size_t c = 0; c < N; ++c) {
                if ((r ^ c) & ~(size_t(1) << q)) continue;
                m(r, c) = g((r >> q) & 1, (c >> q) & 1);
            }
        return m;
    };
    Eigen::VectorXcd ex = Eigen::VectorXcd::Zero(N);
    ex(0) = 1;
    QC::Gates::RyGate<> ry;
    QC::Gates::RzGate<> rz;
    for (size_t q = 0; q < n; ++q) {
        ry.SetTheta(0.4 + 0.3 * double(q));
        ex = kron1(ry.getRawOperatorMatrix(), q) * ex;
        rz.SetTheta(0.7 - 0.2 * double(q));
        ex = kron1(rz.getRawOperatorMatrix(), q) * ex;
    }
    for (size_t q = 0; q + 1 < n; ++q) {  // CNOT: target q + 1, control q
        Eigen::VectorXcd t = ex;
        for (size_t s = 0; s < N; ++s)
            if (s >> q & 1) t(s) = ex(s ^ (size_t(1) << (q + 1)));
        ex = t;
    }
    // a global phase and scale the normalisation must remove
    for (size_t s = 0; s < N; ++s) ex(s) *= cd(0.0, 3.0);
    sim.AdjustPhaseAndNormalize(ex);  // static, called through the object
    if (std::abs(ex(0).imag()) > 1e-15 || ex(0).real() <= 0) return false;
    double nrm = 0;
    for (size_t s = 0; s < N; ++s) nrm += std::norm(ex(s));
    if (std::abs(nrm - 1) > 1e-14) return false;

    sim.setToBasisState(0);
    sim.Execute();
    sim.AdjustPhaseAndNormalize();
    const double fidelity = sim.stateFidelity(ex);
    if (std::abs(fidelity - 1) > 1e-12) return false;
    const Eigen::VectorXcd got = sim.getRegisterStorage();
    for (size_t s = 0; s < N; ++s)
        if (!approx(got(s), ex(s), 1e-12)) return false;
    // fidelity against an orthogonal state
    Eigen::VectorXcd orth = Eigen::VectorXcd::Zero(N);
    orth(0) = -std::conj(ex(1));
    orth(1) = std::conj(ex(0));
    if (sim.stateFidelity(orth) > 1e-24) return false;

    // text dump: one "index re im" (or "index probability") line per state
    const std::string path = "/tmp/qcs_facade_dump.txt";
    if (!sim.writeToFile(path)) return false;
    if (!sim.writeToFile(path, false, true)) return false;  // appended
    std::FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    size_t idx = 0, lines = 0;
    double re = 0, im = 0;
    bool ok = true;
    for (; lines < N && std::fscanf(f, "%zu %lf %lf", &idx, &re, &im) == 3; ++lines)
        ok = ok && idx == lines && approx(cd(re, im), got(lines), 1e-15);
    for (size_t k = 0; k < N && std::fscanf(f, "%zu %lf", &idx, &re) == 2; ++k, ++lines)
        ok = ok && idx == k && std::abs(re - std::norm(got(k))) < 1e-15;
    std::fclose(f);
    std::remove(path.c_str());
    if (!ok || lines != 2 * N) return false;
    sim.displayRegister();
    return true;
}

// Compute / Uncompute of a 17-qubit QFT recorded gate by gate: 136
// controlled phases and 8 swaps replay as generic 4x4 blocks, each staging
// its matrix in the compiled program's parameter memory.
static bool recordedQftUncompute() {
    const size_t n = 17;
    QC::QubitRegister<> reg(n);
    QC::Gates::HadamardGate<> h;
    QC::Gates::RyGate<> ry;
    for (size_t q = 0; q < n; ++q) {
        ry.SetTheta(0.1 + 0.05 * double(q));
        reg.ApplyGate(ry, q);
    }
    const Eigen::VectorXcd in = reg.getRegisterStorage();
    QC::SubAlgo::QuantumFourierTransform<> qft(n);
    reg.ComputeStart();
    qft.QFT(reg);
    reg.ComputeEnd();
    const Eigen::VectorXcd once = reg.getRegisterStorage();
    for (int rep = 0; rep < 2; ++rep) {
        reg.Uncompute();
        const Eigen::VectorXcd back = reg.getRegisterStorage();
        for (size_t s = 0; s < (size_t(1) << n); ++s)
            if (!approx(back(s), in(s), 1e-12)) return false;
        reg.Compute();
        const Eigen::VectorXcd again = reg.getRegisterStorage();
        for (size_t s = 0; s < (size_t(1) << n); ++s)
            if (!approx(again(s), once(s), 1e-12)) return false;
    }
    // the recorded sequence is the QFT: compare with the fused passes
    QC::QubitRegister<> fused(n);
    for (size_t q = 0; q < n; ++q) {
        ry.SetTheta(0.1 + 0.05 * double(q));
        fused.ApplyGate(ry, q);
    }
    qft.QFT(fused);
    const Eigen::VectorXcd f = fused.getRegisterStorage();
    for (size_t s = 0; s < (size_t(1) << n); ++s)
        if (!approx(f(s), once(s), 1e-12)) return false;
    return true;
}

// CliffordTests.cpp:208-258: registers created, driven and measured on many
// std::async threads at once (one register per shot), each applying the same
// random Clifford+T circuit. Every thread's state before the measurements
// must equal the one computed alone, and measuring every qubit must leave a
// basis state.
static bool concurrentRegisters() {
    const size_t n = 10, ngates = 200;
    std::mt19937_64 gen(5);
    std::uniform_int_distribution<int> kind(0, 7);
    std::uniform_int_distribution<size_t> qd(0, n - 1);
    std::vector<int> gk(ngates);
    std::vector<size_t> q1(ngates), q2(ngates);
    for (size_t j = 0; j < ngates; ++j) {
        gk[j] = kind(gen);
        q1[j] = qd(gen);
        q2[j] = (q1[j] + 1 + qd(gen) % (n - 1)) % n;
    }
    auto run = [&](QC::QubitRegister<>& reg) {
        QC::Gates::HadamardGate<> h;
        QC::Gates::PauliXGate<> x;
        QC::Gates::PauliYGate<> y;
        QC::Gates::PauliZGate<> z;
        QC::Gates::SGate<> sg;
        QC::Gates::TGate<> t;
        QC::Gates::CNOTGate<> cx;
        QC::Gates::ControlledZGate<> cz;
        for (size_t j = 0; j < ngates; ++j) {
            switch (gk[j]) {
                case 0: reg.ApplyGate(h, q1[j]); break;
                case 1: reg.ApplyGate(x, q1[j]); break;
                case 2: reg.ApplyGate(y, q1[j]); break;
                case 3: reg.ApplyGate(z, q1[j]); break;
                case 4: reg.ApplyGate(sg, q1[j]); break;
                case 5: reg.ApplyGate(t, q1[j]); break;
                case 6: reg.ApplyGate(cx, q1[j], q2[j]); break;
                default: reg.ApplyGate(cz, q1[j], q2[j]); break;
            }
        }
    };
    QC::QubitRegister<> alone(n);
    run(alone);
    const Eigen::VectorXcd want = alone.getRegisterStorage();
    std::atomic<int> bad{0};
    std::vector<std::future<void>> tasks;
    for (int th = 0; th < 8; ++th)
        tasks.push_back(std::async(std::launch::async, [&]() {
            for (int shot = 0; shot < 4; ++shot) {
                QC::QubitRegister<> reg(n);
                run(reg);
                const Eigen::VectorXcd got = reg.getRegisterStorage();
                for (size_t i = 0; i < (size_t(1) << n); ++i)
                    if (!approx(got(i), want(i), 1e-13)) ++bad;
                size_t v = 0;
                for (size_t q = 0; q < n; ++q)
                    if (reg.MeasureQubit(q)) v |= size_t(1) << q;
                const Eigen::VectorXcd post = reg.getRegisterStorage();
                if (std::abs(std::abs(post(v)) - 1.0) > 1e-12) ++bad;
            }
        }));
    for (auto& t : tasks) t.get();
    return bad.load() == 0;
}

int main() {
    const std::vector<std::pair<const char*, bool (*)()>> tests = {
        {"measurementCollapse", measurementCollapse},
        {"expectationValues", expectationValues},
        {"gatesUnitary", gatesUnitary},
        {"entanglementAndTruthTables", entanglementAndTruthTables},
        {"qftClosedForm", qftClosedForm},
        {"qftAdder", qftAdder},
        {"bernsteinVazirani", bernsteinVazirani},
        {"invalidArguments", invalidArguments},
        {"stateManagement", stateManagement},
        {"samplingAndOperators", samplingAndOperators},
        {"computeUncomputeReplay", computeUncomputeReplay},
        {"isingModelQaoa", isingModelQaoa},
        {"maxCutQaoa", maxCutQaoa},
        {"algorithmInterface", algorithmInterface},
        {"operatorsAndPauli", operatorsAndPauli},
        {"recordedQftUncompute", recordedQftUncompute},
        {"concurrentRegisters", concurrentRegisters},
    };
    int failed = 0;
    for (const auto& t : tests) {
        std::cout << t.first << " ... " << std::flush;
        bool ok = false;
        try {
            ok = t.second();
        } catch (const std::exception& e) {
            std::cout << "exception: " << e.what() << " ";
        }
        std::cout << (ok ? "PASS" : "FAIL") << std::endl;
        failed += ok ? 0 : 1;
    }
    std::cout << (failed ? "FAILED " : "ALL PASSED ") << failed << std::endl;
    return failed;
}
