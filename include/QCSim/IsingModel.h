// IsingModel.h - the Ising model and its QAOA subalgorithm with the QCSim API
// (reference: QCSim/IsingModel.h:19-687), the optimiser's energies computed
// exactly by the batched engine (qcs_qaoa_*) instead of 100000-shot
// estimates.
//
// IsingModel keeps the reference's data model: spins (bit 1 = spin -1), site
// fields h, one neighbour set per site holding the larger index of each
// edge, E = -(sum h z + sum J z z) (IsingModel.h:28-212).
//
// QAOAIsingSubalgorithm::Execute runs the reference's loop (IsingModel.h:
// 220-280): GradientDescentStep (central differences with epsilon, ADAM-W
// with beta1, beta2, lambda and stepSize, :580-634) then the energy at the
// new angles, until |E - Eold| <= deltaE; each step's 4p shifted circuits
// are one batched evaluation. It then runs the circuit with the best angles
// on the register, gate by gate as Exec (:557-567) does, draws
// nrMeasurements samples and returns the sampled state with the lowest
// E(state) * count / nrMeasurements. With the extra mixing (SetMixing(true),
// :538-553), which the batched engine does not run, the optimiser's circuits
// run gate by gate on a scratch register, with exact FP64 energies. Unlike
// the reference it prints nothing.
#pragma once

#include <cmath>
#include <limits>
#include <memory>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "QuantumAlgorithm.h"

namespace Models {

template <typename type1 = size_t, typename type2 = size_t> class PairHash {
public:
    size_t operator()(const std::pair<type1, type2>& t) const {
        return std::hash<type1>()(t.first) * 31 + std::hash<type2>()(t.second);
    }
};

class IsingModel {
public:
    struct Interaction {
        size_t i = 0;
        size_t j = 0;
        double J = 1.;
    };

    void Clear() {
        spins.clear();
        h.clear();
        neighbours.clear();
        interactions.clear();
    }

    // self-interactions and sites beyond the spin count are ignored
    void AddInteraction(size_t i, size_t j, double J) {
        if (i == j) return;
        if (i > j) std::swap(i, j);
        if (j >= spins.size()) return;
        neighbours[i].insert(j);
        interactions[std::make_pair(i, j)] = J;
    }

    void Set(const std::vector<bool>& s, const std::vector<Interaction>& ints = {},
             const std::vector<double>& hvals = {}) {
        Clear();
        spins = s;
        h = hvals;
        h.resize(s.size(), 0.);
        for (const auto& it : ints) AddInteraction(it.i, it.j, it.J);
    }

    void Set(const std::vector<Interaction>& ints = {}, const std::vector<double>& hvals = {}) {
        Clear();
        h = hvals;
        spins.resize(h.size(), false);
        for (const auto& it : ints) AddInteraction(it.i, it.j, it.J);
    }

    void SetH(size_t i, double hval) {
        if (i < spins.size()) h[i] = hval;
    }
    double GetH(size_t i) const { return i < h.size() ? h[i] : 0.; }
    void SetSpin(size_t i, bool s) {
        if (i < spins.size()) spins[i] = s;
    }
    bool GetSpin(size_t i) const { return i < spins.size() ? bool(spins[i]) : false; }

    double Energy() {
        double e = 0;
        for (size_t i = 0; i < spins.size(); ++i) {
            const int zi = GetTrueSpin(spins[i]);
            e += GetH(i) * zi;
            const auto it = neighbours.find(i);
            if (it == neighbours.end()) continue;
            for (size_t j : it->second) e += interactions[std::make_pair(i, j)] * zi * GetTrueSpin(spins[j]);
        }
        return -e;
    }

    void SetState(size_t state) {
        for (size_t pos = 0; pos < spins.size(); ++pos, state >>= 1) spins[pos] = (state & 1) == 1;
    }

    double Energy(size_t state) {
        SetState(state);
        return Energy();
    }

    // first state of minimal energy (exhaustive)
    size_t GetMinEnergyState() {
        const size_t n = size_t(1) << spins.size();
        double best = std::numeric_limits<double>::max();
        size_t arg = 0;
        for (size_t s = 0; s < n; ++s) {
            const double e = Energy(s);
            if (best > e) {
                best = e;
                arg = s;
            }
        }
        return arg;
    }

    // every state within 1e-9 of the minimal energy (exhaustive)
    std::set<size_t> GetMinEnergyStates() {
        std::set<size_t> res;
        const size_t n = size_t(1) << spins.size();
        double best = std::numeric_limits<double>::max();
        for (size_t s = 0; s < n; ++s) {
            const double e = Energy(s);
            if (best > e && std::abs(best - e) > 1E-9) {
                best = e;
                res.clear();
                res.insert(s);
            } else if (std::abs(best - e) <= 1E-9) {
                res.insert(s);
            }
        }
        return res;
    }

    const std::vector<bool>& GetSpins() const { return spins; }
    const std::vector<double>& GetH() const { return h; }
    const std::unordered_map<size_t, std::unordered_set<size_t>>& GetNeighbours() const { return neighbours; }
    const std::unordered_map<std::pair<size_t, size_t>, double, PairHash<>>& GetInteractions() const {
        return interactions;
    }

protected:
    static int GetTrueSpin(bool s) { return s ? -1 : 1; }

    std::vector<bool> spins;
    std::vector<double> h;
    std::unordered_map<size_t, std::unordered_set<size_t>> neighbours;
    std::unordered_map<std::pair<size_t, size_t>, double, PairHash<>> interactions;
};

template <class VectorClass = Eigen::VectorXcd, class MatrixClass = Eigen::MatrixXcd>
class QAOAIsingSubalgorithm : public QC::QuantumSubAlgorithm<VectorClass, MatrixClass> {
public:
    using RegisterClass = QC::QubitRegister<VectorClass, MatrixClass>;

    size_t Execute(RegisterClass& reg) override {
        double Eold = std::numeric_limits<double>::max();
        double Emin = Eold;
        double E = 0;
        std::vector<double> gamma = gammaStart, beta = betaStart;
        std::vector<double> optGamma = gamma, optBeta = beta;
        const size_t p = gamma.size();
        sGamma.assign(p, 0.);
        mGamma.assign(p, 0.);
        sBeta.assign(p, 0.);
        mBeta.assign(p, 0.);
        stepNr = 0;
        Batches b = MakeBatches(reg, p);
        for (int i = 0; std::abs(E - Eold) > deltaE && i < 100000; ++i) {
            Eold = E;
            auto gb = GradientDescentStep(b, gamma, beta, epsilon, stepSize);
            gamma = gb.first;
            beta = gb.second;
            E = Evaluate(b, false, 1, p, gamma, beta)[0];
            if (E < Emin) {
                Emin = E;
                optGamma = gamma;
                optBeta = beta;
            }
        }
        Exec(reg, optGamma, optBeta);
        const auto res = reg.RepeatedMeasure(nrMeasurements);
        Emin = std::numeric_limits<double>::max();
        size_t state = 0;
        for (const auto& r : res) {
            const double e = Energy(r.first) * static_cast<double>(r.second) / double(nrMeasurements);
            if (e < Emin) {
                Emin = e;
                state = r.first;
            }
        }
        sGamma.clear();
        mGamma.clear();
        sBeta.clear();
        mBeta.clear();
        return state;
    }

    void Clear() { model.Clear(); }
    void AddInteraction(size_t i, size_t j, double J) { model.AddInteraction(i, j, J); }
    void Set(const std::vector<bool>& s, const std::vector<IsingModel::Interaction>& ints = {},
             const std::vector<double>& hvals = {}) {
        model.Set(s, ints, hvals);
    }
    void Set(const std::vector<IsingModel::Interaction>& ints = {}, const std::vector<double>& hvals = {}) {
        model.Set(ints, hvals);
    }
    void SetH(size_t i, double hval) { model.SetH(i, hval); }
    double GetH(size_t i) const { return model.GetH(i); }
    void SetSpin(size_t i, bool s) { model.SetSpin(i, s); }
    bool GetSpin(size_t i) const { return model.GetSpin(i); }
    double Energy() { return model.Energy(); }
    void SetState(size_t state) { model.SetState(state); }
    double Energy(size_t state) { return model.Energy(state); }
    size_t GetMinEnergyState() { return model.GetMinEnergyState(); }
    std::set<size_t> GetMinEnergyStates() { return model.GetMinEnergyStates(); }
    const std::vector<bool>& GetSpins() const { return model.GetSpins(); }
    const std::vector<double>& GetH() const { return model.GetH(); }
    const std::unordered_map<size_t, std::unordered_set<size_t>>& GetNeighbours() const {
        return model.GetNeighbours();
    }
    const std::unordered_map<std::pair<size_t, size_t>, double, PairHash<>>& GetInteractions() const {
        return model.GetInteractions();
    }

    double GetDeltaE() const { return deltaE; }
    void SetDeltaE(double dE) { deltaE = dE; }
    const std::vector<double>& GetGammaStart() const { return gammaStart; }
    void SetGammaStart(double gamma, size_t index = 0) {
        if (index < gammaStart.size()) gammaStart[index] = gamma;
    }
    const std::vector<double>& GetBetaStart() const { return betaStart; }
    void SetBetaStart(double beta, size_t index = 0) {
        if (index < betaStart.size()) betaStart[index] = beta;
    }
    size_t GetP() const { return gammaStart.size(); }
    void SetP(size_t p) {
        if (p < 1) return;
        gammaStart.resize(p, 1.);
        betaStart.resize(p, 1.);
    }
    double GetEpsilon() const { return epsilon; }
    void SetEpsilon(double eps) { epsilon = eps; }
    double GetStepSize() const { return stepSize; }
    void SetStepSize(double step) { stepSize = step; }
    void SetBeta1(double b1) { beta1 = b1; }
    double GetBeta1() const { return beta1; }
    void SetBeta2(double b2) { beta2 = b2; }
    double GetBeta2() const { return beta2; }
    void SetLambda(double l) { lambda = l; }
    double GetLambda() const { return lambda; }
    double GetNrMeasurements() const { return double(nrMeasurements); }
    void SetNrMeasurements(double nr) { nrMeasurements = size_t(nr); }
    void SetMixing(bool better) { betterMixing = better; }
    bool GetMixing() const { return betterMixing; }

    // Exact evaluations spent by the last Execute (4p + 1 per step).
    size_t GetExactEvaluations() const { return evaluations; }

protected:
    struct BatchDeleter {
        void operator()(qcs_qaoa* q) const { qcs_qaoa_destroy(q); }
    };
    using BatchPtr = std::unique_ptr<qcs_qaoa, BatchDeleter>;
    // The batched engine runs the plain mixer; with the extra mixing the
    // circuits run gate by gate on a scratch register instead.
    struct Batches {
        BatchPtr shifted;  // 4p copies of the instance
        BatchPtr single;
        std::unique_ptr<RegisterClass> scratch;
    };

    // Edge list of the model for the engine (the order of commuting cost
    // terms does not matter).
    void EdgeList(std::vector<uint32_t>& ei, std::vector<uint32_t>& ej, std::vector<double>& J) const {
        for (const auto& site : model.GetNeighbours())
            for (size_t j : site.second) {
                ei.push_back(uint32_t(site.first));
                ej.push_back(uint32_t(j));
                J.push_back(model.GetInteractions().at(std::make_pair(site.first, j)));
            }
    }

    Batches MakeBatches(const RegisterClass& reg, size_t p) {
        const size_t nq = reg.getNrQubits();
        evaluations = 0;
        if (betterMixing) {
            Batches b;
            b.scratch.reset(new RegisterClass(nq));
            return b;
        }
        std::vector<uint32_t> ei, ej;
        std::vector<double> J, h(nq);
        EdgeList(ei, ej, J);
        for (size_t i = 0; i < nq; ++i) h[i] = model.GetH(i);
        auto make = [&](uint32_t copies) {
            std::vector<uint32_t> nqv(copies, uint32_t(nq)), nev(copies, uint32_t(ei.size())), EI, EJ;
            std::vector<double> JJ, HH;
            for (uint32_t c = 0; c < copies; ++c) {
                EI.insert(EI.end(), ei.begin(), ei.end());
                EJ.insert(EJ.end(), ej.begin(), ej.end());
                JJ.insert(JJ.end(), J.begin(), J.end());
                HH.insert(HH.end(), h.begin(), h.end());
            }
            qcs_qaoa* q = nullptr;
            QC::detail::qcs_check(qcs_qaoa_create(copies, nqv.data(), nev.data(), EI.data(), EJ.data(), JJ.data(),
                                                  HH.data(), uint32_t(p), &q));
            return BatchPtr(q);
        };
        Batches b;
        b.shifted = make(uint32_t(4 * p));
        b.single = make(1);
        return b;
    }

    // Exact energies of `copies` angle sets (rows of G and B, p each).
    std::vector<double> Evaluate(Batches& b, bool shifted, size_t copies, size_t p, const std::vector<double>& G,
                                 const std::vector<double>& B) {
        std::vector<double> e(copies);
        evaluations += copies;
        if (b.scratch) {
            for (size_t c = 0; c < copies; ++c) {
                const std::vector<double> g(G.begin() + c * p, G.begin() + (c + 1) * p);
                const std::vector<double> bt(B.begin() + c * p, B.begin() + (c + 1) * p);
                Exec(*b.scratch, g, bt);
                e[c] = EnergyExpectationValue(*b.scratch);
            }
            return e;
        }
        QC::detail::qcs_check(
            qcs_qaoa_evaluate(shifted ? b.shifted.get() : b.single.get(), G.data(), B.data(), e.data()));
        return e;
    }

    // The circuit of the reference's Exec on the register: H on all qubits;
    // per layer CNOT, Rz(2 gamma J), CNOT per edge, Rz(gamma h) per site
    // with h != 0, Rx(2 beta) per qubit (and the optional extra mixing).
    void Exec(RegisterClass& reg, const std::vector<double>& gamma, const std::vector<double>& beta) {
        reg.setToBasisState(0);
        const size_t nq = reg.getNrQubits();
        for (size_t q = 0; q < nq; ++q) reg.ApplyGate(hgate, q);
        for (size_t i = 0; i < gamma.size(); ++i) {
            for (const auto& site : model.GetNeighbours())
                for (size_t j : site.second) {
                    rz.SetTheta(2 * gamma[i] * model.GetInteractions().at(std::make_pair(site.first, j)));
                    reg.ApplyGate(cnot, site.first, j);
                    reg.ApplyGate(rz, site.first);
                    reg.ApplyGate(cnot, site.first, j);
                }
            for (size_t q = 0; q < nq; ++q) {
                const double hv = model.GetH(q);
                if (hv == 0.) continue;
                rz.SetTheta(gamma[i] * hv);
                reg.ApplyGate(rz, q);
            }
            rx.SetTheta(2 * beta[i]);
            for (size_t q = 0; q < nq; ++q) reg.ApplyGate(rx, q);
            if (betterMixing) {
                ry.SetTheta(2 * beta[i]);
                const size_t last = nq - 1;
                for (size_t q = 0; q < last; ++q) reg.ApplyGate(cnot, q + 1, q);
                reg.ApplyGate(cnot, 0, last);
                for (size_t q = 0; q < nq; ++q) reg.ApplyGate(ry, q);
            }
        }
    }

    // Exact <H> of the register's state (FP64 reduction on the device).
    double EnergyExpectationValue(RegisterClass& reg) {
        std::vector<uint32_t> ei, ej;
        std::vector<double> J, h(reg.getNrQubits());
        EdgeList(ei, ej, J);
        for (size_t i = 0; i < h.size(); ++i) h[i] = model.GetH(i);
        double e = 0;
        QC::detail::qcs_check(
            qcs_ising_energy(reg.engineHandle(), uint32_t(ei.size()), ei.data(), ej.data(), J.data(), h.data(), &e));
        return e;
    }

    // One ADAM-W step from central differences (IsingModel.h:580-634); the
    // 4p shifted circuits are evaluated together.
    std::pair<std::vector<double>, std::vector<double>> GradientDescentStep(Batches& b, const std::vector<double>& gamma,
                                                                            const std::vector<double>& beta,
                                                                            double eps, double alpha) {
        const size_t p = gamma.size();
        const double twoEps = 2. * eps;
        std::vector<double> G(4 * p * p), B(4 * p * p);
        for (size_t c = 0; c < 4 * p; ++c)
            for (size_t l = 0; l < p; ++l) {
                G[c * p + l] = gamma[l];
                B[c * p + l] = beta[l];
            }
        for (size_t i = 0; i < p; ++i) {
            G[(4 * i) * p + i] -= eps;
            G[(4 * i + 1) * p + i] += eps;
            B[(4 * i + 2) * p + i] -= eps;
            B[(4 * i + 3) * p + i] += eps;
        }
        const std::vector<double> E = Evaluate(b, true, 4 * p, p, G, B);
        ++stepNr;
        const double div1 = 1. / (1. - std::pow(beta1, stepNr));
        const double div2 = 1. / (1. - std::pow(beta2, stepNr));
        std::pair<std::vector<double>, std::vector<double>> res;
        res.first.resize(p);
        res.second.resize(p);
        for (size_t i = 0; i < p; ++i) {
            double grad = (E[4 * i + 1] - E[4 * i]) / twoEps;
            mGamma[i] = (beta1 * mGamma[i] - (1. - beta1) * grad) * div1;
            sGamma[i] = (beta2 * sGamma[i] + (1. - beta2) * grad * grad) * div2;
            res.first[i] = gamma[i] + alpha * (mGamma[i] / std::sqrt(sGamma[i] + 1E-8) - lambda * gamma[i]);
            grad = (E[4 * i + 3] - E[4 * i + 2]) / twoEps;
            mBeta[i] = (beta1 * mBeta[i] - (1. - beta1) * grad) * div1;
            sBeta[i] = (beta2 * sBeta[i] + (1. - beta2) * grad * grad) * div2;
            res.second[i] = beta[i] + alpha * (mBeta[i] / std::sqrt(sBeta[i] + 1E-8) - lambda * beta[i]);
        }
        return res;
    }

    IsingModel model;

    QC::Gates::HadamardGate<MatrixClass> hgate;
    QC::Gates::CNOTGate<MatrixClass> cnot;
    QC::Gates::RxGate<MatrixClass> rx;
    QC::Gates::RyGate<MatrixClass> ry;
    QC::Gates::RzGate<MatrixClass> rz;

    double deltaE = 0.0005;
    std::vector<double> gammaStart = {1.};
    std::vector<double> betaStart = {1.};
    std::vector<double> sGamma, mGamma, sBeta, mBeta;
    double beta1 = 0.8;
    double beta2 = 0.95;
    double lambda = 0.001;  // 0 gives Adam
    int stepNr = 0;
    double epsilon = 0.001;
    double stepSize = 0.0005;
    size_t nrMeasurements = 100000;
    bool betterMixing = false;
    size_t evaluations = 0;
};

}  // namespace Models
