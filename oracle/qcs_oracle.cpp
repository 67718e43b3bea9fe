// qcs_oracle.cpp - CPU oracle for the statevector engine. TEST INFRASTRUCTURE
// ONLY: it is loaded by tests/, by __graft_entry__.smoke() as the checker and
// by bench.py's cpu_baseline leg. The product (qcsim_amd/) never calls it.
//
// A fresh C++17 + OpenMP restatement of the reference's statevector path
// (aromanro/QCSim, GPL-3; no source copied). It keeps the reference's
// structure where it decides numerics or cost:
//   - per-gate loops over all basis states, a separate results buffer that is
//     swapped in afterwards (QubitRegister.h:451-479);
//   - the trait-driven dispatch of QubitRegisterCalculator.h:39-227,599-639;
//   - serial prefix copies of the control=0 part in controlled gates
//     (QubitRegisterCalculator.h:377-378,400-401,618-619,664-665,690-691);
//   - serial FP64 CDF scans with prob = 1 - u and `prob <= accum`
//     (QubitRegister.h:169-195,619-642; QubitRegisterCalculator.h:948-1254);
//   - OpenMP `parallel for` twins above 8192 amplitudes
//     (QubitRegisterCalculator.h:1276-1278, QubitRegister.h:454).
// Which reference lines each function follows is noted beside it.
//
// Pinning (tests/test_oracle_*.py): dense Kronecker products restated from
// SimpleGates.h:211-232,311-333,411-434 for n <= 8 and every kernel family;
// closed forms (Bell/GHZ, truth tables, QFT == sqrt(N) ifft); the known
// answers of MeasurementsTests.cpp:34-167, Tests.cpp:665-723,
// IsingTests.cpp:14; numpy's sequential cumsum for the CDF searches.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include <omp.h>

namespace {

using cplx = std::complex<double>;

constexpr uint32_t T_CONTROLLED = 0x01u, T_CTRL_Q1 = 0x02u, T_DIAG = 0x04u,
                   T_ANTIDIAG = 0x08u, T_SWAP = 0x10u, T_ISWAP = 0x20u,
                   T_ISWAPDAG = 0x40u;

// Reference threshold for the OpenMP twins (QubitRegisterCalculator.h:1276).
constexpr uint64_t kOmpLimit = 8192;

struct Mat {
    int dim = 0;
    std::vector<cplx> v;
    Mat(int d, const double* m) : dim(d), v(size_t(d) * d) {
        for (int i = 0; i < d * d; ++i) v[i] = cplx(m[2 * i], m[2 * i + 1]);
    }
    cplx operator()(int r, int c) const { return v[size_t(r) * dim + c]; }
};

struct Reg {
    uint32_t n;
    uint64_t A;
    std::vector<cplx> psi;  // registerStorage
    std::vector<cplx> res;  // resultsStorage
    bool omp = true;        // SetMultithreading
};

// Loop helper: the reference writes every per-state loop twice, serial and
// `#pragma omp parallel for` (default static schedule). One body, two drivers.
template <class F>
inline void for_states(bool par, uint64_t begin, uint64_t end, F&& f) {
    if (par) {
#pragma omp parallel for schedule(static)
        for (long long s = (long long)begin; s < (long long)end; ++s) f(uint64_t(s));
    } else {
        for (uint64_t s = begin; s < end; ++s) f(s);
    }
}

// ---- 1-qubit gates: QubitRegisterCalculator.h:39-135 ----------------------
void one_qubit(Reg& R, uint32_t traits, const Mat& m, uint64_t b, bool& swap_storage) {
    const bool par = R.omp && R.A >= kOmpLimit;
    cplx* psi = R.psi.data();
    cplx* out = R.res.data();
    if (traits & T_DIAG) {  // K1 (:41-52 / :89-102), in place
        swap_storage = false;
        const cplx d0 = m(0, 0), d1 = m(1, 1);
        for_states(par, 0, R.A, [&](uint64_t s) { psi[s] *= (s & b) ? d1 : d0; });
        return;
    }
    const uint64_t nb = ~b;
    if (traits & T_ANTIDIAG) {  // K2 (:57-67 / :107-119)
        const cplx m10 = m(1, 0), m01 = m(0, 1);
        for_states(par, 0, R.A, [&](uint64_t s) {
            out[s] = (s & b) ? m10 * psi[s & nb] : m01 * psi[s | b];
        });
        return;
    }
    // K3 generic (:68-79 / :120-133)
    for_states(par, 0, R.A, [&](uint64_t s) {
        const int row = (s & b) ? 1 : 0;
        out[s] = m(row, 0) * psi[s & nb] + m(row, 1) * psi[s | b];
    });
}

// ---- 2-qubit gates: QubitRegisterCalculator.h:137-177, 230-597 -----------
// b = qubit bit (local bit 0), c = controllingQubit1 bit (local bit 1).
void two_qubit(Reg& R, uint32_t traits, const Mat& m, uint64_t b, uint64_t c, bool& swap_storage) {
    const bool par = R.omp && R.A >= kOmpLimit;
    cplx* psi = R.psi.data();
    cplx* out = R.res.data();
    const uint64_t both = b | c;
    if (traits & (T_SWAP | T_ISWAP | T_ISWAPDAG)) {  // K4/K5 (:230-355), in place
        swap_storage = false;
        const bool phase = (traits & (T_SWAP)) == 0;
        const cplx ph = (traits & T_ISWAP) && !(traits & T_SWAP) ? cplx(0, 1) : cplx(0, -1);
        for_states(par, std::min(b, c), R.A, [&](uint64_t s) {
            if ((s & b) != 0 || (s & c) == 0) return;  // representative q=0, c=1
            const uint64_t o = s ^ both;
            std::swap(psi[s], psi[o]);
            if (phase) { psi[s] *= ph; psi[o] *= ph; }
        });
        return;
    }
    if (traits & T_CONTROLLED) {
        if (traits & T_DIAG) {  // K6 (:357-371 / :468-488), in place
            swap_storage = false;
            const cplx d2 = m(2, 2), d3 = m(3, 3);
            for_states(par, c, R.A, [&](uint64_t s) {
                if ((s & c) == 0) return;
                psi[s] *= (s & b) ? d3 : d2;
            });
            return;
        }
        // K7 / K8: serial prefix copy first, even in the OMP twin (:377-378,
        // :400-401, :498-499, :527-528).
        for (uint64_t s = 0; s < c; ++s) out[s] = psi[s];
        if (traits & T_ANTIDIAG) {  // K7 (:373-393 / :490-516)
            const cplx m32 = m(3, 2), m23 = m(2, 3);
            const uint64_t nb = ~b;
            for_states(par, c, R.A, [&](uint64_t s) {
                if ((s & c) == 0) { out[s] = psi[s]; return; }
                out[s] = (s & b) ? m32 * psi[s & nb] : m23 * psi[s | b];
            });
            return;
        }
        const uint64_t nb = ~b;  // K8 (:395-417 / :518-546)
        for_states(par, c, R.A, [&](uint64_t s) {
            if ((s & c) == 0) { out[s] = psi[s]; return; }
            const int row = 2 | ((s & b) ? 1 : 0);
            out[s] = m(row, 2) * psi[(s & nb) | c] + m(row, 3) * psi[s | both];
        });
        return;
    }
    // K9 generic 4x4 (:419-465 / :548-597)
    const uint64_t nb = ~b, nc = ~c;
    for_states(par, 0, R.A, [&](uint64_t s) {
        const int row = ((s & c) ? 2 : 0) | ((s & b) ? 1 : 0);
        const uint64_t s00 = s & nb & nc;
        out[s] = m(row, 0) * psi[s00] + m(row, 1) * psi[s00 | b] + m(row, 2) * psi[s00 | c] +
                 m(row, 3) * psi[s00 | both];
    });
}

// ---- 3-qubit gates: QubitRegisterCalculator.h:179-227, 599-939 -----------
// b = qubit (bit 0), b2 = controllingQubit1 (bit 1), c = controllingQubit2 (bit 2).
void three_qubit(Reg& R, uint32_t traits, const Mat& m, uint64_t b, uint64_t b2, uint64_t c,
                 bool& swap_storage) {
    const bool par = R.omp && R.A >= kOmpLimit;
    cplx* psi = R.psi.data();
    cplx* out = R.res.data();
    if (traits & T_SWAP) {  // K10 Fredkin (:181-194 / :203-222), in place
        swap_storage = false;
        const uint64_t bb = b | b2;
        for_states(par, std::max(c, std::min(b, b2)), R.A, [&](uint64_t s) {
            if ((s & c) == 0 || (s & b2) == 0 || (s & b) != 0) return;
            std::swap(psi[s], psi[s ^ bb]);
        });
        return;
    }
    if (traits & T_CONTROLLED) {
        if (traits & T_CTRL_Q1) {  // two controls: c and b2 (:601-609)
            const uint64_t cc = c | b2;
            const uint64_t limit = std::max(c, b2);
            if (traits & T_DIAG) {  // K11 (:642-656 / :812-833)
                swap_storage = false;
                const cplx d6 = m(6, 6), d7 = m(7, 7);
                for_states(par, limit, R.A, [&](uint64_t s) {
                    if ((s & cc) != cc) return;
                    psi[s] *= (s & b) ? d7 : d6;
                });
                return;
            }
            for (uint64_t s = 0; s < limit; ++s) out[s] = psi[s];  // serial prefix
            const uint64_t nb = ~b;
            if (traits & T_ANTIDIAG) {  // K12 Toffoli (:658-680 / :835-863)
                const cplx m76 = m(7, 6), m67 = m(6, 7);
                for_states(par, limit, R.A, [&](uint64_t s) {
                    if ((s & cc) != cc) { out[s] = psi[s]; return; }
                    out[s] = (s & b) ? m76 * psi[s & nb] : m67 * psi[s | b];
                });
                return;
            }
            // K13 CC generic (:682-707 / :865-896)
            for_states(par, limit, R.A, [&](uint64_t s) {
                if ((s & cc) != cc) { out[s] = psi[s]; return; }
                const int row = 6 | ((s & b) ? 1 : 0);
                out[s] = m(row, 6) * psi[(s & nb) | cc] + m(row, 7) * psi[s | cc | b];
            });
            return;
        }
        // K14 single control, 4x4 block rows 4..7 (:610-638 / :775-809)
        for (uint64_t s = 0; s < c; ++s) out[s] = psi[s];
        const uint64_t nb = ~b, nb2 = ~b2;
        for_states(par, c, R.A, [&](uint64_t s) {
            if ((s & c) == 0) { out[s] = psi[s]; return; }
            const int row = 4 | ((s & b2) ? 2 : 0) | ((s & b) ? 1 : 0);
            const uint64_t s0 = (s & nb & nb2) | c;
            out[s] = m(row, 4) * psi[s0] + m(row, 5) * psi[s0 | b] + m(row, 6) * psi[s0 | b2] +
                     m(row, 7) * psi[s0 | b | b2];
        });
        return;
    }
    // K15 generic 8x8 (:710-762 / :898-939)
    const uint64_t clear = ~(b | b2 | c);
    for_states(par, 0, R.A, [&](uint64_t s) {
        const int row = ((s & c) ? 4 : 0) | ((s & b2) ? 2 : 0) | ((s & b) ? 1 : 0);
        const uint64_t s0 = s & clear;
        cplx acc = 0;
        for (int col = 0; col < 8; ++col) {
            const uint64_t idx = s0 | ((col & 1) ? b : 0) | ((col & 2) ? b2 : 0) | ((col & 4) ? c : 0);
            acc += m(row, col) * psi[idx];
        }
        out[s] = acc;
    });
}

// Serial FP64 CDF scan: first i with prob <= accum, else `dflt`
// (QubitRegister.h:171-190, 621-639; QubitRegisterCalculator.h:955-970).
uint64_t cdf_search(const Reg& R, double prob, uint64_t dflt) {
    double accum = 0;
    for (uint64_t i = 0; i < R.A; ++i) {
        accum += std::norm(R.psi[i]);
        if (prob <= accum) return i;
    }
    return dflt;
}

thread_local char g_err[256];

}  // namespace

extern "C" {

void* orc_create(uint32_t n) {
    if (n == 0 || n > 34) return nullptr;
    Reg* R = new Reg;
    R->n = n;
    R->A = 1ull << n;
    R->psi.assign(R->A, cplx(0, 0));
    R->res.resize(R->A);  // resultsStorage is allocated too (QubitRegister.h:24)
    R->psi[0] = 1;
    return R;
}

void orc_destroy(void* h) { delete static_cast<Reg*>(h); }

void orc_set_multithreading(void* h, int on) { static_cast<Reg*>(h)->omp = on != 0; }

int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_set_state(void* h, const double* a) {
    Reg& R = *static_cast<Reg*>(h);
    std::memcpy(static_cast<void*>(R.psi.data()), a, R.A * sizeof(cplx));
}

void orc_get_state(void* h, double* a) {
    Reg& R = *static_cast<Reg*>(h);
    std::memcpy(a, R.psi.data(), R.A * sizeof(cplx));
}

// QubitRegister::ApplyGate with OPTIMIZED_TENSOR_PRODUCT (QubitRegister.h:434-486).
// Returns 0, or 1 for the invalid_argument cases of CheckQubits (:677-690).
int orc_apply_gate(void* h, uint32_t k, uint32_t traits, uint32_t q, uint32_t c1, uint32_t c2,
                   const double* mat) {
    Reg& R = *static_cast<Reg*>(h);
    if (q >= R.n) { std::snprintf(g_err, sizeof g_err, "Qubit number is too high"); return 1; }
    if (k == 2) {
        if (c1 >= R.n) { std::snprintf(g_err, sizeof g_err, "Controlling qubit number is too high"); return 1; }
        if (q == c1) { std::snprintf(g_err, sizeof g_err, "Qubit and controlling qubit are the same"); return 1; }
    } else if (k == 3) {
        if (c1 >= R.n || c2 >= R.n) { std::snprintf(g_err, sizeof g_err, "Controlling qubit number is too high"); return 1; }
        if (q == c1 || q == c2 || c1 == c2) { std::snprintf(g_err, sizeof g_err, "Qubits must be different"); return 1; }
    } else if (k != 1) {
        std::snprintf(g_err, sizeof g_err, "gate arity must be 1..3");
        return 1;
    }
    const Mat m(1 << k, mat);
    bool swap_storage = true;
    const uint64_t b = 1ull << q;
    if (k == 1) one_qubit(R, traits, m, b, swap_storage);
    else if (k == 2) two_qubit(R, traits, m, b, 1ull << c1, swap_storage);
    else three_qubit(R, traits, m, b, 1ull << c1, 1ull << c2, swap_storage);
    if (swap_storage) R.psi.swap(R.res);
    return 0;
}

// Dense k-qubit block on arbitrary qubits (k <= 5): the Kronecker-embedded
// definition of SimpleGates.h:211-232/311-333/411-434 generalised to k bits,
// evaluated per state like the K3/K9/K15 loops (results buffer + swap).
int orc_apply_dense(void* h, uint32_t k, const uint32_t* qubits, const double* mat) {
    Reg& R = *static_cast<Reg*>(h);
    if (k < 1 || k > 5) return 1;
    for (uint32_t i = 0; i < k; ++i) {
        if (qubits[i] >= R.n) return 1;
        for (uint32_t j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return 1;
    }
    const int dim = 1 << k;
    const Mat m(dim, mat);
    uint64_t bits[5], mask = 0;
    for (uint32_t i = 0; i < k; ++i) { bits[i] = 1ull << qubits[i]; mask |= bits[i]; }
    const bool par = R.omp && R.A >= kOmpLimit;
    cplx* psi = R.psi.data();
    cplx* out = R.res.data();
    for_states(par, 0, R.A, [&](uint64_t s) {
        int row = 0;
        for (uint32_t i = 0; i < k; ++i) row |= (s & bits[i]) ? (1 << i) : 0;
        const uint64_t s0 = s & ~mask;
        cplx acc = 0;
        for (int col = 0; col < dim; ++col) {
            uint64_t idx = s0;
            for (uint32_t i = 0; i < k; ++i) idx |= (col >> i & 1) ? bits[i] : 0;
            acc += m(row, col) * psi[idx];
        }
        out[s] = acc;
    });
    R.psi.swap(R.res);
    return 0;
}

// MeasureAll (QubitRegister.h:169-195): default 2^n - 1, collapse via
// setToBasisState (:74-80).
uint64_t orc_measure_all(void* h, double u) {
    Reg& R = *static_cast<Reg*>(h);
    const uint64_t st = cdf_search(R, 1. - u, R.A - 1);
    std::fill(R.psi.begin(), R.psi.end(), cplx(0, 0));
    R.psi[st] = 1;
    return st;
}

// Measure(first, second) (QubitRegister.h:208-224) -> Calc::MeasureQubit /
// Calc::Measure (QubitRegisterCalculator.h:948-1059, 1124-1225).
uint64_t orc_measure(void* h, uint32_t first, uint32_t second, double u) {
    Reg& R = *static_cast<Reg*>(h);
    const double prob = 1. - u;
    const bool par = R.omp && R.A >= kOmpLimit;
    const uint64_t lowmask = (1ull << first) - 1;
    const uint64_t mask = ((1ull << (second + 1)) - 1) - lowmask;
    const uint64_t pstate = cdf_search(R, prob, 0);
    const uint64_t meas = pstate & mask;
    // outcome norm (:981-987 / :1152-1158), OMP reduction in the twin
    double accum = 0;
    if (par) {
#pragma omp parallel for reduction(+ : accum) schedule(static)
        for (long long s = (long long)meas; s < (long long)R.A; ++s)
            if ((uint64_t(s) & mask) == meas) accum += std::norm(R.psi[s]);
    } else {
        for (uint64_t s = meas; s < R.A; ++s)
            if ((s & mask) == meas) accum += std::norm(R.psi[s]);
    }
    const double scale = 1. / std::sqrt(accum);
    cplx* psi = R.psi.data();
    for_states(par, 0, R.A, [&](uint64_t s) { psi[s] *= ((s & mask) == meas) ? scale : 0.; });
    return meas >> first;
}

// MeasureNoCollapse() (QubitRegister.h:619-642, default 0) and the protected
// MeasureNoCollapse(first, second) (:705-713 -> Calc:1061-1086, 1227-1254).
uint64_t orc_measure_no_collapse(void* h, int all, uint32_t first, uint32_t second, double u) {
    Reg& R = *static_cast<Reg*>(h);
    const uint64_t st = cdf_search(R, 1. - u, 0);
    if (all) return st;
    const uint64_t lowmask = (1ull << first) - 1;
    const uint64_t mask = ((1ull << (second + 1)) - 1) - lowmask;
    return (st & mask) >> first;
}

// RepeatedMeasure family (QubitRegister.h:227-429): truncated CDF built with
// the 2-way loop (truncation once accum > 1 - DBL_EPSILON), lower_bound per
// shot. nshots == 1 takes the MeasureNoCollapse shortcut, including the
// range overload's second mask-and-shift of an already shifted value
// (:337-338 on Calc:1253).
void orc_sample(void* h, int all, uint32_t first, uint32_t second, uint64_t nshots,
                const double* u, uint64_t* out) {
    Reg& R = *static_cast<Reg*>(h);
    const uint64_t lowmask = (1ull << first) - 1;
    const uint64_t mask = ((1ull << (second + 1)) - 1) - lowmask;
    if (nshots == 0) return;
    if (nshots == 1) {
        if (all) { out[0] = orc_measure_no_collapse(h, 1, 0, 0, u[0]); return; }
        const uint64_t m = orc_measure_no_collapse(h, 0, first, second, u[0]);
        out[0] = (m & mask) >> first;
        return;
    }
    std::vector<double> cdf(R.A);
    double accum = 0;
    size_t len = R.A;
    for (uint64_t i = 0; i < R.A; ++i) {
        accum += std::norm(R.psi[i]);
        cdf[i] = accum;
        if (accum > 1.0 - std::numeric_limits<double>::epsilon()) { len = i + 1; break; }
    }
    cdf.resize(len);
    for (uint64_t s = 0; s < nshots; ++s) {
        const double prob = 1. - u[s];
        const uint64_t meas = std::lower_bound(cdf.begin(), cdf.end(), prob) - cdf.begin();
        out[s] = all ? meas : (meas & mask) >> first;
    }
}

// GetQubitProbability (QubitRegister.h:592-598; Calc:1088-1122).
double orc_qubit_probability(void* h, uint32_t q) {
    Reg& R = *static_cast<Reg*>(h);
    const uint64_t b = 1ull << q;
    double accum = 0;
    if (R.omp && R.A >= kOmpLimit) {
#pragma omp parallel for reduction(+ : accum) schedule(static)
        for (long long s = (long long)b; s < (long long)R.A; ++s)
            if (uint64_t(s) & b) accum += std::norm(R.psi[s]);
    } else {
        for (uint64_t s = b; s < R.A; ++s)
            if (s & b) accum += std::norm(R.psi[s]);
    }
    return accum;
}

double orc_norm2(void* h) {
    Reg& R = *static_cast<Reg*>(h);
    double a = 0;
    for (uint64_t s = 0; s < R.A; ++s) a += std::norm(R.psi[s]);
    return a;
}

// Normalize (QubitRegister.h:124-130).
void orc_normalize(void* h) {
    Reg& R = *static_cast<Reg*>(h);
    const double nrm = std::sqrt(orc_norm2(h));
    if (nrm < 1e-20) return;
    const double inv = 1. / nrm;
    for (auto& x : R.psi) x *= inv;
}

// <a|b> (stateFidelity :527-534, ExpectationValue :646-660).
void orc_inner(void* ha, void* hb, double* re_im) {
    Reg& a = *static_cast<Reg*>(ha);
    Reg& b = *static_cast<Reg*>(hb);
    cplx acc = 0;
    for (uint64_t s = 0; s < a.A; ++s) acc += std::conj(a.psi[s]) * b.psi[s];
    re_im[0] = acc.real();
    re_im[1] = acc.imag();
}

// QFT / IQFT as the reference's gate sequence (QuantumFourierTransform.h:35-87,
// QubitsSwapper.h:23-33, range clamp QuantumAlgorithm.h:228-229).
void orc_qft(void* h, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    Reg& R = *static_cast<Reg*>(h);
    eq = std::max(sq, std::min(R.n - 1, eq));
    const double hs = 1. / std::sqrt(2.);
    const double H[8] = {hs, 0, hs, 0, hs, 0, -hs, 0};
    auto cphase = [&](uint32_t cur, uint32_t ctrl, double phase) {
        double cp[32] = {0};
        for (int i = 0; i < 4; ++i) cp[2 * (i * 4 + i)] = 1;
        const cplx p = std::polar(1., phase);
        cp[2 * 15] = p.real();
        cp[2 * 15 + 1] = p.imag();
        orc_apply_gate(h, 2, T_CONTROLLED | T_DIAG, cur, ctrl, 0, cp);
    };
    auto swaps = [&]() {
        double sw[32] = {0};
        sw[2 * 0] = 1; sw[2 * 6] = 1; sw[2 * 9] = 1; sw[2 * 15] = 1;
        for (uint32_t a = sq, b = eq; a < b; ++a, --b) orc_apply_gate(h, 2, T_SWAP, a, b, 0, sw);
    };
    const int s = int(sq), e = int(eq);
    if (!inverse) {
        orc_apply_gate(h, 1, 0, eq, 0, 0, H);
        for (int cur = e; cur > s; --cur) {
            double phase = M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) { cphase(cur, ctrl, phase); phase *= 0.5; }
            orc_apply_gate(h, 1, 0, cur - 1, 0, 0, H);
        }
        if (do_swap) swaps();
    } else {
        if (do_swap) swaps();
        for (int cur = s + 1; cur <= e; ++cur) {
            orc_apply_gate(h, 1, 0, cur - 1, 0, 0, H);
            double phase = -M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) { cphase(cur, ctrl, phase); phase *= 0.5; }
        }
        orc_apply_gate(h, 1, 0, eq, 0, 0, H);
    }
    (void)R;
}

// Ising energy of one basis state, E(s) = -(sum h_i z_i + sum J_ij z_i z_j),
// z = +1 for bit 0 (IsingModel.h:106-119, 203-206).
double orc_ising_state_energy(uint64_t s, uint32_t n, uint32_t ne, const uint32_t* ei,
                              const uint32_t* ej, const double* J, const double* hv) {
    double e = 0;
    for (uint32_t i = 0; i < n; ++i) e += hv[i] * ((s >> i & 1) ? -1. : 1.);
    for (uint32_t k = 0; k < ne; ++k) {
        const double zi = (s >> ei[k] & 1) ? -1. : 1.;
        const double zj = (s >> ej[k] & 1) ? -1. : 1.;
        e += J[k] * zi * zj;
    }
    return -e;
}

// Exact sum_s |psi_s|^2 E(s): the expectation the reference estimates with
// 100000 shots (IsingModel.h:569-578).
double orc_ising_energy(void* h, uint32_t ne, const uint32_t* ei, const uint32_t* ej,
                        const double* J, const double* hv) {
    Reg& R = *static_cast<Reg*>(h);
    double acc = 0;
    for (uint64_t s = 0; s < R.A; ++s)
        acc += std::norm(R.psi[s]) * orc_ising_state_energy(s, R.n, ne, ei, ej, J, hv);
    return acc;
}

// QAOAIsingSubalgorithm::Exec (IsingModel.h:557-567): |0> -> H on all, then per
// layer ApplyIsingOperator (:491-531: per edge CNOT(q1 <- q2), Rz(2 gamma J) on
// q1, CNOT; per site Rz(gamma h) skipped when h == 0) and ApplyMixingOperator
// (:533-554: Rx(2 beta) on all, optional CNOT ring + Ry(2 beta)). Edges are
// given as (i < j) pairs in the caller's order.
void orc_qaoa_exec(void* h, uint32_t ne, const uint32_t* ei, const uint32_t* ej, const double* J,
                   const double* hv, uint32_t p, const double* gamma, const double* beta,
                   int better_mixing) {
    Reg& R = *static_cast<Reg*>(h);
    std::fill(R.psi.begin(), R.psi.end(), cplx(0, 0));
    R.psi[0] = 1;
    const double hs = 1. / std::sqrt(2.);
    const double H[8] = {hs, 0, hs, 0, hs, 0, -hs, 0};
    double cnot[32] = {0};
    cnot[0] = 1; cnot[2 * 5] = 1; cnot[2 * 11] = 1; cnot[2 * 14] = 1;
    auto rz = [](double th, double* m) {
        const cplx a = std::polar(1., -th * 0.5), b = std::polar(1., th * 0.5);
        m[0] = a.real(); m[1] = a.imag(); m[2] = m[3] = m[4] = m[5] = 0; m[6] = b.real(); m[7] = b.imag();
    };
    auto rx = [](double th, double* m) {
        const double c = std::cos(th * 0.5), s = std::sin(th * 0.5);
        m[0] = c; m[1] = 0; m[2] = 0; m[3] = -s; m[4] = 0; m[5] = -s; m[6] = c; m[7] = 0;
    };
    auto ry = [](double th, double* m) {
        const double c = std::cos(th * 0.5), s = std::sin(th * 0.5);
        m[0] = c; m[1] = 0; m[2] = -s; m[3] = 0; m[4] = s; m[5] = 0; m[6] = c; m[7] = 0;
    };
    for (uint32_t q = 0; q < R.n; ++q) orc_apply_gate(h, 1, 0, q, 0, 0, H);
    double m[8];
    for (uint32_t l = 0; l < p; ++l) {
        for (uint32_t k = 0; k < ne; ++k) {
            rz(2 * gamma[l] * J[k], m);
            orc_apply_gate(h, 2, T_CONTROLLED | T_ANTIDIAG, ei[k], ej[k], 0, cnot);
            orc_apply_gate(h, 1, T_DIAG, ei[k], 0, 0, m);
            orc_apply_gate(h, 2, T_CONTROLLED | T_ANTIDIAG, ei[k], ej[k], 0, cnot);
        }
        for (uint32_t q = 0; q < R.n; ++q) {
            if (hv[q] == 0.) continue;
            rz(gamma[l] * hv[q], m);
            orc_apply_gate(h, 1, T_DIAG, q, 0, 0, m);
        }
        rx(2 * beta[l], m);
        for (uint32_t q = 0; q < R.n; ++q) orc_apply_gate(h, 1, 0, q, 0, 0, m);
        if (better_mixing) {
            ry(2 * beta[l], m);
            const uint32_t last = R.n - 1;
            for (uint32_t q = 0; q < last; ++q)
                orc_apply_gate(h, 2, T_CONTROLLED | T_ANTIDIAG, q + 1, q, 0, cnot);
            orc_apply_gate(h, 2, T_CONTROLLED | T_ANTIDIAG, 0, last, 0, cnot);
            for (uint32_t q = 0; q < R.n; ++q) orc_apply_gate(h, 1, 0, q, 0, 0, m);
        }
    }
}

const char* orc_last_error(void) { return g_err; }

// Host memory roofline for the CPU baseline: the STREAM triad
// a[i] = b[i] + s * c[i] (McCalpin's STREAM benchmark, counted as 3 x 8 B
// per element) over n doubles per array with `threads` OpenMP threads (0:
// all), first-touched by the same static schedule; best of `reps` passes,
// in GB/s.
double orc_stream_triad(uint64_t n, int reps, int threads) {
    const int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    // uninitialised storage: the first touch below places the pages
    std::unique_ptr<double[]> a(new double[n]), b(new double[n]), c(new double[n]);
    double* pa = a.get();
    double* pb = b.get();
    double* pc = c.get();
    const int64_t N = int64_t(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        pa[i] = 0;
        pb[i] = 1;
        pc[i] = 2;
    }
    double best = 0;
    for (int r = 0; r < reps; ++r) {
        const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) pa[i] = pb[i] + 3.0 * pc[i];
        const double dt = omp_get_wtime() - t0;
        best = std::max(best, 24.0 * double(n) / dt / 1e9);
    }
    omp_set_num_threads(saved);
    volatile double sink = pa[N / 2];
    (void)sink;
    return best;
}

}  // extern "C"
