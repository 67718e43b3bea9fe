// autotune.hip - launch-variant selection for the streaming gate kernels.
//
// The in-place gate kernels come in variants (orbits per thread, nontemporal
// access, XCD-aware block order). Their speed depends on the lowest qubit a
// gate touches (it sets how a wave's accesses interleave and how far apart
// the two halves of each orbit lie) and on the state size. qcs_autotune()
// times every variant on the register's own buffer for each qubit and keeps
// the fastest in a process-wide table keyed by (precision, n, lowest qubit).
// Timing uses the generic 1-qubit kernel with the identity matrix, which
// moves exactly the bytes of a real gate and leaves every finite amplitude
// unchanged in value. Untuned entries fall back to the defaults measured on
// MI355X: for complex128 k_1q one amplitude per lane (k_1q_split) with
// nontemporal access and XCD-ordered blocks, the fastest layout for 27 of
// the 30 targets at 30 qubits; otherwise (complex64, where 8-byte lanes
// stream slower, and the other kernels) plain loads for qubits 0-2,
// nontemporal above, both with two orbits per thread and XCD-ordered blocks.
#include <algorithm>
#include <atomic>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace qcs {

// {orbits per thread, nontemporal (0: no, 1: loads and stores, 2: loads,
// 3: stores), block regions (0: none, 1: 8, 2: 64), log2 distance of a
// thread's orbits (k_1q; 0: adjacent chunks), one amplitude per lane (k_1q;
// k_1q_split)}; ids are persisted in tuning caches, so new variants go at
// the end.
static const Variant kVariants[] = {{2, 1, 1}, {2, 0, 1}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {2, 1, 0},
                                    {4, 0, 1}, {4, 1, 1}, {1, 1, 2}, {2, 1, 2}, {2, 0, 2}, {1, 2, 1},
                                    {1, 3, 1}, {1, 3, 0}, {2, 3, 1}, {2, 1, 1, 12}, {2, 1, 1, 16},
                                    {2, 1, 1, 20}, {4, 1, 1, 16}, {2, 3, 1, 16}, {2, 1, 1, 14},
                                    {2, 1, 1, 18}, {4, 1, 1, 12}, {2, 1, 2, 16}, {4, 1, 1, 14},
                                    {1, 1, 1, 0, 1}, {1, 1, 0, 0, 1}, {1, 0, 1, 0, 1}, {1, 3, 1, 0, 1},
                                    {1, 1, 2, 0, 1}, {1, 1, 1, 0, 2}, {1, 1, 0, 0, 2}, {1, 3, 1, 0, 2}};
constexpr int kNumVariants = int(sizeof(kVariants) / sizeof(kVariants[0]));
constexpr int kMaxN = 40;

static std::atomic<int8_t> g_table[2][kMaxN + 1][kMaxN];
static std::once_flag g_init;

static void init_table() {
    for (auto& a : g_table)
        for (auto& b : a)
            for (auto& c : b) c.store(-1, std::memory_order_relaxed);
}

int variant_count() { return kNumVariants; }
Variant variant_at(int i) { return kVariants[i]; }

void set_tuned(int prec, uint32_t n, uint32_t q, int vid) {
    std::call_once(g_init, init_table);
    if (n <= kMaxN && q < kMaxN) g_table[prec][n][q].store(int8_t(vid), std::memory_order_relaxed);
}

Variant pick_variant(const qcs_register* r, uint32_t lowbit, bool k1q) {
    std::call_once(g_init, init_table);
    if (r->n <= kMaxN && lowbit < kMaxN) {
        const int vid = g_table[r->prec][r->n][lowbit].load(std::memory_order_relaxed);
        if (vid >= 0) return kVariants[vid];
    }
    if (k1q && r->prec == QCS_COMPLEX128) return Variant{1, 1, 1, 0, 1};
    return lowbit < 3 ? Variant{2, 0, 1} : Variant{2, 1, 1};
}

}  // namespace qcs

using namespace qcs;

extern "C" int qcs_autotune(qcs_register* r, uint32_t reps, double* gbps) {
    if (!r) return set_error(QCS_EINVAL, "null register");
    if (reps == 0) reps = 3;
    const hcplx ident[4] = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
    const double bytes = 2.0 * double(r->amp_bytes()) * double(r->A);
    for (uint32_t q = 0; q < r->n; ++q) {
        float best = 1e30f;
        int best_id = 0;
        for (int v = 0; v < kNumVariants; ++v) {
            int e = k1q_variant(r, 0, q, nullptr, 0, ident, kVariants[v]);  // warm-up
            if (e) return e;
            if (hipEventRecord(r->ev0, r->stream) != hipSuccess) return set_error(QCS_EDEVICE, "event");
            for (uint32_t i = 0; i < reps; ++i) {
                e = k1q_variant(r, 0, q, nullptr, 0, ident, kVariants[v]);
                if (e) return e;
            }
            float ms = 0;
            if (hipEventRecord(r->ev1, r->stream) != hipSuccess || hipEventSynchronize(r->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, r->ev0, r->ev1) != hipSuccess)
                return set_error(QCS_EDEVICE, "autotune timing failed");
            ms /= float(reps);
            if (gbps) gbps[size_t(q) * kNumVariants + v] = bytes / (double(ms) * 1e6);
            if (ms < best) {
                best = ms;
                best_id = v;
            }
        }
        set_tuned(r->prec, r->n, q, best_id);
    }
    return QCS_OK;
}

extern "C" int qcs_autotune_variants(void) { return kNumVariants; }

extern "C" int qcs_set_tuned_variant(int precision, uint32_t n, uint32_t q, int vid) {
    if (precision != QCS_COMPLEX128 && precision != QCS_COMPLEX64) return set_error(QCS_EINVAL, "unknown precision");
    if (n > kMaxN || q >= n) return set_error(QCS_EINVAL, "qubit out of range");
    if (vid < -1 || vid >= kNumVariants) return set_error(QCS_EINVAL, "unknown variant");
    set_tuned(precision, n, q, vid);
    return QCS_OK;
}

extern "C" int qcs_get_tuned_variant(int precision, uint32_t n, uint32_t q) {
    std::call_once(g_init, init_table);
    if ((precision != QCS_COMPLEX128 && precision != QCS_COMPLEX64) || n > kMaxN || q >= kMaxN) return -1;
    return g_table[precision][n][q].load(std::memory_order_relaxed);
}

// ---------------------------------------------------------------------------
// Fusion width. A circuit runs fastest at the fusion width whose blocks cost
// least in total: wider blocks mean fewer passes over the state, but a dense
// k-qubit pass costs more than a plain gate once its 8 * 2^k FLOP per
// amplitude approach the matrix cores' rate. The cost of each pass kind is
// measured on the register's own buffer (identity matrices: the state keeps
// its values) and kept per (precision, n): t[0] for a generic 1-qubit gate,
// t[k] for a dense k-qubit block (k = 2..5). A gate that changes a fraction
// f of the amplitudes is priced at f * t[0] (2- and 3-qubit gates without
// control structure at t[k]); the width with the lowest predicted total is
// used.

namespace qcs {
int kdense(qcs_register* r, uint32_t k, const uint32_t* qubits, const double* mat_host);
int fuse(const qcs_gate* gates, uint32_t n, uint32_t width, std::vector<qcs_gate>& out);

static std::mutex g_fuse_mu;
static double g_fuse_cost[2][kMaxN + 1][6];  // ms; 0 = not measured

static int time_launches(qcs_register* r, uint32_t reps, const std::function<int()>& launch, double* ms) {
    if (const int e = launch()) return e;  // warm-up
    if (hipEventRecord(r->ev0, r->stream) != hipSuccess) return set_error(QCS_EDEVICE, "event");
    for (uint32_t i = 0; i < reps; ++i)
        if (const int e = launch()) return e;
    float t = 0;
    if (hipEventRecord(r->ev1, r->stream) != hipSuccess || hipEventSynchronize(r->ev1) != hipSuccess ||
        hipEventElapsedTime(&t, r->ev0, r->ev1) != hipSuccess)
        return set_error(QCS_EDEVICE, "fusion autotune timing failed");
    *ms = double(t) / reps;
    return QCS_OK;
}

// Fraction of the amplitudes a reference gate changes (its traits decide
// which kernel runs): 1 for generic, antidiagonal and two-entry diagonal
// gates; 1/2 for phase gates (m00 = 1), single-control gates and swaps;
// 1/4 for two-control gates, controlled phases and Fredkin; 1/8 for CCZ.
static double gate_fraction(const qcs_gate& g) {
    const uint32_t t = g.traits;
    const bool ctrl = t & QCS_TRAIT_CONTROLLED;
    auto one = [&](int i) { return g.mat[2 * i] == 1.0 && g.mat[2 * i + 1] == 0.0; };
    if (g.k == 1) return ((t & QCS_TRAIT_DIAGONAL) && one(0)) ? 0.5 : 1.0;
    if (g.k == 2) {
        if (t & (QCS_TRAIT_SWAP | QCS_TRAIT_ISWAP | QCS_TRAIT_ISWAPDAG)) return 0.5;
        if (ctrl) return ((t & QCS_TRAIT_DIAGONAL) && one(2 * 4 + 2)) ? 0.25 : 0.5;
        return -1;  // generic 4x4
    }
    if (t & QCS_TRAIT_SWAP) return 0.25;  // Fredkin
    if (ctrl && (t & QCS_TRAIT_CTRL_Q1)) return ((t & QCS_TRAIT_DIAGONAL) && one(6 * 8 + 6)) ? 0.125 : 0.25;
    if (ctrl) return -1;  // one control, 4x4 block: priced as a dense 2-qubit pass
    return -1;            // generic 8x8
}

static double block_cost(const qcs_gate& g, const double* t) {
    if ((g.flags & QCS_GATE_DENSE) || g.k > 3) return g.k == 1 ? t[0] : t[g.k];
    const double f = gate_fraction(g);
    if (f > 0) return f * t[0];
    const bool one_ctrl = g.k == 3 && (g.traits & QCS_TRAIT_CONTROLLED);
    return t[one_ctrl ? 2 : g.k];
}

int autotune_fusion(qcs_register* r, uint32_t reps, double* out6) {
    if (reps == 0) reps = 2;
    double t[6] = {0, 0, 0, 0, 0, 0};
    const hcplx ident1[4] = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
    const uint32_t n = r->n;
    const uint32_t q1[3] = {0, n / 2, n - 1};
    double sum = 0;
    for (uint32_t q : q1) {
        double ms = 0;
        if (const int e = time_launches(r, reps, [&] { return k1q(r, 0, q, nullptr, 0, ident1); }, &ms)) return e;
        sum += ms;
    }
    t[0] = sum / 3;
    std::vector<double> ident;
    for (uint32_t k = 2; k <= 5; ++k) {
        if (k > n) {
            t[k] = t[k - 1] * 2;  // no such block on this register
            continue;
        }
        const uint32_t D = 1u << k;
        ident.assign(size_t(2) * D * D, 0.0);
        for (uint32_t i = 0; i < D; ++i) ident[2 * (i * D + i)] = 1.0;
        uint32_t sets[3][5];
        for (uint32_t i = 0; i < k; ++i) {
            sets[0][i] = i;                          // lowest qubits
            sets[1][i] = n - k + i;                  // highest qubits
            sets[2][i] = uint32_t(uint64_t(i) * (n - 1) / (k - 1));  // spread
        }
        sum = 0;
        for (auto& qs : sets) {
            double ms = 0;
            if (const int e = time_launches(r, reps, [&] { return kdense(r, k, qs, ident.data()); }, &ms)) return e;
            sum += ms;
        }
        t[k] = sum / 3;
    }
    t[1] = t[0];
    {
        std::lock_guard<std::mutex> lock(g_fuse_mu);
        if (n <= kMaxN)
            for (int i = 0; i < 6; ++i) g_fuse_cost[r->prec][n][i] = t[i];
    }
    if (out6)
        for (int i = 0; i < 6; ++i) out6[i] = t[i];
    return QCS_OK;
}

// The fused sequence of the cheapest width (0 = unfused, else 2..5) under
// the measured costs; measures them first when this (precision, n) has none.
int pick_fusion(qcs_register* r, const qcs_gate* gates, uint32_t n, std::vector<qcs_gate>& out, uint32_t* width) {
    double t[6];
    bool have = false;
    {
        std::lock_guard<std::mutex> lock(g_fuse_mu);
        if (r->n <= kMaxN && g_fuse_cost[r->prec][r->n][0] > 0) {
            for (int i = 0; i < 6; ++i) t[i] = g_fuse_cost[r->prec][r->n][i];
            have = true;
        }
    }
    if (!have)
        if (const int e = autotune_fusion(r, 2, t)) return e;
    double best = 0;
    for (uint32_t i = 0; i < n; ++i) best += block_cost(gates[i], t);
    out.assign(gates, gates + n);
    *width = 0;
    // the widths are fused on host threads of their own (the pass is host
    // code and shares no state between calls), then priced in order
    const uint32_t wmax = std::min<uint32_t>(5, r->n);
    std::vector<std::vector<qcs_gate>> f(6);
    int err[6] = {0, 0, 0, 0, 0, 0};
    {
        std::vector<std::thread> pool;
        for (uint32_t w = 2; w <= wmax; ++w) pool.emplace_back([&, w] { err[w] = fuse(gates, n, w, f[w]); });
        for (auto& th : pool) th.join();
    }
    for (uint32_t w = 2; w <= wmax; ++w) {
        if (err[w]) return fuse(gates, n, w, f[w]);  // again on this thread, for its error message
        double c = 0;
        for (const qcs_gate& g : f[w]) c += block_cost(g, t);
        if (c < best) {
            best = c;
            *width = w;
            out.swap(f[w]);
        }
    }
    return QCS_OK;
}

}  // namespace qcs

extern "C" int qcs_autotune_fusion(qcs_register* r, uint32_t reps, double* ms) {
    if (!r) return set_error(QCS_EINVAL, "null register");
    return autotune_fusion(r, reps, ms);
}

extern "C" int qcs_set_fusion_costs(int precision, uint32_t n, const double* ms) {
    if (precision != QCS_COMPLEX128 && precision != QCS_COMPLEX64) return set_error(QCS_EINVAL, "unknown precision");
    if (n == 0 || n > kMaxN || !ms) return set_error(QCS_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(g_fuse_mu);
    for (int i = 0; i < 6; ++i) g_fuse_cost[precision][n][i] = ms[i];
    return QCS_OK;
}

extern "C" int qcs_get_fusion_costs(int precision, uint32_t n, double* ms) {
    if (precision != QCS_COMPLEX128 && precision != QCS_COMPLEX64) return set_error(QCS_EINVAL, "unknown precision");
    if (n == 0 || n > kMaxN || !ms) return set_error(QCS_EINVAL, "bad arguments");
    std::lock_guard<std::mutex> lock(g_fuse_mu);
    for (int i = 0; i < 6; ++i) ms[i] = g_fuse_cost[precision][n][i];
    return QCS_OK;
}
