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
#include <atomic>
#include <mutex>

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
