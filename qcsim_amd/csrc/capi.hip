// capi.hip - the extern "C" boundary (include/qcsim_amd.h).
//
// Validation, dispatch and host<->device plumbing. Gate dispatch mirrors
// QubitRegister::ApplyGate (QubitRegister.h:434-486) and the trait tests of
// QubitRegisterCalculator.h:39-227,599-639, mapped onto in-place kernels:
//   k=1: diagonal -> phase on the bit=1 half when m00 == 1 (P, S, T, Z), else
//        both entries; antidiagonal; generic.
//   k=2: SWAP / iSWAP / iSWAP-dagger -> pair exchange on the q != c quarter
//        pairs; controlled diagonal / antidiagonal / generic -> 1-qubit
//        update on the control=1 half (phase-only when m22 == 1: CZ, CP);
//        otherwise a 4x4 block.
//   k=3: Fredkin -> exchange on the control=1 quarter; two-control gates ->
//        1-qubit update on the control=11 quarter (phase-only for CCZ);
//        single-control -> 4x4 block on the control=1 half; else 8x8.
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <functional>
#include <vector>

#include "common.h"

namespace qcs {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int check_launch(qcs_register* r, const char* what) {
    r->launches++;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(QCS_EDEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
    return QCS_OK;
}

int qft_sequence(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse);
int fuse(const qcs_gate* gates, uint32_t n, uint32_t width, std::vector<qcs_gate>& out);
int fuse_stream(const qcs_gate* gates, uint32_t n, uint32_t width, const std::function<int(const qcs_gate&)>& emit);

int stage_upload(qcs_register* r, const void* src, size_t bytes, const void** dev) {
    if (bytes > r->mat_bytes) return set_error(QCS_EINVAL, "parameter block too large");
    if (r->sink) {  // capturing a program: the program owns the parameter memory
        ParamSink& ps = *r->sink;
        const size_t need = (bytes + 255) & ~size_t(255);
        if (ps.used + need > ps.cap) return set_error(QCS_EINVAL, "program parameter space exhausted");
        std::memcpy(ps.host + ps.used, src, bytes);
        *dev = ps.dev + ps.used;
        ps.used += need;
        return QCS_OK;
    }
    const uint32_t slot = r->mat_next++ % qcs_register::kMatSlots;
    if (hipEventSynchronize(r->mat_ev[slot]) != hipSuccess) return set_error(QCS_EDEVICE, "staging wait failed");
    char* h = static_cast<char*>(r->hmat) + slot * r->mat_bytes;
    char* d = static_cast<char*>(r->mat) + slot * r->mat_bytes;
    std::memcpy(h, src, bytes);
    if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, r->stream) != hipSuccess ||
        hipEventRecord(r->mat_ev[slot], r->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "parameter upload failed");
    *dev = d;
    return QCS_OK;
}
}  // namespace qcs

using namespace qcs;

#define QCS_REQUIRE(cond, ...) \
    do {                       \
        if (!(cond)) return set_error(QCS_EINVAL, __VA_ARGS__); \
    } while (0)

static uint64_t device_cap() {
    const char* s = std::getenv("QCS_MAX_DEVICE_BYTES");
    return s ? std::strtoull(s, nullptr, 10) : 0;
}

extern "C" {

const char* qcs_last_error(void) { return g_err.c_str(); }
const char* qcs_version(void) { return "qcsim_amd 0.1 gfx950"; }

int qcs_create(uint32_t nqubits, int precision, qcs_register** out) {
    QCS_REQUIRE(out != nullptr, "null output handle");
    QCS_REQUIRE(nqubits > 0 && nqubits <= 36, "number of qubits must be in 1..36");
    QCS_REQUIRE(precision == QCS_COMPLEX128 || precision == QCS_COMPLEX64, "unknown precision");
    *out = nullptr;
    auto* r = new qcs_register;
    r->n = nqubits;
    r->A = 1ull << nqubits;
    r->prec = precision;
    if (hipGetDevice(&r->device) != hipSuccess) {
        delete r;
        return set_error(QCS_EDEVICE, "no HIP device");
    }
    const size_t bytes = r->A * r->amp_bytes();
    const uint64_t nchunks = r->A > (1ull << kChunkLog) ? (r->A >> kChunkLog) : 1;
    const uint64_t cap = device_cap();
    if (cap && bytes > cap) {
        delete r;
        return set_error(QCS_ENOMEM, "state of %zu bytes exceeds QCS_MAX_DEVICE_BYTES", bytes);
    }
    auto fail = [&](int code, const char* what) {
        if (r->psi) (void)hipFree(r->psi);
        if (r->red) (void)hipFree(r->red);
        if (r->chunk) (void)hipFree(r->chunk);
        if (r->dout) (void)hipFree(r->dout);
        if (r->mat) (void)hipFree(r->mat);
        if (r->xfer) (void)hipFree(r->xfer);
        if (r->hmat) (void)hipHostFree(r->hmat);
        if (r->hpin) (void)hipHostFree(r->hpin);
        for (hipEvent_t ev : r->mat_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (r->stream) (void)hipStreamDestroy(r->stream);
        delete r;
        return set_error(code, "%s", what);
    };
    if (hipMalloc(&r->psi, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(QCS_ENOMEM, "hipMalloc of the state failed");
    }
    r->mat_bytes = sizeof(double) * 64 * 64;  // real embedding of a 32x32 complex block
    if (hipMalloc(reinterpret_cast<void**>(&r->red), sizeof(double) * kRedDoubles) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&r->chunk), sizeof(double) * (nchunks + (nchunks + 2047) / 2048)) !=
            hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&r->dout), sizeof(uint64_t) * 64) != hipSuccess ||
        hipMalloc(&r->mat, r->mat_bytes * qcs_register::kMatSlots) != hipSuccess) {
        (void)hipGetLastError();
        return fail(QCS_ENOMEM, "hipMalloc of scratch failed");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&r->hpin), sizeof(double) * 64) != hipSuccess ||
        hipHostMalloc(&r->hmat, r->mat_bytes * qcs_register::kMatSlots) != hipSuccess)
        return fail(QCS_ENOMEM, "hipHostMalloc failed");
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&r->ev0) != hipSuccess || hipEventCreate(&r->ev1) != hipSuccess)
        return fail(QCS_EDEVICE, "stream/event creation failed");
    for (hipEvent_t& ev : r->mat_ev)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ev, r->stream) != hipSuccess)
            return fail(QCS_EDEVICE, "staging event creation failed");
    int e = kfill(r, 0, 0);
    if (!e) e = kset_one(r, 0, 1, 0);
    if (!e && hipStreamSynchronize(r->stream) != hipSuccess) e = set_error(QCS_EDEVICE, "init sync failed");
    if (e) {
        qcs_destroy(r);
        return e;
    }
    *out = r;
    return QCS_OK;
}

int qcs_destroy(qcs_register* r) {
    if (!r) return QCS_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    (void)hipFree(r->psi);
    (void)hipFree(r->red);
    (void)hipFree(r->chunk);
    (void)hipFree(r->dout);
    (void)hipFree(r->mat);
    if (r->xfer) (void)hipFree(r->xfer);
    (void)hipHostFree(r->hmat);
    (void)hipHostFree(r->hpin);
    for (hipEvent_t ev : r->mat_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return QCS_OK;
}

uint32_t qcs_get_nqubits(const qcs_register* r) { return r ? r->n : 0; }
int qcs_get_precision(const qcs_register* r) { return r ? r->prec : -1; }
uint64_t qcs_launch_count(const qcs_register* r) { return r ? r->launches : 0; }
uint32_t qcs_last_fuse_width(const qcs_register* r) { return r ? r->last_fuse_width : 0; }

int qcs_copy_state(qcs_register* dst, const qcs_register* src) {
    QCS_REQUIRE(dst && src, "null register");
    QCS_REQUIRE(dst->A == src->A && dst->prec == src->prec, "registers differ in size or precision");
    if (dst == src) return QCS_OK;
    // The copy runs on dst's stream after src's pending work; it completes
    // before returning, so either register may be modified right after.
    if (hipStreamSynchronize(src->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
    if (hipMemcpyAsync(dst->psi, src->psi, src->A * src->amp_bytes(), hipMemcpyDeviceToDevice, dst->stream) !=
            hipSuccess ||
        hipStreamSynchronize(dst->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "device copy failed");
    return QCS_OK;
}

int qcs_swap_state(qcs_register* a, qcs_register* b) {
    QCS_REQUIRE(a && b, "null register");
    QCS_REQUIRE(a->A == b->A && a->prec == b->prec, "registers differ in size or precision");
    if (hipStreamSynchronize(a->stream) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "sync failed");
    std::swap(a->psi, b->psi);
    return QCS_OK;
}

int qcs_set_state(qcs_register* r, const double* amps, uint64_t n_amps) {
    QCS_REQUIRE(r && amps, "null argument");
    QCS_REQUIRE(n_amps == r->A, "state size mismatch");
    if (r->prec == QCS_COMPLEX128) {
        if (hipMemcpyAsync(r->psi, amps, r->A * 16, hipMemcpyHostToDevice, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "upload failed");
        return hipStreamSynchronize(r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "sync failed");
    }
    return convert_upload(r, amps);
}

int qcs_get_state(qcs_register* r, double* amps, uint64_t n_amps) {
    QCS_REQUIRE(r && amps, "null argument");
    QCS_REQUIRE(n_amps == r->A, "state size mismatch");
    if (r->prec == QCS_COMPLEX128) {
        if (hipMemcpyAsync(amps, r->psi, r->A * 16, hipMemcpyDeviceToHost, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "download failed");
        return hipStreamSynchronize(r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "sync failed");
    }
    return convert_download(r, amps);
}

int qcs_get_amplitude(qcs_register* r, uint64_t idx, double* re_im) {
    QCS_REQUIRE(r && re_im, "null argument");
    re_im[0] = re_im[1] = 0;
    if (idx >= r->A) return QCS_OK;  // getBasisStateAmplitude returns 0 (:63)
    const size_t ab = r->amp_bytes();
    if (hipMemcpyAsync(r->hpin, static_cast<char*>(r->psi) + idx * ab, ab, hipMemcpyDeviceToHost, r->stream) !=
            hipSuccess ||
        hipStreamSynchronize(r->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "amplitude read failed");
    if (r->prec == QCS_COMPLEX128) {
        re_im[0] = r->hpin[0];
        re_im[1] = r->hpin[1];
    } else {
        const float* f = reinterpret_cast<const float*>(r->hpin);
        re_im[0] = f[0];
        re_im[1] = f[1];
    }
    return QCS_OK;
}

int qcs_set_amplitude(qcs_register* r, uint64_t idx, double re, double im) {
    QCS_REQUIRE(r, "null register");
    if (idx >= r->A) return QCS_OK;  // silent no-op (:114)
    return kset_one(r, idx, re, im);
}

int qcs_set_basis_state(qcs_register* r, uint64_t idx) {
    QCS_REQUIRE(r, "null register");
    if (idx >= r->A) return QCS_OK;  // silent no-op (:76)
    int e = kfill(r, 0, 0);
    return e ? e : kset_one(r, idx, 1, 0);
}

int qcs_fill(qcs_register* r, double re, double im) {
    QCS_REQUIRE(r, "null register");
    return kfill(r, re, im);
}

int qcs_scale(qcs_register* r, double re, double im) {
    QCS_REQUIRE(r, "null register");
    return kscale(r, re, im);
}

static int check_qubits(const qcs_register* r, uint32_t k, uint32_t q, uint32_t c1, uint32_t c2) {
    // CheckQubits, QubitRegister.h:677-690 (same messages)
    if (r->n == 0) return set_error(QCS_EINVAL, "Qubit number is zero");
    if (r->n <= q) return set_error(QCS_EINVAL, "Qubit number is too high");
    if (k == 2) {
        if (r->n <= c1) return set_error(QCS_EINVAL, "Controlling qubit number is too high");
        if (q == c1) return set_error(QCS_EINVAL, "Qubit and controlling qubit are the same");
    } else if (k == 3) {
        if (r->n <= c1 || r->n <= c2) return set_error(QCS_EINVAL, "Controlling qubit number is too high");
        if (q == c1 || q == c2 || c1 == c2) return set_error(QCS_EINVAL, "Qubits must be different");
    } else if (k != 1) {
        return set_error(QCS_EINVAL, "gate must act on 1, 2 or 3 qubits");
    }
    return QCS_OK;
}

static inline hcplx M(const double* mat, int dim, int row, int col) {
    const int i = row * dim + col;
    return hcplx{mat[2 * i], mat[2 * i + 1]};
}

static inline bool is_one(hcplx z) { return z.re == 1.0 && z.im == 0.0; }

int qcs_apply_gate(qcs_register* r, uint32_t k, uint32_t traits, uint32_t q, uint32_t c1, uint32_t c2,
                   const double* mat) {
    QCS_REQUIRE(r && mat, "null argument");
    int e = check_qubits(r, k, q, c1, c2);
    if (e) return e;
    if (k == 1) {
        if (traits & QCS_TRAIT_DIAGONAL) {
            const hcplx d0 = M(mat, 2, 0, 0), d1 = M(mat, 2, 1, 1);
            if (is_one(d0)) return kphase(r, &q, 1, d1);
            const hcplx m4[4] = {d0, {0, 0}, {0, 0}, d1};
            return k1q(r, 2, q, nullptr, 0, m4);
        }
        const hcplx m4[4] = {M(mat, 2, 0, 0), M(mat, 2, 0, 1), M(mat, 2, 1, 0), M(mat, 2, 1, 1)};
        return k1q(r, (traits & QCS_TRAIT_ANTIDIAGONAL) ? 1 : 0, q, nullptr, 0, m4);
    }
    if (k == 2) {
        if (traits & QCS_TRAIT_SWAP) return kswap(r, q, c1, nullptr, 0, hcplx{1, 0}, false);
        if (traits & QCS_TRAIT_ISWAP) return kswap(r, q, c1, nullptr, 0, hcplx{0, 1}, true);
        if (traits & QCS_TRAIT_ISWAPDAG) return kswap(r, q, c1, nullptr, 0, hcplx{0, -1}, true);
        if (traits & QCS_TRAIT_CONTROLLED) {
            const hcplx blk[4] = {M(mat, 4, 2, 2), M(mat, 4, 2, 3), M(mat, 4, 3, 2), M(mat, 4, 3, 3)};
            if (traits & QCS_TRAIT_DIAGONAL) {
                if (is_one(blk[0])) {
                    const uint32_t bits[2] = {q, c1};
                    return kphase(r, bits, 2, blk[3]);
                }
                return k1q(r, 2, q, &c1, 1, blk);
            }
            return k1q(r, (traits & QCS_TRAIT_ANTIDIAGONAL) ? 1 : 0, q, &c1, 1, blk);
        }
        hcplx m16[16];
        for (int i = 0; i < 16; ++i) m16[i] = hcplx{mat[2 * i], mat[2 * i + 1]};
        return k2q(r, q, c1, nullptr, 0, m16);
    }
    // k == 3: qubit = bit 0, c1 = bit 1, c2 = bit 2
    if (traits & QCS_TRAIT_SWAP) return kswap(r, q, c1, &c2, 1, hcplx{1, 0}, false);
    if (traits & QCS_TRAIT_CONTROLLED) {
        if (traits & QCS_TRAIT_CTRL_Q1) {
            const hcplx blk[4] = {M(mat, 8, 6, 6), M(mat, 8, 6, 7), M(mat, 8, 7, 6), M(mat, 8, 7, 7)};
            const uint32_t ctrls[2] = {c1, c2};
            if (traits & QCS_TRAIT_DIAGONAL) {
                if (is_one(blk[0])) {
                    const uint32_t bits[3] = {q, c1, c2};
                    return kphase(r, bits, 3, blk[3]);
                }
                return k1q(r, 2, q, ctrls, 2, blk);
            }
            return k1q(r, (traits & QCS_TRAIT_ANTIDIAGONAL) ? 1 : 0, q, ctrls, 2, blk);
        }
        hcplx m16[16];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m16[i * 4 + j] = M(mat, 8, 4 + i, 4 + j);
        return k2q(r, q, c1, &c2, 1, m16);
    }
    hcplx m64[64];
    for (int i = 0; i < 64; ++i) m64[i] = hcplx{mat[2 * i], mat[2 * i + 1]};
    return k3q(r, q, c1, c2, m64);
}

int qcs_apply_dense(qcs_register* r, uint32_t k, const uint32_t* qubits, const double* mat) {
    QCS_REQUIRE(r && qubits && mat, "null argument");
    QCS_REQUIRE(k >= 1 && k <= 5, "dense block must act on 1..5 qubits");
    for (uint32_t i = 0; i < k; ++i) {
        QCS_REQUIRE(qubits[i] < r->n, "Qubit number is too high");
        for (uint32_t j = 0; j < i; ++j) QCS_REQUIRE(qubits[i] != qubits[j], "Qubits must be different");
    }
    return kdense(r, k, qubits, mat);
}

int qcs_apply_circuit(qcs_register* r, const qcs_gate* gates, uint32_t n, uint32_t fuse_width) {
    QCS_REQUIRE(r && (gates || n == 0), "null argument");
    QCS_REQUIRE(fuse_width <= 5 || fuse_width == QCS_FUSE_AUTO, "fusion width must be 0..5 or QCS_FUSE_AUTO");
    r->last_fuse_width = fuse_width == QCS_FUSE_AUTO ? 0 : fuse_width;
    auto apply = [r](const qcs_gate& g) {
        const bool dense = (g.flags & QCS_GATE_DENSE) || g.k > 3;
        return dense ? qcs_apply_dense(r, g.k, g.qubits, g.mat)
                     : qcs_apply_gate(r, g.k, g.traits, g.qubits[0], g.qubits[1], g.qubits[2], g.mat);
    };
    if (fuse_width < 2) {
        for (uint32_t i = 0; i < n; ++i)
            if (const int e = apply(gates[i])) return e;
        return QCS_OK;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < gates[i].k && j < 5; ++j)
            QCS_REQUIRE(gates[i].qubits[j] < r->n, "gate %u: qubit out of range", i);
    if (fuse_width != QCS_FUSE_AUTO) {
        // each block is launched as soon as the fusion pass has built it, so
        // that the rest of the pass runs on the host while the GPU applies it
        r->last_fuse_width = fuse_width;
        return fuse_stream(gates, n, fuse_width, apply);
    }
    std::vector<qcs_gate> fused;
    uint32_t picked = 0;
    if (const int e = pick_fusion(r, gates, n, fused, &picked)) return e;
    r->last_fuse_width = picked;
    for (const qcs_gate& g : fused)
        if (const int e = apply(g)) return e;
    return QCS_OK;
}

int qcs_apply_diagonal(qcs_register* r, const double* diag, uint64_t n) {
    QCS_REQUIRE(r && diag, "null argument");
    QCS_REQUIRE(n == r->A, "diagonal size mismatch");
    return apply_diagonal_host(r, diag);
}

int qcs_apply_matrix(qcs_register* r, const double* mat, uint64_t dim) {
    QCS_REQUIRE(r && mat, "null argument");
    QCS_REQUIRE(dim == r->A, "operator size mismatch");
    return apply_matrix_host(r, mat);
}

int qcs_inner_host(qcs_register* r, const double* amps, uint64_t n_amps, double* re_im) {
    QCS_REQUIRE(r && amps && re_im, "null argument");
    QCS_REQUIRE(n_amps == r->A, "state size mismatch");
    return inner_host(r, amps, re_im);
}

int qcs_pauli_expectation(qcs_register* r, uint32_t nterms, const uint64_t* xmask, const uint64_t* zmask,
                          double* out) {
    QCS_REQUIRE(r, "null register");
    if (nterms == 0) return QCS_OK;
    QCS_REQUIRE(xmask && zmask && out, "null argument");
    return pauli_expectation(r, nterms, xmask, zmask, out);
}

int qcs_qft(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    QCS_REQUIRE(r, "null register");
    QCS_REQUIRE(sq < r->n, "start qubit out of range");
    // QuantumSubAlgorithmOnSubregister clamps the end qubit (QuantumAlgorithm.h:228-229)
    const uint32_t e = std::max(sq, std::min(r->n - 1, eq));
    return qft(r, sq, e, do_swap, inverse);
}

static inline uint64_t range_mask(uint32_t first, uint32_t second) {
    const uint64_t low = (1ull << first) - 1;
    const uint64_t upto = second >= 63 ? ~0ull : ((1ull << (second + 1)) - 1);
    return upto - low;
}

}  // extern "C"

namespace qcs {
int measure_no_collapse_impl(qcs_register* r, int all, uint32_t first, uint32_t second, double u,
                             uint64_t* outcome, const ShotDev* dev) {
    QCS_REQUIRE(r && outcome, "null argument");
    if (!all) QCS_REQUIRE(first <= second && second < r->n, "measured qubit range out of bounds");
    // prob = 1 - u (QubitRegister.h:621), formed on the device
    uint64_t st = 0;
    const int e = cdf_search(r, 1, &u, 1, 0, 0, &st, 1, dev);
    if (e) return e;
    *outcome = all ? st : (st & range_mask(first, second)) >> first;
    return QCS_OK;
}

int sample_impl(qcs_register* r, int all, uint32_t first, uint32_t second, uint64_t nshots, const double* u,
                uint64_t* outcomes, const ShotDev* dev) {
    QCS_REQUIRE(r && (u || nshots == 0) && (outcomes || nshots == 0), "null argument");
    if (!all) QCS_REQUIRE(first <= second && second < r->n, "measured qubit range out of bounds");
    if (nshots == 0) return QCS_OK;
    const uint64_t mask = all ? ~0ull : range_mask(first, second);
    const uint32_t shift = all ? 0 : first;
    if (nshots == 1) {
        // single-shot shortcut (QubitRegister.h:233-238, 335-340)
        uint64_t m = 0;
        const int e = measure_no_collapse_impl(r, all, first, second, u[0], &m, dev);
        if (e) return e;
        outcomes[0] = all ? m : (m & mask) >> shift;
        return QCS_OK;
    }
    // truncation point: first i with cdf(i) > 1 - DBL_EPSILON (QubitRegister.h:250-253)
    const double thr = 1.0 - DBL_EPSILON;
    uint64_t t = 0;
    int e = cdf_search(r, 1, &thr, 0, 1, r->A, &t, 1, dev);
    if (e) return e;
    const uint64_t len = t >= r->A ? r->A : t + 1;
    // prob = 1 - u per shot (QubitRegister.h:262), formed on the device
    e = cdf_search(r, nshots, u, 1, 0, r->A, outcomes, 0, dev);  // the CDF just built
    if (e) return e;
    for (uint64_t s = 0; s < nshots; ++s) {
        const uint64_t v = outcomes[s] < len ? outcomes[s] : len;  // lower_bound on [0, len)
        outcomes[s] = all ? v : (v & mask) >> shift;
    }
    return QCS_OK;
}
}  // namespace qcs

extern "C" {

int qcs_measure_all(qcs_register* r, double u, uint64_t* outcome) {
    QCS_REQUIRE(r && outcome, "null argument");
    const double prob = 1. - u;  // QubitRegister.h:171
    uint64_t st = 0;
    int e = cdf_search(r, 1, &prob, 0, 0, r->A - 1, &st, 1);
    if (e) return e;
    e = kfill(r, 0, 0);
    if (!e) e = kset_one(r, st, 1, 0);
    *outcome = st;
    return e;
}

int qcs_measure(qcs_register* r, uint32_t first, uint32_t second, double u, uint64_t* outcome) {
    QCS_REQUIRE(r && outcome, "null argument");
    QCS_REQUIRE(first <= second && second < r->n, "measured qubit range out of bounds");
    const double prob = 1. - u;  // QubitRegister.h:210
    uint64_t pstate = 0;
    int e = cdf_search(r, 1, &prob, 0, 0, 0, &pstate, 1);
    if (e) return e;
    const uint64_t mask = range_mask(first, second);
    const uint64_t meas = pstate & mask;
    double accum = 0;
    e = rsum_sq(r, first, second - first + 1, meas, &accum);
    if (e) return e;
    e = kcollapse(r, mask, meas, 1. / std::sqrt(accum));
    *outcome = meas >> first;
    return e;
}

int qcs_measure_no_collapse(qcs_register* r, int all, uint32_t first, uint32_t second, double u,
                            uint64_t* outcome) {
    return measure_no_collapse_impl(r, all, first, second, u, outcome, nullptr);
}

int qcs_sample(qcs_register* r, int all, uint32_t first, uint32_t second, uint64_t nshots, const double* u,
               uint64_t* outcomes) {
    return sample_impl(r, all, first, second, nshots, u, outcomes, nullptr);
}

int qcs_qubit_probability(qcs_register* r, uint32_t q, double* p) {
    QCS_REQUIRE(r && p, "null argument");
    QCS_REQUIRE(q < r->n, "Qubit number is too high");
    return rsum_sq(r, q, 1, 1ull << q, p);
}

int qcs_norm2(qcs_register* r, double* n2) {
    QCS_REQUIRE(r && n2, "null argument");
    return rsum_sq(r, 0, 0, 0, n2);
}

int qcs_normalize(qcs_register* r) {
    QCS_REQUIRE(r, "null register");
    double n2 = 0;
    const int e = rsum_sq(r, 0, 0, 0, &n2);
    if (e) return e;
    const double nrm = std::sqrt(n2);
    if (nrm < 1e-20) return QCS_OK;  // QubitRegister.h:127
    return kscale(r, 1. / nrm, 0);
}

int qcs_inner(qcs_register* a, qcs_register* b, double* re_im) {
    QCS_REQUIRE(a && b && re_im, "null argument");
    QCS_REQUIRE(a->A == b->A, "registers differ in size");
    return rinner(a, b, re_im, nullptr);
}

int qcs_compare(qcs_register* a, qcs_register* b, double* max_abs, double* inner_re_im) {
    QCS_REQUIRE(a && b && max_abs && inner_re_im, "null argument");
    QCS_REQUIRE(a->A == b->A, "registers differ in size");
    return rinner(a, b, inner_re_im, max_abs);
}

int qcs_ising_energy(qcs_register* r, uint32_t n_edges, const uint32_t* ei, const uint32_t* ej, const double* J,
                     const double* h, double* energy) {
    QCS_REQUIRE(r && energy && h && (n_edges == 0 || (ei && ej && J)), "null argument");
    for (uint32_t k = 0; k < n_edges; ++k) QCS_REQUIRE(ei[k] < r->n && ej[k] < r->n, "edge out of range");
    return rising(r, n_edges, ei, ej, J, h, energy);
}

int qcs_synchronize(qcs_register* r) {
    QCS_REQUIRE(r, "null register");
    return hipStreamSynchronize(r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "sync failed");
}

int qcs_timer_begin(qcs_register* r) {
    QCS_REQUIRE(r, "null register");
    return hipEventRecord(r->ev0, r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "event record");
}

int qcs_timer_end(qcs_register* r, float* ms) {
    QCS_REQUIRE(r && ms, "null argument");
    if (hipEventRecord(r->ev1, r->stream) != hipSuccess || hipEventSynchronize(r->ev1) != hipSuccess ||
        hipEventElapsedTime(ms, r->ev0, r->ev1) != hipSuccess)
        return set_error(QCS_EDEVICE, "event timing failed");
    return QCS_OK;
}

}  // extern "C"
