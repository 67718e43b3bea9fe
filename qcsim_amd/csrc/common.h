// common.h - internal types of the gfx950 statevector engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "qcsim_amd.h"

// Device-side complex types: amplitudes are interleaved (re, im) pairs,
// 16 B (complex128) or 8 B (complex64), so one amplitude is one
// global_load_dwordx4 / dwordx2.
template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using cv = typename cvec<T>::type;

struct hcplx {
    double re, im;
};

// Gate matrices passed by value in kernel arguments (row-major).
template <typename T, int D> struct MatD {
    T re[D * D];
    T im[D * D];
};

// Positions (ascending) at which zero bits are inserted into an orbit index
// to obtain the index of the orbit's first amplitude.
struct Ins {
    int n;
    uint8_t p[6];
};

__device__ __forceinline__ uint64_t insert_zeros(uint64_t o, const Ins& ins) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        if (i < ins.n) {
            const uint64_t low = (1ull << ins.p[i]) - 1;
            o = ((o & ~low) << 1) | (o & low);
        }
    }
    return o;
}

// Contiguous run of `len` zero bits inserted at `pos` (range measurements).
__device__ __forceinline__ uint64_t insert_run(uint64_t o, uint32_t pos, uint32_t len) {
    const uint64_t low = (1ull << pos) - 1;
    return ((o & ~low) << len) | (o & low);
}

// Native clang vectors for the nontemporal builtins (HIP's double2/float2
// are structs, which the builtins do not take).
template <typename T> struct nvec;
template <> struct nvec<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct nvec<float> { typedef float type __attribute__((ext_vector_type(2))); };

// Load / store one amplitude; NT = nontemporal (streaming, no reuse).
template <typename T, bool NT>
__device__ __forceinline__ cv<T> ldamp(const cv<T>* p) {
    if (NT) {
        const auto v = __builtin_nontemporal_load(reinterpret_cast<const typename nvec<T>::type*>(p));
        cv<T> r;
        r.x = v.x;
        r.y = v.y;
        return r;
    }
    return *p;
}

template <typename T, bool NT>
__device__ __forceinline__ void stamp(cv<T>* p, cv<T> a) {
    if (NT) {
        typename nvec<T>::type v;
        v.x = a.x;
        v.y = a.y;
        __builtin_nontemporal_store(v, reinterpret_cast<typename nvec<T>::type*>(p));
    } else {
        *p = a;
    }
}

// Block index remap so that the 8 XCDs (blocks are dispatched round-robin
// over them) each stream one contiguous eighth of the state.
// Nontemporal loads in the LDS-staged passes (QFT, dense blocks, bit
// reversal); their stores are always nontemporal. Measured on MI355X at 30
// qubits: plain loads are slower (QFT 44.2 against 41.8 ms, fused 4-qubit
// blocks 4.6 % slower).
#ifndef QCS_STAGED_NT_LOADS
#define QCS_STAGED_NT_LOADS 1
#endif
constexpr bool kStagedNtLoads = QCS_STAGED_NT_LOADS != 0;

template <int R>
__device__ __forceinline__ uint64_t block_id() {
    // R = 0 or 1: launch order. R = 8: block b -> region b % 8 (the XCD it
    // is dispatched to), so each XCD walks one contiguous eighth. R > 8:
    // R contiguous regions, R / 8 per XCD, visited in parallel.
    if (R > 1 && (gridDim.x % R) == 0) return uint64_t(blockIdx.x % R) * (gridDim.x / R) + (blockIdx.x / R);
    return blockIdx.x;
}

template <typename T>
__device__ __forceinline__ cv<T> cmul(T ar, T ai, cv<T> x) {
    cv<T> r;
    r.x = ar * x.x - ai * x.y;
    r.y = ar * x.y + ai * x.x;
    return r;
}

// FP64 reductions over waves of 64 and blocks of 256 (fixed order).
namespace qcs {
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block (256 threads = 4 waves) sum; result valid in thread 0.
__device__ __forceinline__ double block_sum256(double v, double* lds4) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) t = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
    __syncthreads();
    return t;
}

template <typename T>
__device__ __forceinline__ double norm2(cv<T> a) {
    const double x = double(a.x), y = double(a.y);
    return x * x + y * y;
}

}  // namespace qcs

// Device parameter space of a program being captured: uploads of dense
// matrices go to dev + used (copied from the host mirror once the capture
// ends) instead of the staging ring.
struct ParamSink {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0, used = 0;
};

// Engine register: one HBM buffer of 2^n amplitudes, updated in place,
// plus FP64 reduction scratch and its own HIP stream.
struct qcs_register {
    uint32_t n = 0;
    uint64_t A = 0;
    int prec = QCS_COMPLEX128;
    int device = 0;
    void* psi = nullptr;        // 2^n amplitudes
    double* red = nullptr;      // reduction partials (kRedDoubles)
    double* chunk = nullptr;    // per-chunk CDF (A / kChunk doubles, >= 1)
    uint64_t* dout = nullptr;   // search results / small device outputs
    // Parameter staging ring for dense matrices: kMatSlots slots of
    // mat_bytes each, pinned on the host and mirrored on the device. A host
    // slot is refilled only after the event of its previous copy completed.
    static constexpr int kMatSlots = 8;
    void* mat = nullptr;        // device slots
    void* hmat = nullptr;       // pinned host slots
    hipEvent_t mat_ev[kMatSlots] = {};
    uint32_t mat_next = 0;
    size_t mat_bytes = 0;
    double* hpin = nullptr;     // pinned host scalars
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t launches = 0;
    bool dense_no_mfma = false;  // route dense blocks to the non-MFMA kernel
    ParamSink* sink = nullptr;   // set while a program is captured
    // Transfer buffer for host vectors streamed through the device in
    // chunks (complex64 uploads/downloads, diagonal operators, inner
    // products with a host vector, dense operator panels): allocated on
    // first use, at most kXferBytes, kept until the register is destroyed.
    void* xfer = nullptr;
    size_t xfer_bytes = 0;
    uint32_t last_fuse_width = 0;  // width the last qcs_apply_circuit used

    size_t amp_bytes() const { return prec == QCS_COMPLEX128 ? 16 : 8; }
};

namespace qcs {

// A launch covers at most kMaxBlocks blocks of 256 (2^30 work-items: a
// grid's size in work-items must stay below 2^32); larger one-shot grids
// (complex64 states of 33 qubits and up) run as several launches.
constexpr uint64_t kMaxBlocks = 1ull << 22;
constexpr int kRedBlocks = 4096;     // fixed grid of reductions (deterministic)
constexpr int kRedThreads = 256;
// Pauli-string expectations: at most kPauliTerms terms per pass, partials
// of at most kPauliBlocks blocks.
constexpr int kPauliTerms = 16;
constexpr int kPauliBlocks = 2048;
// reduction scratch per register: kRedBlocks * 3 partials, or the Pauli
// partials, then 64 result slots
constexpr int kRedDoubles = (kRedBlocks * 3 > kPauliTerms * kPauliBlocks ? kRedBlocks * 3 : kPauliTerms * kPauliBlocks) + 64;
constexpr int kRedResult = kRedDoubles - 64;  // first result slot
constexpr size_t kXferBytes = size_t(64) << 20;
constexpr uint32_t kChunkLog = 11;   // CDF chunk = 2048 amplitudes
constexpr int kScanThreads = 256;

int set_error(int code, const char* fmt, ...);
int check_launch(qcs_register* r, const char* what);
// Copies `bytes` (<= mat_bytes) from host memory into the next staging slot
// and enqueues its upload on the register's stream; *dev receives the device
// slot. `src` may be released as soon as the call returns.
int stage_upload(qcs_register* r, const void* src, size_t bytes, const void** dev);

// Launch variant of the streaming gate kernels: orbits per thread (1, 2, 4),
// nontemporal access and XCD-aware block order. Chosen per (precision,
// qubit count, lowest touched qubit) by the autotuner (autotune.hip), with a
// measured default when untuned.
struct Variant {
    uint8_t p, nt, xcd;
    uint8_t sp = 0;  // k_1q only: log2 distance of a thread's orbits (0: kThreads)
    uint8_t x = 0;   // k_1q only: one amplitude per lane, partner 1: through LDS, 2: by lane shuffle
};
Variant pick_variant(const qcs_register* r, uint32_t lowbit, bool k1q = false);
int variant_count();
Variant variant_at(int i);
void set_tuned(int prec, uint32_t n, uint32_t q, int vid);

// gate_kernels.hip
int k1q_variant(qcs_register* r, int mode, uint32_t target, const uint32_t* ctrls, int nctrl,
                const hcplx* m, Variant v);
int k1q(qcs_register* r, int mode, uint32_t target, const uint32_t* ctrls, int nctrl,
        const hcplx* m /* 4 */);
int kphase(qcs_register* r, const uint32_t* bits, int nbits, hcplx ph);
int kswap(qcs_register* r, uint32_t a, uint32_t b, const uint32_t* ctrls, int nctrl, hcplx ph,
          bool phased);
int k2q(qcs_register* r, uint32_t t0, uint32_t t1, const uint32_t* ctrls, int nctrl,
        const hcplx* m /* 16 */);
int k3q(qcs_register* r, uint32_t t0, uint32_t t1, uint32_t t2, const hcplx* m /* 64 */);
int kdiag_full(qcs_register* r, const double* diag_dev);
int kfill(qcs_register* r, double re, double im);
int kscale(qcs_register* r, double re, double im);
int kset_one(qcs_register* r, uint64_t idx, double re, double im);

// dense_kernels.hip: dense k-qubit blocks (k <= 5) on arbitrary qubits
int kdense(qcs_register* r, uint32_t k, const uint32_t* qubits, const double* mat_host);

// reduce_kernels.hip (results in FP64)
int rsum_sq(qcs_register* r, uint32_t first, uint32_t len, uint64_t val, double* out);
int rinner(qcs_register* a, qcs_register* b, double* re_im, double* max_abs);
// Device buffers for the shots of a CDF search: `cap` values in, `cap`
// outcomes out. A search without them allocates its own, stream-ordered.
struct ShotDev {
    double* in = nullptr;
    uint64_t* out = nullptr;
    uint64_t cap = 0;
};
// CDF search: out[s] = first index whose cumulative probability reaches
// prob_s (exceeds it with `strict`), or `dflt`, where prob_s = vals_host[s],
// or 1 - vals_host[s] (computed on the device) with `one_minus`. build_cdf = 0
// reuses the chunk CDF built by the previous search on this register, which
// is only valid if the state has not changed since (the norms and rsum_sq
// also use that buffer). `dev` (may be null) supplies the shot buffers.
int cdf_search(qcs_register* r, uint64_t nshots, const double* vals_host, int one_minus, int strict,
               uint64_t dflt, uint64_t* out_host, int build_cdf, const ShotDev* dev = nullptr);
// qcs_measure_no_collapse / qcs_sample with the shot buffers of `dev`
// (null: allocated per call)
int measure_no_collapse_impl(qcs_register* r, int all, uint32_t first, uint32_t second, double u,
                             uint64_t* outcome, const ShotDev* dev);
int sample_impl(qcs_register* r, int all, uint32_t first, uint32_t second, uint64_t nshots, const double* u,
                uint64_t* outcomes, const ShotDev* dev);
int kcollapse(qcs_register* r, uint64_t mask, uint64_t val, double scale);
int rising(qcs_register* r, uint32_t ne, const uint32_t* ei, const uint32_t* ej,
           const double* J, const double* h, double* energy);

// qft.hip
int qft(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse);

// autotune.hip: the fused sequence of the cheapest fusion width
int pick_fusion(qcs_register* r, const qcs_gate* gates, uint32_t n, std::vector<qcs_gate>& out, uint32_t* width);

// expect_kernels.hip
// Makes r->xfer at least `bytes` long (at most kXferBytes).
int ensure_xfer(qcs_register* r, size_t bytes);
int fetch(qcs_register* r, const double* dev, int n, double* host);
int pauli_expectation(qcs_register* r, uint32_t nterms, const uint64_t* xmask, const uint64_t* zmask, double* out);
int inner_host(qcs_register* r, const double* amps, double* re_im);
int apply_diagonal_host(qcs_register* r, const double* diag);
int apply_matrix_host(qcs_register* r, const double* mat);
int convert_upload(qcs_register* r, const double* amps);
int convert_download(qcs_register* r, double* amps);

}  // namespace qcs
