// vqe.hip - batched VQE energies: many small registers of uneven size, each
// prepared by the reference's hardware-efficient ansatz and scored with the
// exact FP64 energy of a Pauli-string Hamiltonian, in one launch per size
// class.
//
// Ansatz (VQE::PauliStringVQE::Ansatz, VQE.h:78-114), from |0...0>: with
// parameters p[0..P) (nothing at all when P < n), the first layer puts
// Ry(p[2q]) then Rz(p[2q+1]) on qubits q = 0, 1, ... while two parameters
// remain; then, while two parameters remain, an entangling ladder
// (alternately CNOT(target q+1, control q) for q = 0..n-2 plus, for n > 2,
// CNOT(target 0, control n-1); and CNOT(target q-1, control q) for
// q = n-1..1 plus, for n > 2, CNOT(target n-1, control 0)) followed by
// another Ry / Rz layer.
//
// Energy: sum_t c_t <psi|P_t|psi> with P_t = i^{|x_t & z_t|} X^{x_t} Z^{z_t},
// evaluated as in k_pauli (one pair sum per term); for a VQE term this is
// the exact value of the estimate that EstimateEnergy samples (VQE.h:116-142
// after the change of basis of :33-44).
//
// A block owns one instance: its 2^n complex128 amplitudes live in LDS
// (n <= 12: 64 KiB), each gate is one barrier-separated pass over the LDS
// state, and the final state is written to the instance's slot of an arena
// in HBM. Instances are grouped by size class (n <= 8, <= 10, <= 12) so that
// small ones do not reserve the largest LDS footprint.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace qcs {

constexpr uint32_t kVqeMaxQubits = 12;

struct VInst {
    uint64_t off;       // first amplitude of the instance in the arena
    uint32_t n;         // qubits
    uint32_t nterms;    // Hamiltonian terms
    uint32_t toff;      // first term
    uint32_t nparams;   // ansatz parameters (set per evaluation)
    uint32_t poff;      // first parameter
    uint32_t pad;
};

// y = Rz(phi) Ry(theta) x on the pairs (i, i | 2^q) of the LDS state.
__device__ void vqe_rot(double2* s, uint32_t n, uint32_t q, double theta, double phi) {
    double sy, cy, sz, cz;
    sincos(0.5 * theta, &sy, &cy);
    sincos(0.5 * phi, &sz, &cz);
    // Ry = [[cy, -sy], [sy, cy]]; Rz = diag(e^{-i phi/2}, e^{i phi/2})
    const uint32_t half = 1u << (n - 1), qb = 1u << q;
    for (uint32_t p = threadIdx.x; p < half; p += blockDim.x) {
        const uint32_t i0 = ((p >> q) << (q + 1)) | (p & (qb - 1)), i1 = i0 | qb;
        const double2 a = s[i0], b = s[i1];
        const double2 u = make_double2(cy * a.x - sy * b.x, cy * a.y - sy * b.y);
        const double2 v = make_double2(sy * a.x + cy * b.x, sy * a.y + cy * b.y);
        // times e^{-i phi/2} = (cz, -sz) and e^{i phi/2} = (cz, sz)
        s[i0] = make_double2(cz * u.x + sz * u.y, cz * u.y - sz * u.x);
        s[i1] = make_double2(cz * v.x - sz * v.y, cz * v.y + sz * v.x);
    }
    __syncthreads();
}

// CNOT: swap the amplitudes of (i, i | 2^t) where bit c of i is set.
__device__ void vqe_cnot(double2* s, uint32_t n, uint32_t t, uint32_t c) {
    const uint32_t half = 1u << (n - 1), tb = 1u << t, cb = 1u << c;
    for (uint32_t p = threadIdx.x; p < half; p += blockDim.x) {
        const uint32_t i0 = ((p >> t) << (t + 1)) | (p & (tb - 1));
        if (!(i0 & cb)) continue;
        const double2 a = s[i0];
        s[i0] = s[i0 | tb];
        s[i0 | tb] = a;
    }
    __syncthreads();
}

// Fixed-order sum over the block (blockDim 256); valid in every thread.
__device__ __forceinline__ double vqe_block_sum(double v, double* red) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void k_vqe(const VInst* __restrict__ insts, const uint32_t* __restrict__ order,
                                             const double* __restrict__ params, const uint64_t* __restrict__ xm,
                                             const uint64_t* __restrict__ zm, const double* __restrict__ coef,
                                             double2* __restrict__ arena, double* __restrict__ energy) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    double2* s = reinterpret_cast<double2*>(smem_raw);
    __shared__ double red[4];
    const uint32_t id = order[blockIdx.x];
    const VInst I = insts[id];
    const uint32_t n = I.n, A = 1u << n;
    for (uint32_t i = threadIdx.x; i < A; i += blockDim.x) s[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
    __syncthreads();
    const double* p = params + I.poff;
    const uint32_t P = I.nparams;
    if (P >= n) {
        uint32_t pos = 0;
        for (uint32_t q = 0; q < n && pos + 1 < P; ++q, pos += 2) vqe_rot(s, n, q, p[pos], p[pos + 1]);
        bool flip = false;
        while (pos + 1 < P) {
            if (flip) {
                for (uint32_t q = n - 1; q > 0; --q) vqe_cnot(s, n, q - 1, q);
                if (n > 2) vqe_cnot(s, n, n - 1, 0);
            } else {
                for (uint32_t q = 0; q + 1 < n; ++q) vqe_cnot(s, n, q + 1, q);
                if (n > 2) vqe_cnot(s, n, 0, n - 1);
            }
            flip = !flip;
            for (uint32_t q = 0; q < n && pos + 1 < P; ++q, pos += 2) vqe_rot(s, n, q, p[pos], p[pos + 1]);
        }
    }
    double e = 0;
    for (uint32_t t = 0; t < I.nterms; ++t) {
        const uint64_t x = xm[I.toff + t], z = zm[I.toff + t];
        const uint32_t ny = uint32_t(__popcll(x & z));
        double acc = 0;
        if (x == 0) {
            for (uint32_t i = threadIdx.x; i < A; i += blockDim.x) {
                const double2 v = s[i];
                const double w = v.x * v.x + v.y * v.y;
                acc += (__popcll(i & z) & 1) ? -w : w;
            }
        } else {
            const uint32_t h = 63 - __clzll(x);
            for (uint32_t o = threadIdx.x; o < A / 2; o += blockDim.x) {
                const uint32_t i = uint32_t(insert_run(o, h, 1));
                const double2 u = s[i], w = s[i ^ uint32_t(x)];
                const double v = (ny & 1) ? u.x * w.y - u.y * w.x : u.x * w.x + u.y * w.y;
                acc += (__popcll(i & z) & 1) ? -v : v;
            }
        }
        double tot = vqe_block_sum(acc, red);
        if (x != 0) tot *= ((((ny & 1) ? ny - 1 : ny) / 2) & 1) ? -2.0 : 2.0;
        e += coef[I.toff + t] * tot;
    }
    if (threadIdx.x == 0) energy[id] = e;
    for (uint32_t i = threadIdx.x; i < A; i += blockDim.x) stamp<double, true>(arena + I.off + i, s[i]);
}

}  // namespace qcs

using namespace qcs;

struct qcs_vqe {
    uint32_t ninst = 0;
    std::vector<VInst> hinst;
    std::vector<uint32_t> order;      // instances sorted by size class
    uint32_t cls_begin[4] = {0, 0, 0, 0};
    uint32_t cls_bits[3] = {8, 10, 12};
    VInst* dinst = nullptr;
    uint32_t* dorder = nullptr;
    uint64_t* dxm = nullptr;
    uint64_t* dzm = nullptr;
    double* dcoef = nullptr;
    double* dparams = nullptr;
    size_t params_cap = 0;
    double2* arena = nullptr;
    double* denergy = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
};

extern "C" {

int qcs_vqe_destroy(qcs_vqe* v) {
    if (!v) return QCS_OK;
    if (v->stream) (void)hipStreamSynchronize(v->stream);
    for (void* p : {static_cast<void*>(v->dinst), static_cast<void*>(v->dorder), static_cast<void*>(v->dxm),
                    static_cast<void*>(v->dzm), static_cast<void*>(v->dcoef), static_cast<void*>(v->dparams),
                    static_cast<void*>(v->arena), static_cast<void*>(v->denergy)})
        if (p) (void)hipFree(p);
    if (v->ev0) (void)hipEventDestroy(v->ev0);
    if (v->ev1) (void)hipEventDestroy(v->ev1);
    if (v->stream) (void)hipStreamDestroy(v->stream);
    delete v;
    return QCS_OK;
}

int qcs_vqe_create(uint32_t ninst, const uint32_t* nqubits, const uint32_t* nterms, const uint64_t* xmask,
                   const uint64_t* zmask, const double* coef, qcs_vqe** out) {
    if (!out || !ninst || !nqubits || !nterms) return set_error(QCS_EINVAL, "null argument");
    *out = nullptr;
    auto* v = new qcs_vqe;
    v->ninst = ninst;
    v->hinst.resize(ninst);
    uint64_t off = 0, toff = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t n = nqubits[i];
        if (n < 1 || n > kVqeMaxQubits) {
            delete v;
            return set_error(QCS_EINVAL, "instance %u: %u qubits (batched VQE holds 1..%u)", i, n, kVqeMaxQubits);
        }
        VInst& I = v->hinst[i];
        I.off = off;
        I.n = n;
        I.nterms = nterms[i];
        I.toff = uint32_t(toff);
        for (uint32_t t = 0; t < nterms[i]; ++t)
            if ((xmask[toff + t] | zmask[toff + t]) >> n) {
                delete v;
                return set_error(QCS_EINVAL, "instance %u term %u: mask beyond the register", i, t);
            }
        off += uint64_t(1) << n;
        toff += nterms[i];
        if (toff > 0xffffffffull) {
            delete v;
            return set_error(QCS_EINVAL, "too many terms");
        }
    }
    if (toff && (!xmask || !zmask || !coef)) {
        delete v;
        return set_error(QCS_EINVAL, "null term arrays");
    }
    // size classes: n <= 8, <= 10, <= 12
    for (uint32_t c = 0; c < 3; ++c) {
        v->cls_begin[c] = uint32_t(v->order.size());
        for (uint32_t i = 0; i < ninst; ++i) {
            const uint32_t n = v->hinst[i].n;
            const uint32_t lo = c ? v->cls_bits[c - 1] : 0;
            if (n > lo && n <= v->cls_bits[c]) v->order.push_back(i);
        }
    }
    v->cls_begin[3] = uint32_t(v->order.size());
    const size_t nt = std::max<uint64_t>(toff, 1);
    if (hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&v->ev0) != hipSuccess || hipEventCreate(&v->ev1) != hipSuccess) {
        qcs_vqe_destroy(v);
        return set_error(QCS_EDEVICE, "stream creation failed");
    }
    if (hipMalloc(reinterpret_cast<void**>(&v->dinst), sizeof(VInst) * ninst) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->dorder), sizeof(uint32_t) * ninst) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->dxm), sizeof(uint64_t) * nt) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->dzm), sizeof(uint64_t) * nt) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->dcoef), sizeof(double) * nt) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->arena), sizeof(double2) * off) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&v->denergy), sizeof(double) * ninst) != hipSuccess) {
        (void)hipGetLastError();
        qcs_vqe_destroy(v);
        return set_error(QCS_ENOMEM, "batched VQE buffers");
    }
    if ((toff && (hipMemcpy(v->dxm, xmask, sizeof(uint64_t) * toff, hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(v->dzm, zmask, sizeof(uint64_t) * toff, hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(v->dcoef, coef, sizeof(double) * toff, hipMemcpyHostToDevice) != hipSuccess)) ||
        hipMemcpy(v->dorder, v->order.data(), sizeof(uint32_t) * ninst, hipMemcpyHostToDevice) != hipSuccess) {
        qcs_vqe_destroy(v);
        return set_error(QCS_EDEVICE, "batched VQE upload failed");
    }
    *out = v;
    return QCS_OK;
}

int qcs_vqe_evaluate(qcs_vqe* v, const uint32_t* nparams, const double* params, double* energy) {
    if (!v || !nparams || !energy) return set_error(QCS_EINVAL, "null argument");
    uint64_t poff = 0;
    for (uint32_t i = 0; i < v->ninst; ++i) {
        v->hinst[i].nparams = nparams[i];
        v->hinst[i].poff = uint32_t(poff);
        poff += nparams[i];
        if (poff > 0xffffffffull) return set_error(QCS_EINVAL, "too many parameters");
    }
    if (poff && !params) return set_error(QCS_EINVAL, "null parameters");
    if (poff > v->params_cap) {
        if (hipStreamSynchronize(v->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
        if (v->dparams) (void)hipFree(v->dparams);
        v->dparams = nullptr;
        v->params_cap = 0;
        if (hipMalloc(reinterpret_cast<void**>(&v->dparams), sizeof(double) * poff) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(QCS_ENOMEM, "parameter buffer");
        }
        v->params_cap = poff;
    }
    if ((poff && hipMemcpyAsync(v->dparams, params, sizeof(double) * poff, hipMemcpyHostToDevice, v->stream) !=
                     hipSuccess) ||
        hipMemcpyAsync(v->dinst, v->hinst.data(), sizeof(VInst) * v->ninst, hipMemcpyHostToDevice, v->stream) !=
            hipSuccess ||
        hipEventRecord(v->ev0, v->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "batched VQE upload failed");
    for (uint32_t c = 0; c < 3; ++c) {
        const uint32_t b0 = v->cls_begin[c], cnt = v->cls_begin[c + 1] - b0;
        if (!cnt) continue;
        const size_t shm = sizeof(double2) << v->cls_bits[c];
        hipLaunchKernelGGL(k_vqe, dim3(cnt), dim3(256), shm, v->stream, v->dinst, v->dorder + b0, v->dparams,
                           v->dxm, v->dzm, v->dcoef, v->arena, v->denergy);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(QCS_EDEVICE, "k_vqe launch failed: %s", hipGetErrorString(e));
    }
    if (hipEventRecord(v->ev1, v->stream) != hipSuccess ||
        hipMemcpyAsync(energy, v->denergy, sizeof(double) * v->ninst, hipMemcpyDeviceToHost, v->stream) !=
            hipSuccess ||
        hipStreamSynchronize(v->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "batched VQE evaluation failed");
    (void)hipEventElapsedTime(&v->last_ms, v->ev0, v->ev1);
    return QCS_OK;
}

int qcs_vqe_last_ms(qcs_vqe* v, float* ms) {
    if (!v || !ms) return set_error(QCS_EINVAL, "null argument");
    *ms = v->last_ms;
    return QCS_OK;
}

int qcs_vqe_get_state(qcs_vqe* v, uint32_t inst, double* amps) {
    if (!v || !amps) return set_error(QCS_EINVAL, "null argument");
    if (inst >= v->ninst) return set_error(QCS_EINVAL, "instance out of range");
    const VInst& I = v->hinst[inst];
    if (hipMemcpy(amps, v->arena + I.off, sizeof(double2) << I.n, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(QCS_EDEVICE, "state download failed");
    return QCS_OK;
}

}  // extern "C"
