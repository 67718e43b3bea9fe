// dense_kernels.hip - dense k-qubit blocks (k <= 5) on arbitrary qubits.
//
// A block is a 2^k x 2^k complex matrix acting on the amplitudes that differ
// only in the block's k qubits (local bit i <-> qubits[i]); this is the
// Kronecker embedding of SimpleGates.h:211-232/311-333/411-434 generalised
// to k bits. k <= 3 uses the gate kernels; k = 4, 5 stage the matrix in LDS
// and keep the orbit's 2^k amplitudes in registers.
#include <algorithm>
#include <vector>

#include "common.h"

namespace qcs {

template <int K> struct QBits {
    uint64_t b[K];
};

template <typename T, int K>
__global__ __launch_bounds__(256) void k_dense(cv<T>* __restrict__ psi, uint64_t norb, Ins ins, QBits<K> qb,
                                              const T* __restrict__ mat) {
    constexpr int D = 1 << K;
    __shared__ T sre[D * D], sim[D * D];
    for (int i = threadIdx.x; i < D * D; i += 256) {
        sre[i] = mat[i];
        sim[i] = mat[D * D + i];
    }
    __syncthreads();
    const uint64_t o = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (o >= norb) return;
    const uint64_t i0 = insert_zeros(o, ins);
    cv<T> x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        uint64_t idx = i0;
#pragma unroll
        for (int q = 0; q < K; ++q) idx |= (c >> q & 1) ? qb.b[q] : 0;
        x[c] = psi[idx];
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const T mr = sre[r * D + c], mi = sim[r * D + c];
            yr += mr * x[c].x - mi * x[c].y;
            yi += mr * x[c].y + mi * x[c].x;
        }
        uint64_t idx = i0;
#pragma unroll
        for (int q = 0; q < K; ++q) idx |= (r >> q & 1) ? qb.b[q] : 0;
        cv<T> y;
        y.x = yr;
        y.y = yi;
        psi[idx] = y;
    }
}

template <typename T, int K>
static int kdense_big(qcs_register* r, const uint32_t* qubits, const double* mat_host) {
    constexpr int D = 1 << K;
    // matrix converted to T, split re[] then im[], uploaded from pageable
    // memory (the copy has consumed the host buffer when the call returns).
    std::vector<T> hm(2 * D * D);
    for (int i = 0; i < D * D; ++i) {
        hm[i] = T(mat_host[2 * i]);
        hm[D * D + i] = T(mat_host[2 * i + 1]);
    }
    if (hipMemcpyAsync(r->mat, hm.data(), sizeof(T) * hm.size(), hipMemcpyHostToDevice, r->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "dense matrix upload failed");
    Ins ins{};
    uint32_t tmp[K];
    QBits<K> qb;
    for (int i = 0; i < K; ++i) {
        tmp[i] = qubits[i];
        qb.b[i] = 1ull << qubits[i];
    }
    std::sort(tmp, tmp + K);
    ins.n = K;
    for (int i = 0; i < K; ++i) ins.p[i] = uint8_t(tmp[i]);
    const uint64_t norb = r->A >> K;
    const unsigned g = unsigned((norb + 255) / 256);
    hipLaunchKernelGGL((k_dense<T, K>), dim3(g), dim3(256), 0, r->stream, static_cast<cv<T>*>(r->psi), norb, ins,
                       qb, static_cast<const T*>(r->mat));
    return check_launch(r, "k_dense");
}

int kdense(qcs_register* r, uint32_t k, const uint32_t* qubits, const double* mat_host) {
    const int D = 1 << k;
    if (k <= 3) {
        hcplx m[64];
        for (int i = 0; i < D * D; ++i) m[i] = hcplx{mat_host[2 * i], mat_host[2 * i + 1]};
        if (k == 1) return k1q(r, 0, qubits[0], nullptr, 0, m);
        if (k == 2) return k2q(r, qubits[0], qubits[1], nullptr, 0, m);
        return k3q(r, qubits[0], qubits[1], qubits[2], m);
    }
    const bool f64 = r->prec == QCS_COMPLEX128;
    if (k == 4) return f64 ? kdense_big<double, 4>(r, qubits, mat_host) : kdense_big<float, 4>(r, qubits, mat_host);
    return f64 ? kdense_big<double, 5>(r, qubits, mat_host) : kdense_big<float, 5>(r, qubits, mat_host);
}

}  // namespace qcs
