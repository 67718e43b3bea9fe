// dense_kernels.hip - dense k-qubit blocks (k <= 5) on arbitrary qubits.
//
// A block is a 2^k x 2^k complex matrix acting on the amplitudes that differ
// only in the block's k qubits (local bit i <-> qubits[i]); this is the
// Kronecker embedding of SimpleGates.h:211-232/311-333/411-434 generalised
// to k bits. k <= 3 uses the gate kernels. k = 4, 5 run on the matrix cores
// (k_dense_mfma): a workgroup stages 64 orbits ([Re x; Im x] as 2D x 64
// reals) in LDS and each wave multiplies 16 of them with
// v_mfma_f64_16x16x4f64 (complex128) or v_mfma_f32_16x16x4f32 (complex64),
// as three real D x D products (Gauss's product of complex numbers, applied
// to matrices), the factors' fragments held in registers. 8 * 2^k FLOP per
// amplitude for the complex product; the MFMA units execute 6 * 2^k. k_dense (matrix in LDS, orbit in registers, no MFMA) is the
// fallback for states too small to tile.
#include <algorithm>
#include <cstdlib>
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

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = mf_d4;
    __device__ static acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // output element (lane, reg) -> row within the 16x16 tile
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    using acc_t = mf_f4;
    __device__ static acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// 16-byte global accesses: one complex128 or two adjacent complex64
// amplitudes per lane (MI355X streams 16 B lanes faster than 8 B ones).
template <typename T> struct V16;
template <> struct V16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct V16<float> { typedef float type __attribute__((ext_vector_type(4))); };

template <typename T>
__device__ __forceinline__ typename V16<T>::type ld16(const cv<T>* p) {
    const auto* q = reinterpret_cast<const typename V16<T>::type*>(p);
    if (kStagedNtLoads) return __builtin_nontemporal_load(q);
    return *q;
}

template <typename T>
__device__ __forceinline__ void st16(cv<T>* p, typename V16<T>::type v) {
    __builtin_nontemporal_store(v, reinterpret_cast<typename V16<T>::type*>(p));
}

// Tile geometry: the block's K qubits plus the 6 lowest other bits form the
// 11 (K=5) / 10 (K=4) "tile positions", ascending. Element e of a tile
// deposits its bits into those positions, so consecutive lanes touch
// consecutive addresses wherever the positions are contiguous from bit 0.
struct TileGeom {
    uint32_t npos;
    uint8_t pos[12];
    uint8_t role[12];  // < 0x80: local block bit; 0x80 | j: orbit bit j
};

// complex64: at most 128 VGPRs (four waves per SIMD, what its LDS tiles
// allow); the compiler then keeps the accumulators in VGPRs without spills
// (147 registers and three waves otherwise; 5-qubit blocks on 32 qubits
// measured 2-6 % faster). complex128 is held to two waves by its 66 KiB of
// LDS tiles and keeps the default.
template <typename T, int K, int NC>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1) void k_dense_mfma(cv<T>* __restrict__ psi, uint64_t ntiles, TileGeom tg,
                                                    const T* __restrict__ R, int xcd) {
    // NC groups of 64 orbit columns per tile; wave w multiplies columns
    // 64 g + 16 w .. 64 g + 16 w + 15 of every group g
    constexpr int D = 1 << K, RD = 2 * D, NB = 64 * NC, E = D * NB, EPT = E / 256, LDX = NB + 2;
    // VP amplitudes per 16-byte access: elements VP e' .. VP e' + VP - 1 of
    // the tile are adjacent in memory, since tile position 0 is always
    // qubit 0 (a block qubit or the lowest orbit bit)
    constexpr int VP = sizeof(T) == 4 ? 2 : 1, NV = EPT / VP;
    using vt = typename V16<T>::type;
    constexpr int NT = D / 16, NS = D / 4;
    // sX holds a tile's inputs, sY its results: a block synchronises twice
    // per tile (inputs written; results written), since results never
    // overwrite inputs that another wave may still be reading.
    __shared__ __attribute__((aligned(16))) T sX[RD * LDX];
    __shared__ __attribute__((aligned(16))) T sY[RD * LDX];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // per-thread element map (the same for every tile): the first element
    // of each access; the second (VP = 2) differs in element bit 0
    uint64_t off[NV];
    int row[NV], col[NV];
    const int rb = (tg.role[0] & 0x80) ? 0 : 1 << tg.role[0];
    const int cb = (tg.role[0] & 0x80) ? 1 << (tg.role[0] & 0x7f) : 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        const uint32_t e = uint32_t(VP * (tid + 256 * j));
        uint64_t o = 0;
        int c = 0, ol = 0;
        for (uint32_t p = 0; p < tg.npos; ++p) {
            const uint32_t bit = (e >> p) & 1u;
            o |= uint64_t(bit) << tg.pos[p];
            const uint8_t rl = tg.role[p];
            if (rl & 0x80) ol |= int(bit) << (rl & 0x7f);
            else c |= int(bit) << rl;
        }
        off[j] = o;
        row[j] = c;
        col[j] = ol;
    }
    // Fragments of the three D x D factors (Re M, Im M, Re M + Im M):
    // af[m][t][s] = factor m at row 16 t + lane % 16, column 4 s + lane / 16
    T af[3][NT][NS];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                af[m][t][s] = R[(m * D + 16 * t + (lane & 15)) * D + 4 * s + (lane >> 4)];
    // The next tile's amplitudes are loaded into registers while the
    // current one is multiplied (tiles are disjoint and owned by one block).
    auto tile_base = [&](uint64_t tile) {
        uint64_t b = tile;
        for (uint32_t p = 0; p < tg.npos; ++p) {
            const uint64_t low = (1ull << tg.pos[p]) - 1;
            b = ((b & ~low) << 1) | (b & low);
        }
        return b;
    };
    vt pre[NV];
    // with xcd, the blocks of one XCD (dispatched round-robin over 8) take
    // consecutive tiles
    uint64_t tile = xcd ? block_id<8>() : blockIdx.x, base = 0;
    if (tile < ntiles) {
        base = tile_base(tile);
#pragma unroll
        for (int j = 0; j < NV; ++j) pre[j] = ld16<T>(psi + (base | off[j]));
    }
    for (; tile < ntiles; tile += gridDim.x) {
        const uint64_t cur = base;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            sX[row[j] * LDX + col[j]] = pre[j][0];
            sX[(D + row[j]) * LDX + col[j]] = pre[j][1];
            if constexpr (VP == 2) {
                sX[(row[j] | rb) * LDX + (col[j] | cb)] = pre[j][2];
                sX[(D + (row[j] | rb)) * LDX + (col[j] | cb)] = pre[j][3];
            }
        }
        const uint64_t next = tile + gridDim.x;
        if (next < ntiles) {
            base = tile_base(next);
#pragma unroll
            for (int j = 0; j < NV; ++j) pre[j] = ld16<T>(psi + (base | off[j]));
        }
        __syncthreads();
        // Three real products per complex one: with M = A + iB and
        // x = X + iY, P1 = A X, P2 = B Y, P3 = (A + B)(X + Y); then
        // Re Mx = P1 - P2 and Im Mx = P3 - P1 - P2.
        using acc_t = typename Mfma<T>::acc_t;
        acc_t p1[NC][NT], p2[NC][NT], p3[NC][NT];
#pragma unroll
        for (int g = 0; g < NC; ++g)
#pragma unroll
            for (int t = 0; t < NT; ++t) p1[g][t] = p2[g][t] = p3[g][t] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int g = 0; g < NC; ++g)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int c = 64 * g + 16 * w + (lane & 15), rr = 4 * s + (lane >> 4);
                const T bx = sX[rr * LDX + c], by = sX[(D + rr) * LDX + c], bs = bx + by;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    p1[g][t] = Mfma<T>::step(af[0][t][s], bx, p1[g][t]);
                    p2[g][t] = Mfma<T>::step(af[1][t][s], by, p2[g][t]);
                    p3[g][t] = Mfma<T>::step(af[2][t][s], bs, p3[g][t]);
                }
            }
#pragma unroll
        for (int g = 0; g < NC; ++g)
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rw = 16 * t + Mfma<T>::row(lane, r), c = 64 * g + 16 * w + (lane & 15);
                    sY[rw * LDX + c] = p1[g][t][r] - p2[g][t][r];
                    sY[(D + rw) * LDX + c] = p3[g][t][r] - p1[g][t][r] - p2[g][t][r];
                }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            vt v;
            v[0] = sY[row[j] * LDX + col[j]];
            v[1] = sY[(D + row[j]) * LDX + col[j]];
            if constexpr (VP == 2) {
                v[2] = sY[(row[j] | rb) * LDX + (col[j] | cb)];
                v[3] = sY[(D + (row[j] | rb)) * LDX + (col[j] | cb)];
            }
            st16<T>(psi + (cur | off[j]), v);
        }
        // no barrier here: the next tile's inputs go to sX (no wave reads it
        // any more, all passed the barrier above), and its results reach sY
        // only after the next barrier, once every wave has stored this one
    }
}

template <typename T, int K>
static int kdense_mfma(qcs_register* r, const uint32_t* qubits, const double* mat_host) {
    // complex128: 64-orbit tiles (NC = 1); 128-orbit tiles measured no
    // faster for 4-qubit blocks at 30 qubits (6.77 against 6.6-6.9 ms per
    // launch) at lower occupancy. complex64: 128-orbit tiles (NC = 2), so
    // that a tile's runs are at least 128 amplitudes (1 KiB) long: the
    // 80-layer 32-qubit circuit fused to width 5 took 4.41 s against 4.56 s.
    constexpr int NC = sizeof(T) == 4 ? 2 : 1, OB = NC == 2 ? 7 : 6;
    constexpr int D = 1 << K;
    // factors Re M, Im M and Re M + Im M, stacked (3 D x D, row-major)
    std::vector<T> R(size_t(3) * D * D);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            const double re = mat_host[2 * (i * D + j)], im = mat_host[2 * (i * D + j) + 1];
            R[size_t(i) * D + j] = T(re);
            R[size_t(D + i) * D + j] = T(im);
            R[size_t(2 * D + i) * D + j] = T(re + im);
        }
    const void* dmat = nullptr;
    if (const int e = stage_upload(r, R.data(), sizeof(T) * R.size(), &dmat)) return e;
    TileGeom tg{};
    uint64_t gmask = 0;
    for (int i = 0; i < K; ++i) gmask |= 1ull << qubits[i];
    // Orbit bits of a tile: the lowest OB bits outside the block, except
    // that when those make the tile's run from bit 0 between 128 and 1024
    // amplitudes long, the last one moves to the first free bit at or above
    // 12 (the tile then covers two ranges that far apart). Measured on
    // MI355X at 30 qubits: 6-15 % faster for such tiles, slower for runs of
    // 64 or 2048 amplitudes.
    uint32_t orb[OB], no = 0;
    for (uint32_t b = 0; no < uint32_t(OB); ++b)
        if (!(gmask >> b & 1)) orb[no++] = b;
    {
        uint32_t run = 0;  // log2 of the contiguous run from bit 0
        while (run < 64 && ((gmask >> run & 1) || std::find(orb, orb + OB, run) != orb + OB)) ++run;
        uint32_t b = std::max(12u, orb[OB - 1] + 1);
        while (b < r->n && (gmask >> b & 1)) ++b;
        if (run >= 7 && run <= 10 && b < r->n) orb[OB - 1] = b;
    }
    uint32_t all[K + OB];
    for (int i = 0; i < K; ++i) all[i] = qubits[i];
    for (int j = 0; j < OB; ++j) all[K + j] = orb[j];
    std::sort(all, all + K + OB);
    tg.npos = K + OB;
    for (uint32_t p = 0; p < tg.npos; ++p) {
        tg.pos[p] = uint8_t(all[p]);
        tg.role[p] = 0xff;
        for (int i = 0; i < K; ++i)
            if (qubits[i] == all[p]) tg.role[p] = uint8_t(i);
        for (int j = 0; j < OB; ++j)
            if (orb[j] == all[p]) tg.role[p] = uint8_t(0x80 | j);
    }
    // complex64 tiles load amplitude pairs (2v, 2v + 1): position 0 must be bit 0
    if (tg.pos[0] != 0) return set_error(QCS_EDEVICE, "dense tile geometry without qubit 0");
    const uint64_t ntiles = r->A >> (K + OB);
    const unsigned grid = unsigned(ntiles < 4096 ? ntiles : 4096);
    static const int xcd = [] {
        const char* v = std::getenv("QCS_DENSE_XCD");  // measurement switch, default on
        return v ? std::atoi(v) : 1;
    }();
    hipLaunchKernelGGL((k_dense_mfma<T, K, NC>), dim3(grid), dim3(256), 0, r->stream, static_cast<cv<T>*>(r->psi),
                       ntiles, tg, static_cast<const T*>(dmat), xcd);
    return check_launch(r, "k_dense_mfma");
}

template <typename T, int K>
static int kdense_big(qcs_register* r, const uint32_t* qubits, const double* mat_host) {
    constexpr uint32_t kMinQubits = K + (sizeof(T) == 4 ? 7 : 6);  // one MFMA tile
    if (r->n >= kMinQubits && !r->dense_no_mfma) return kdense_mfma<T, K>(r, qubits, mat_host);
    constexpr int D = 1 << K;
    // matrix converted to T, split re[] then im[], through the staging ring
    std::vector<T> hm(2 * D * D);
    for (int i = 0; i < D * D; ++i) {
        hm[i] = T(mat_host[2 * i]);
        hm[D * D + i] = T(mat_host[2 * i + 1]);
    }
    const void* dmat = nullptr;
    if (const int e = stage_upload(r, hm.data(), sizeof(T) * hm.size(), &dmat)) return e;
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
                       qb, static_cast<const T*>(dmat));
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
