// gate_kernels.hip - in-place gate kernels for gfx950.
//
// Every kernel enumerates "orbits": the set of amplitudes a gate mixes.
// Orbit o is expanded into its first amplitude index by inserting zero bits
// at the positions of the gate's target and control qubits and OR-ing the
// control mask; the other amplitudes of the orbit differ only in target bits.
// A work-item owns a whole orbit, so updates are in place and race-free, and
// control=0 amplitudes are never read or written (the reference copies them,
// QubitRegisterCalculator.h:377-378,400-401). Consecutive lanes take
// consecutive orbits, so for target qubits >= 6 each wave streams contiguous
// 1 KiB (FP64) runs per orbit member; for low targets a wave's members
// interleave inside one contiguous 2-4 KiB span, still whole cache lines.
// k_1q_split is the exception: there a block owns 128 whole orbits and each
// lane one amplitude of them, exchanged through LDS.
#include <algorithm>
#include <vector>

#include "common.h"

namespace qcs {

constexpr int kThreads = 256;

static Ins make_ins(const uint32_t* pos, int n) {
    Ins ins{};
    uint32_t tmp[6];
    for (int i = 0; i < n; ++i) tmp[i] = pos[i];
    std::sort(tmp, tmp + n);
    ins.n = n;
    for (int i = 0; i < n; ++i) ins.p[i] = uint8_t(tmp[i]);
    return ins;
}

static uint64_t grid_for(uint64_t norb, int per_block) {
    return (norb + per_block - 1) / per_block;
}

// Larger sweeps than kMaxBlocks blocks run as several launches, each told
// its first block.
#define QCS_CHUNKED(NBLK, LAUNCH1)                                                    \
    for (uint64_t boff = 0; boff < (NBLK); boff += kMaxBlocks) {                      \
        const unsigned g = unsigned(std::min<uint64_t>(kMaxBlocks, (NBLK) - boff));   \
        LAUNCH1;                                                                      \
    }

// ---------------------------------------------------------------------------
// 1-qubit update on the orbit {i0, i0|bit} (plus controls in cmask).
// MODE 0: generic 2x2 (K3, K8, K13); 1: antidiagonal (K2, K7, K12);
// 2: diagonal with both entries (K1, K6, K11).
// sh = 0: a thread's P orbits are kThreads apart (one contiguous range per
// block); sh >= 8: they are 2^sh apart, so each block works on P ranges
// that far from each other at once.
template <typename T, int MODE, int P, int NT, int XCD>
__global__ __launch_bounds__(kThreads) void k_1q(cv<T>* __restrict__ psi, uint64_t norb, Ins ins,
                                                 uint64_t cmask, uint64_t bit, MatD<T, 2> m, uint32_t sh,
                                                 uint64_t boff) {
    const uint64_t bb = block_id<XCD>() + boff;
    uint64_t base, step;
    if (sh) {
        const uint32_t lb = sh - 8;  // blocks per range of 2^sh orbits
        base = ((bb >> lb) << (sh + (P == 4 ? 2 : P == 2 ? 1 : 0))) + ((bb & ((1ull << lb) - 1)) << 8) +
               threadIdx.x;
        step = 1ull << sh;
    } else {
        base = bb * (kThreads * P) + threadIdx.x;
        step = kThreads;
    }
    cv<T> a[P], b[P];
    uint64_t i0[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * step;
        i0[j] = insert_zeros(o, ins) | cmask;
        if (o < norb) {
            a[j] = ldamp<T, NT == 1 || NT == 2>(psi + i0[j]);
            b[j] = ldamp<T, NT == 1 || NT == 2>(psi + (i0[j] | bit));
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * step;
        if (o >= norb) continue;
        cv<T> u, v;
        if (MODE == 0) {
            u.x = m.re[0] * a[j].x - m.im[0] * a[j].y + m.re[1] * b[j].x - m.im[1] * b[j].y;
            u.y = m.re[0] * a[j].y + m.im[0] * a[j].x + m.re[1] * b[j].y + m.im[1] * b[j].x;
            v.x = m.re[2] * a[j].x - m.im[2] * a[j].y + m.re[3] * b[j].x - m.im[3] * b[j].y;
            v.y = m.re[2] * a[j].y + m.im[2] * a[j].x + m.re[3] * b[j].y + m.im[3] * b[j].x;
        } else if (MODE == 1) {
            u = cmul<T>(m.re[1], m.im[1], b[j]);
            v = cmul<T>(m.re[2], m.im[2], a[j]);
        } else {
            u = cmul<T>(m.re[0], m.im[0], a[j]);
            v = cmul<T>(m.re[3], m.im[3], b[j]);
        }
        stamp<T, NT == 1 || NT == 3>(psi + i0[j], u);
        stamp<T, NT == 1 || NT == 3>(psi + (i0[j] | bit), v);
    }
}

// The same update with one amplitude per lane: a block owns 128 orbits (256
// amplitudes), each lane loads one amplitude, takes its partner from LDS
// and stores its own result. Loads and stores are one 16 B (FP64) access
// per lane on consecutive addresses, the shape that streams fastest on
// MI355X (a plain in-place sweep of this shape measured 6.6 TB/s against
// 5.9 TB/s for two elements per lane). `e` indexes the amplitudes with the
// control bits removed (zeros inserted by `insc`, control mask OR-ed back);
// the target is bit `tl` of e. For tl < 7 a block takes 256 consecutive
// values of e (partner lane tid ^ 2^tl); otherwise 128 consecutive ones
// with bit tl clear and the same 128 with it set (partner lane tid ^ 128).
template <typename T, int MODE, int NT, int XCD>
__global__ __launch_bounds__(kThreads) void k_1q_split(cv<T>* __restrict__ psi, Ins insc, uint64_t cmask,
                                                       uint32_t tl, MatD<T, 2> m, uint64_t boff) {
    __shared__ cv<T> s[kThreads];
    const uint64_t bb = block_id<XCD>() + boff;
    const int tid = threadIdx.x;
    uint64_t e;
    int h, partner;
    if (tl < 7) {
        e = bb * kThreads + tid;
        h = (tid >> tl) & 1;
        partner = tid ^ (1 << tl);
    } else {
        const uint64_t o = bb * (kThreads / 2) + (tid & (kThreads / 2 - 1));
        const uint64_t low = (1ull << tl) - 1;
        h = tid >> 7;
        e = ((o & ~low) << 1) | (o & low) | (uint64_t(h) << tl);
        partner = tid ^ (kThreads / 2);
    }
    const uint64_t idx = insert_zeros(e, insc) | cmask;
    const cv<T> a = ldamp<T, NT == 1 || NT == 2>(psi + idx);
    cv<T> u;
    if (MODE == 2) {
        u = h ? cmul<T>(m.re[3], m.im[3], a) : cmul<T>(m.re[0], m.im[0], a);
    } else {
        s[tid] = a;
        __syncthreads();
        const cv<T> b = s[partner];
        const cv<T> x0 = h ? b : a, x1 = h ? a : b;
        if (MODE == 0) {
            // row h of the matrix: (m[2h], m[2h + 1])
            const T r0 = h ? m.re[2] : m.re[0], i0 = h ? m.im[2] : m.im[0];
            const T r1 = h ? m.re[3] : m.re[1], i1 = h ? m.im[3] : m.im[1];
            u.x = r0 * x0.x - i0 * x0.y + r1 * x1.x - i1 * x1.y;
            u.y = r0 * x0.y + i0 * x0.x + r1 * x1.y + i1 * x1.x;
        } else {
            u = h ? cmul<T>(m.re[2], m.im[2], x0) : cmul<T>(m.re[1], m.im[1], x1);
        }
    }
    stamp<T, NT == 1 || NT == 3>(psi + idx, u);
}

// k_1q_split with the partner in the same wave, exchanged by a lane shuffle
// (no LDS, no barrier): for tl < 6 a wave takes 64 consecutive values of e
// (partner lane ^ 2^tl); otherwise 32 consecutive orbits, lanes 0..31 with
// the target bit clear and lanes 32..63 with it set (partner lane ^ 32).
template <typename T, int MODE, int NT, int XCD>
__global__ __launch_bounds__(kThreads) void k_1q_shfl(cv<T>* __restrict__ psi, Ins insc, uint64_t cmask,
                                                      uint32_t tl, MatD<T, 2> m, uint64_t boff) {
    const uint64_t bb = block_id<XCD>() + boff;
    const int tid = threadIdx.x, lane = tid & 63;
    uint64_t e;
    int h, mask;
    if (tl < 6) {
        e = bb * kThreads + tid;
        h = (lane >> tl) & 1;
        mask = 1 << tl;
    } else {
        const uint64_t o = (bb * (kThreads / 64) + (tid >> 6)) * 32 + (lane & 31);
        const uint64_t low = (1ull << tl) - 1;
        h = lane >> 5;
        e = ((o & ~low) << 1) | (o & low) | (uint64_t(h) << tl);
        mask = 32;
    }
    const uint64_t idx = insert_zeros(e, insc) | cmask;
    const cv<T> a = ldamp<T, NT == 1 || NT == 2>(psi + idx);
    cv<T> u;
    if (MODE == 2) {
        u = h ? cmul<T>(m.re[3], m.im[3], a) : cmul<T>(m.re[0], m.im[0], a);
    } else {
        cv<T> b;
        b.x = __shfl_xor(a.x, mask, 64);
        b.y = __shfl_xor(a.y, mask, 64);
        const cv<T> x0 = h ? b : a, x1 = h ? a : b;
        if (MODE == 0) {
            const T r0 = h ? m.re[2] : m.re[0], i0 = h ? m.im[2] : m.im[0];
            const T r1 = h ? m.re[3] : m.re[1], i1 = h ? m.im[3] : m.im[1];
            u.x = r0 * x0.x - i0 * x0.y + r1 * x1.x - i1 * x1.y;
            u.y = r0 * x0.y + i0 * x0.x + r1 * x1.y + i1 * x1.x;
        } else {
            u = h ? cmul<T>(m.re[2], m.im[2], x0) : cmul<T>(m.re[1], m.im[1], x1);
        }
    }
    stamp<T, NT == 1 || NT == 3>(psi + idx, u);
}

// Phase on the single amplitude of each orbit whose listed bits are all 1:
// P/S/T/Z (m00 == 1), CZ/CP/CRz-with-unit-m22 (K6), CCZ (K11).
template <typename T, int P, int NT, int XCD>
__global__ __launch_bounds__(kThreads) void k_phase(cv<T>* __restrict__ psi, uint64_t norb, Ins ins,
                                                    uint64_t cmask, T pr, T pi, uint64_t boff) {
    const uint64_t base = (block_id<XCD>() + boff) * (kThreads * P) + threadIdx.x;
    cv<T> a[P];
    uint64_t i0[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * kThreads;
        i0[j] = insert_zeros(o, ins) | cmask;
        if (o < norb) a[j] = ldamp<T, NT == 1 || NT == 2>(psi + i0[j]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * kThreads;
        if (o < norb) stamp<T, NT == 1 || NT == 3>(psi + i0[j], cmul<T>(pr, pi, a[j]));
    }
}

// Exchange psi[i|ba] <-> psi[i|bb], optionally times a phase (SWAP K4,
// iSWAP/iSWAP-dagger K5, Fredkin K10 with its control in cmask).
template <typename T, bool PHASED, int P, int NT, int XCD>
__global__ __launch_bounds__(kThreads) void k_swap(cv<T>* __restrict__ psi, uint64_t norb, Ins ins,
                                                   uint64_t cmask, uint64_t ba, uint64_t bb, T pr,
                                                   T pi, uint64_t boff) {
    const uint64_t base = (block_id<XCD>() + boff) * (kThreads * P) + threadIdx.x;
    cv<T> a[P], b[P];
    uint64_t i0[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * kThreads;
        i0[j] = insert_zeros(o, ins) | cmask;
        if (o < norb) {
            a[j] = ldamp<T, NT == 1 || NT == 2>(psi + (i0[j] | ba));
            b[j] = ldamp<T, NT == 1 || NT == 2>(psi + (i0[j] | bb));
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint64_t o = base + uint64_t(j) * kThreads;
        if (o >= norb) continue;
        if (PHASED) {
            stamp<T, NT == 1 || NT == 3>(psi + (i0[j] | ba), cmul<T>(pr, pi, b[j]));
            stamp<T, NT == 1 || NT == 3>(psi + (i0[j] | bb), cmul<T>(pr, pi, a[j]));
        } else {
            stamp<T, NT == 1 || NT == 3>(psi + (i0[j] | ba), b[j]);
            stamp<T, NT == 1 || NT == 3>(psi + (i0[j] | bb), a[j]);
        }
    }
}

// Generic 4x4 on target bits (b0 = local bit 0, b1 = local bit 1), with
// optional controls: K9 and K14.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_2q(cv<T>* __restrict__ psi, uint64_t norb, Ins ins,
                                                 uint64_t cmask, uint64_t b0, uint64_t b1,
                                                 MatD<T, 4> m, uint64_t boff) {
    const uint64_t o = (uint64_t(blockIdx.x) + boff) * kThreads + threadIdx.x;
    if (o >= norb) return;
    const uint64_t i0 = insert_zeros(o, ins) | cmask;
    uint64_t idx[4] = {i0, i0 | b0, i0 | b1, i0 | b0 | b1};
    cv<T> x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = psi[idx[c]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            yr += m.re[r * 4 + c] * x[c].x - m.im[r * 4 + c] * x[c].y;
            yi += m.re[r * 4 + c] * x[c].y + m.im[r * 4 + c] * x[c].x;
        }
        cv<T> y;
        y.x = yr;
        y.y = yi;
        psi[idx[r]] = y;
    }
}

// Generic 2^K x 2^K update (K = 2: K9 and K14; K = 3: K15) with one
// amplitude per lane. A block owns a tile of 256 amplitudes of the
// control-stripped index space: the K target bits plus the 8 - K lowest
// other bits (tile positions, ascending, so consecutive lanes hit
// consecutive addresses wherever those positions run from bit 0). Each
// lane puts its amplitude in LDS, gathers the 2^K amplitudes of its orbit
// and computes its own row of the matrix (staged in LDS as Re then Im).
struct NqGeom {
    uint8_t pos[8];    // tile positions in the stripped index, ascending
    uint8_t tslot[3];  // lane-index bit that holds target i
};

template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_nq_split(cv<T>* __restrict__ psi, Ins insc, uint64_t cmask, NqGeom gm,
                                                       const T* __restrict__ mat, uint64_t boff) {
    constexpr int D = 1 << K;
    __shared__ cv<T> s[kThreads];
    __shared__ T sm[2 * D * D];
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * D * D; i += kThreads) sm[i] = mat[i];
    uint64_t e = block_id<8>() + boff;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const uint64_t low = (1ull << gm.pos[p]) - 1;
        e = ((e & ~low) << 1) | (e & low);
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) e |= uint64_t((tid >> p) & 1) << gm.pos[p];
    const uint64_t idx = insert_zeros(e, insc) | cmask;
    s[tid] = ldamp<T, true>(psi + idx);
    int c = 0, tmask = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        c |= ((tid >> gm.tslot[i]) & 1) << i;
        tmask |= 1 << gm.tslot[i];
    }
    __syncthreads();
    T yr = 0, yi = 0;
#pragma unroll
    for (int cp = 0; cp < D; ++cp) {
        int lane = tid & ~tmask;
#pragma unroll
        for (int i = 0; i < K; ++i) lane |= ((cp >> i) & 1) << gm.tslot[i];
        const cv<T> x = s[lane];
        const T mr = sm[c * D + cp], mi = sm[D * D + c * D + cp];
        yr += mr * x.x - mi * x.y;
        yi += mr * x.y + mi * x.x;
    }
    cv<T> y;
    y.x = yr;
    y.y = yi;
    stamp<T, true>(psi + idx, y);
}

// Generic 8x8 (K15): local bits (b0, b1, b2) = (qubit, c1, c2).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_3q(cv<T>* __restrict__ psi, uint64_t norb, Ins ins,
                                                 uint64_t b0, uint64_t b1, uint64_t b2,
                                                 MatD<T, 8> m, uint64_t boff) {
    const uint64_t o = (uint64_t(blockIdx.x) + boff) * kThreads + threadIdx.x;
    if (o >= norb) return;
    const uint64_t i0 = insert_zeros(o, ins);
    uint64_t idx[8];
    cv<T> x[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        idx[c] = i0 | ((c & 1) ? b0 : 0) | ((c & 2) ? b1 : 0) | ((c & 4) ? b2 : 0);
        x[c] = psi[idx[c]];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            yr += m.re[r * 8 + c] * x[c].x - m.im[r * 8 + c] * x[c].y;
            yi += m.re[r * 8 + c] * x[c].y + m.im[r * 8 + c] * x[c].x;
        }
        cv<T> y;
        y.x = yr;
        y.y = yi;
        psi[idx[r]] = y;
    }
}

// Whole-state sweeps (fill, scale, full diagonal): one amplitude per lane
// in one pass, nontemporal, in launches of at most kMaxBlocks blocks that
// start at block `boff`.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill(cv<T>* __restrict__ psi, uint64_t A, T re, T im, uint64_t boff) {
    const uint64_t i = (uint64_t(blockIdx.x) + boff) * kThreads + threadIdx.x;
    if (i >= A) return;
    cv<T> v;
    v.x = re;
    v.y = im;
    stamp<T, true>(psi + i, v);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scale(cv<T>* __restrict__ psi, uint64_t A, T re, T im, uint64_t boff) {
    const uint64_t i = (uint64_t(blockIdx.x) + boff) * kThreads + threadIdx.x;
    if (i >= A) return;
    stamp<T, true>(psi + i, cmul<T>(re, im, ldamp<T, true>(psi + i)));
}

template <typename T>
__global__ void k_set_one(cv<T>* __restrict__ psi, uint64_t idx, T re, T im) {
    cv<T> v;
    v.x = re;
    v.y = im;
    psi[idx] = v;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_diag_full(cv<T>* __restrict__ psi, uint64_t A,
                                                        const double2* __restrict__ d, uint64_t boff) {
    const uint64_t i = (uint64_t(blockIdx.x) + boff) * kThreads + threadIdx.x;
    if (i >= A) return;
    const double2 f = ldamp<double, true>(d + i);
    stamp<T, true>(psi + i, cmul<T>(T(f.x), T(f.y), ldamp<T, true>(psi + i)));
}

// ---------------------------------------------------------------------------
// Host launchers.

// Gates that touch only the amplitudes with some bit set (controls, phase
// bits, the a != b half of a swap) use only part of each 128-byte line when
// that bit is below 3 (complex128) or 4 (complex64). Nontemporal loads then
// lose the untouched part, and the partial stores cost a read-modify-write:
// measured on MI355X at 30 qubits (S on qubit 0, CZ/SWAP on 0-1, CNOT with
// control 1) 8.1-14.7 ms with nontemporal loads against 6.05-6.3 ms with
// plain loads and nontemporal stores. Such launches take plain loads,
// nontemporal stores, one orbit per thread and XCD-ordered blocks (the
// one-amplitude-per-lane choice of k_1q is kept).
static Variant partial_line_variant(const qcs_register* r, Variant v, uint32_t lowest_condition_bit) {
    const uint32_t line_bits = r->prec == QCS_COMPLEX128 ? 3 : 4;
    if (lowest_condition_bit >= line_bits) return v;
    return Variant{1, 3, 1, 0, v.x};
}

// Variant dispatch: (P, NT, block regions) -> template instance. v.nt 0:
// plain access, 1: nontemporal loads and stores, 2: loads only, 3: stores
// only; v.xcd 0: launch order, 1: 8 regions (one per XCD), 2: 64 regions.
#define QCS_VARIANT_SWITCH(v, LAUNCH)                                         \
    do {                                                                      \
        if (v.p == 1 && v.nt == 0 && v.xcd == 0) LAUNCH(1, 0, 0);             \
        else if (v.p == 1 && v.nt == 1 && v.xcd == 0) LAUNCH(1, 1, 0);        \
        else if (v.p == 1 && v.nt == 1 && v.xcd == 1) LAUNCH(1, 1, 8);        \
        else if (v.p == 2 && v.nt == 0 && v.xcd == 1) LAUNCH(2, 0, 8);        \
        else if (v.p == 2 && v.nt == 1 && v.xcd == 0) LAUNCH(2, 1, 0);        \
        else if (v.p == 4 && v.nt == 0 && v.xcd == 1) LAUNCH(4, 0, 8);        \
        else if (v.p == 4 && v.nt == 1 && v.xcd == 1) LAUNCH(4, 1, 8);        \
        else if (v.p == 1 && v.nt == 1 && v.xcd == 2) LAUNCH(1, 1, 64);       \
        else if (v.p == 2 && v.nt == 1 && v.xcd == 2) LAUNCH(2, 1, 64);       \
        else if (v.p == 2 && v.nt == 0 && v.xcd == 2) LAUNCH(2, 0, 64);       \
        else if (v.p == 1 && v.nt == 2 && v.xcd == 1) LAUNCH(1, 2, 8);        \
        else if (v.p == 1 && v.nt == 3 && v.xcd == 1) LAUNCH(1, 3, 8);        \
        else if (v.p == 1 && v.nt == 3 && v.xcd == 0) LAUNCH(1, 3, 0);        \
        else if (v.p == 2 && v.nt == 3 && v.xcd == 1) LAUNCH(2, 3, 8);        \
        else LAUNCH(2, 1, 8);                                                 \
    } while (0)

template <typename T>
static int k1q_t(qcs_register* r, int mode, uint32_t target, const uint32_t* ctrls, int nctrl,
                 const hcplx* m, Variant v) {
    uint32_t pos[6];
    uint64_t cmask = 0;
    uint32_t low = target;
    pos[0] = target;
    for (int i = 0; i < nctrl; ++i) {
        pos[1 + i] = ctrls[i];
        cmask |= 1ull << ctrls[i];
        low = ctrls[i] < low ? ctrls[i] : low;
    }
    const Ins ins = make_ins(pos, 1 + nctrl);
    const uint64_t norb = r->A >> (1 + nctrl);
    MatD<T, 2> mm;
    for (int i = 0; i < 4; ++i) {
        mm.re[i] = T(m[i].re);
        mm.im[i] = T(m[i].im);
    }
    if (v.p == 0) {
        v = pick_variant(r, low, true);
        uint32_t lowc = 64;
        for (int i = 0; i < nctrl; ++i) lowc = ctrls[i] < lowc ? ctrls[i] : lowc;
        v = partial_line_variant(r, v, lowc);
    }
    auto* psi = static_cast<cv<T>*>(r->psi);
    const uint64_t bit = 1ull << target;
    // spread layout only when every block gets v.p full ranges of 2^sh orbits
    uint32_t sh = v.sp;
    if (sh && (v.p < 2 || sh < 8 || norb < (uint64_t(v.p) << sh))) sh = 0;
#define QCS_L1Q(PP, NTT, XX)                                                                                   \
    do {                                                                                                       \
        QCS_CHUNKED(grid_for(norb, kThreads * PP), {                                                           \
            if (mode == 0)                                                                                     \
                hipLaunchKernelGGL((k_1q<T, 0, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, norb, \
                                   ins, cmask, bit, mm, sh, boff);                                             \
            else if (mode == 1)                                                                                \
                hipLaunchKernelGGL((k_1q<T, 1, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, norb, \
                                   ins, cmask, bit, mm, sh, boff);                                             \
            else                                                                                               \
                hipLaunchKernelGGL((k_1q<T, 2, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, norb, \
                                   ins, cmask, bit, mm, sh, boff);                                             \
        })                                                                                                     \
    } while (0)
    if (v.x == 1 && norb >= kThreads / 2) {
        // one amplitude per lane (k_1q_split): needs whole blocks of 128 orbits
        uint32_t tl = target;
        for (int i = 0; i < nctrl; ++i) tl -= ctrls[i] < target ? 1 : 0;
        const Ins insc = make_ins(ctrls, nctrl);
        const uint64_t nblk = norb / (kThreads / 2);
#define QCS_L1S(NTT, XX)                                                                                        \
    do {                                                                                                        \
        QCS_CHUNKED(nblk, {                                                                                     \
            if (mode == 0)                                                                                      \
                hipLaunchKernelGGL((k_1q_split<T, 0, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                        \
            else if (mode == 1)                                                                                 \
                hipLaunchKernelGGL((k_1q_split<T, 1, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                        \
            else                                                                                                \
                hipLaunchKernelGGL((k_1q_split<T, 2, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                        \
        })                                                                                                      \
    } while (0)
        if (v.nt == 1 && v.xcd == 0) QCS_L1S(1, 0);
        else if (v.nt == 0 && v.xcd == 1) QCS_L1S(0, 8);
        else if (v.nt == 3 && v.xcd == 1) QCS_L1S(3, 8);
        else if (v.nt == 1 && v.xcd == 2) QCS_L1S(1, 64);
        else QCS_L1S(1, 8);
#undef QCS_L1S
        return check_launch(r, "k_1q_split");
    }
    if (v.x == 2 && norb >= kThreads / 2) {
        // one amplitude per lane, partner by lane shuffle (k_1q_shfl)
        uint32_t tl = target;
        for (int i = 0; i < nctrl; ++i) tl -= ctrls[i] < target ? 1 : 0;
        const Ins insc = make_ins(ctrls, nctrl);
        const uint64_t nblk = norb / (kThreads / 2);
#define QCS_L1W(NTT, XX)                                                                                       \
    do {                                                                                                       \
        QCS_CHUNKED(nblk, {                                                                                    \
            if (mode == 0)                                                                                     \
                hipLaunchKernelGGL((k_1q_shfl<T, 0, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                       \
            else if (mode == 1)                                                                                \
                hipLaunchKernelGGL((k_1q_shfl<T, 1, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                       \
            else                                                                                               \
                hipLaunchKernelGGL((k_1q_shfl<T, 2, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, insc, \
                                   cmask, tl, mm, boff);                                                       \
        })                                                                                                     \
    } while (0)
        if (v.nt == 1 && v.xcd == 0) QCS_L1W(1, 0);
        else if (v.nt == 3 && v.xcd == 1) QCS_L1W(3, 8);
        else QCS_L1W(1, 8);
#undef QCS_L1W
        return check_launch(r, "k_1q_shfl");
    }
    QCS_VARIANT_SWITCH(v, QCS_L1Q);
#undef QCS_L1Q
    return check_launch(r, "k_1q");
}

int k1q_variant(qcs_register* r, int mode, uint32_t target, const uint32_t* ctrls, int nctrl,
                const hcplx* m, Variant v) {
    return r->prec == QCS_COMPLEX128 ? k1q_t<double>(r, mode, target, ctrls, nctrl, m, v)
                                     : k1q_t<float>(r, mode, target, ctrls, nctrl, m, v);
}

int k1q(qcs_register* r, int mode, uint32_t target, const uint32_t* ctrls, int nctrl,
        const hcplx* m) {
    return k1q_variant(r, mode, target, ctrls, nctrl, m, Variant{0, 0, 0});
}

template <typename T>
static int kphase_t(qcs_register* r, const uint32_t* bits, int nbits, hcplx ph) {
    uint64_t cmask = 0;
    uint32_t low = 63;
    for (int i = 0; i < nbits; ++i) {
        cmask |= 1ull << bits[i];
        low = bits[i] < low ? bits[i] : low;
    }
    const Ins ins = make_ins(bits, nbits);
    const uint64_t norb = r->A >> nbits;
    const Variant v = partial_line_variant(r, pick_variant(r, low), low);
    auto* psi = static_cast<cv<T>*>(r->psi);
#define QCS_LPH(PP, NTT, XX)                                                                                   \
    do {                                                                                                       \
        QCS_CHUNKED(grid_for(norb, kThreads * PP),                                                             \
                    hipLaunchKernelGGL((k_phase<T, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi,   \
                                       norb, ins, cmask, T(ph.re), T(ph.im), boff))                             \
    } while (0)
    QCS_VARIANT_SWITCH(v, QCS_LPH);
#undef QCS_LPH
    return check_launch(r, "k_phase");
}

int kphase(qcs_register* r, const uint32_t* bits, int nbits, hcplx ph) {
    return r->prec == QCS_COMPLEX128 ? kphase_t<double>(r, bits, nbits, ph)
                                     : kphase_t<float>(r, bits, nbits, ph);
}

template <typename T>
static int kswap_t(qcs_register* r, uint32_t a, uint32_t b, const uint32_t* ctrls, int nctrl,
                   hcplx ph, bool phased) {
    uint32_t pos[6] = {a, b};
    uint64_t cmask = 0;
    uint32_t low = a < b ? a : b;
    for (int i = 0; i < nctrl; ++i) {
        pos[2 + i] = ctrls[i];
        cmask |= 1ull << ctrls[i];
        low = ctrls[i] < low ? ctrls[i] : low;
    }
    const Ins ins = make_ins(pos, 2 + nctrl);
    const uint64_t norb = r->A >> (2 + nctrl);
    const Variant v = partial_line_variant(r, pick_variant(r, low), low);
    auto* psi = static_cast<cv<T>*>(r->psi);
#define QCS_LSW(PP, NTT, XX)                                                                                   \
    do {                                                                                                       \
        QCS_CHUNKED(grid_for(norb, kThreads * PP), {                                                           \
            if (phased)                                                                                        \
                hipLaunchKernelGGL((k_swap<T, true, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi,  \
                                   norb, ins, cmask, 1ull << a, 1ull << b, T(ph.re), T(ph.im), boff);          \
            else                                                                                               \
                hipLaunchKernelGGL((k_swap<T, false, PP, NTT, XX>), dim3(g), dim3(kThreads), 0, r->stream, psi, \
                                   norb, ins, cmask, 1ull << a, 1ull << b, T(0), T(0), boff);                  \
        })                                                                                                     \
    } while (0)
    QCS_VARIANT_SWITCH(v, QCS_LSW);
#undef QCS_LSW
    return check_launch(r, "k_swap");
}

int kswap(qcs_register* r, uint32_t a, uint32_t b, const uint32_t* ctrls, int nctrl, hcplx ph,
          bool phased) {
    return r->prec == QCS_COMPLEX128 ? kswap_t<double>(r, a, b, ctrls, nctrl, ph, phased)
                                     : kswap_t<float>(r, a, b, ctrls, nctrl, ph, phased);
}

// k_nq_split launch for targets t[0..K) (qubit order = matrix bit order) and
// controls. Sets *launched = false when the state is too small for a tile
// (the caller then uses the orbit kernel).
template <typename T, int K>
static int knq_split(qcs_register* r, const uint32_t* t, const uint32_t* ctrls, int nctrl, const hcplx* m,
                     bool* launched) {
    *launched = false;
    const uint32_t nstr = r->n - uint32_t(nctrl);
    if (nstr < 8) return QCS_OK;
    constexpr int D = 1 << K;
    uint64_t cmask = 0;
    uint32_t ts[K];
    for (int i = 0; i < nctrl; ++i) cmask |= 1ull << ctrls[i];
    for (int i = 0; i < K; ++i) {
        ts[i] = t[i];
        for (int j = 0; j < nctrl; ++j) ts[i] -= ctrls[j] < t[i] ? 1 : 0;
    }
    uint32_t all[8];
    int na = 0;
    for (int i = 0; i < K; ++i) all[na++] = ts[i];
    for (uint32_t b = 0; na < 8; ++b) {
        bool tgt = false;
        for (int i = 0; i < K; ++i) tgt = tgt || ts[i] == b;
        if (!tgt) all[na++] = b;
    }
    std::sort(all, all + 8);
    NqGeom gm{};
    for (int p = 0; p < 8; ++p) gm.pos[p] = uint8_t(all[p]);
    for (int i = 0; i < K; ++i)
        for (int p = 0; p < 8; ++p)
            if (all[p] == ts[i]) gm.tslot[i] = uint8_t(p);
    std::vector<T> hm(2 * D * D);
    for (int i = 0; i < D * D; ++i) {
        hm[i] = T(m[i].re);
        hm[D * D + i] = T(m[i].im);
    }
    const void* dmat = nullptr;
    if (const int e = stage_upload(r, hm.data(), sizeof(T) * hm.size(), &dmat)) return e;
    const Ins insc = make_ins(ctrls, nctrl);
    const uint64_t nblk = (uint64_t(1) << nstr) / kThreads;
    QCS_CHUNKED(nblk, hipLaunchKernelGGL((k_nq_split<T, K>), dim3(g), dim3(kThreads), 0, r->stream,
                                         static_cast<cv<T>*>(r->psi), insc, cmask, gm, static_cast<const T*>(dmat),
                                         boff));
    *launched = true;
    return check_launch(r, "k_nq_split");
}

template <typename T>
static int k2q_t(qcs_register* r, uint32_t t0, uint32_t t1, const uint32_t* ctrls, int nctrl,
                 const hcplx* m) {
    {
        // measured on MI355X at 30 qubits (complex128): 6.1-6.3 ms against
        // 6.5-7.2 ms for the orbit kernel below
        const uint32_t t[2] = {t0, t1};
        bool launched = false;
        const int e = knq_split<T, 2>(r, t, ctrls, nctrl, m, &launched);
        if (e || launched) return e;
    }
    uint32_t pos[6] = {t0, t1};
    uint64_t cmask = 0;
    for (int i = 0; i < nctrl; ++i) {
        pos[2 + i] = ctrls[i];
        cmask |= 1ull << ctrls[i];
    }
    const Ins ins = make_ins(pos, 2 + nctrl);
    const uint64_t norb = r->A >> (2 + nctrl);
    MatD<T, 4> mm;
    for (int i = 0; i < 16; ++i) {
        mm.re[i] = T(m[i].re);
        mm.im[i] = T(m[i].im);
    }
    QCS_CHUNKED(grid_for(norb, kThreads),
                hipLaunchKernelGGL((k_2q<T>), dim3(g), dim3(kThreads), 0, r->stream, static_cast<cv<T>*>(r->psi),
                                   norb, ins, cmask, 1ull << t0, 1ull << t1, mm, boff));
    return check_launch(r, "k_2q");
}

int k2q(qcs_register* r, uint32_t t0, uint32_t t1, const uint32_t* ctrls, int nctrl,
        const hcplx* m) {
    return r->prec == QCS_COMPLEX128 ? k2q_t<double>(r, t0, t1, ctrls, nctrl, m)
                                     : k2q_t<float>(r, t0, t1, ctrls, nctrl, m);
}

template <typename T>
static int k3q_t(qcs_register* r, uint32_t t0, uint32_t t1, uint32_t t2, const hcplx* m) {
    {
        // measured on MI355X at 30 qubits (complex128): the lane kernel wins
        // when a target is below bit 3 (7.2 against 8.7 ms for qubits 0-2),
        // the orbit kernel otherwise (6.4-6.5 against 6.6-6.8 ms)
        const uint32_t t[3] = {t0, t1, t2};
        bool launched = false;
        if (std::min(t0, std::min(t1, t2)) < 3) {
            const int e = knq_split<T, 3>(r, t, nullptr, 0, m, &launched);
            if (e || launched) return e;
        }
    }
    uint32_t pos[3] = {t0, t1, t2};
    const Ins ins = make_ins(pos, 3);
    const uint64_t norb = r->A >> 3;
    MatD<T, 8> mm;
    for (int i = 0; i < 64; ++i) {
        mm.re[i] = T(m[i].re);
        mm.im[i] = T(m[i].im);
    }
    QCS_CHUNKED(grid_for(norb, kThreads),
                hipLaunchKernelGGL((k_3q<T>), dim3(g), dim3(kThreads), 0, r->stream, static_cast<cv<T>*>(r->psi),
                                   norb, ins, 1ull << t0, 1ull << t1, 1ull << t2, mm, boff));
    return check_launch(r, "k_3q");
}

int k3q(qcs_register* r, uint32_t t0, uint32_t t1, uint32_t t2, const hcplx* m) {
    return r->prec == QCS_COMPLEX128 ? k3q_t<double>(r, t0, t1, t2, m) : k3q_t<float>(r, t0, t1, t2, m);
}

int kfill(qcs_register* r, double re, double im) {
    QCS_CHUNKED(grid_for(r->A, kThreads), {
        if (r->prec == QCS_COMPLEX128)
            hipLaunchKernelGGL((k_fill<double>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<double2*>(r->psi), r->A, re, im, boff);
        else
            hipLaunchKernelGGL((k_fill<float>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<float2*>(r->psi), r->A, float(re), float(im), boff);
    })
    return check_launch(r, "k_fill");
}

int kscale(qcs_register* r, double re, double im) {
    QCS_CHUNKED(grid_for(r->A, kThreads), {
        if (r->prec == QCS_COMPLEX128)
            hipLaunchKernelGGL((k_scale<double>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<double2*>(r->psi), r->A, re, im, boff);
        else
            hipLaunchKernelGGL((k_scale<float>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<float2*>(r->psi), r->A, float(re), float(im), boff);
    })
    return check_launch(r, "k_scale");
}

int kset_one(qcs_register* r, uint64_t idx, double re, double im) {
    if (r->prec == QCS_COMPLEX128)
        hipLaunchKernelGGL((k_set_one<double>), dim3(1), dim3(1), 0, r->stream,
                           static_cast<double2*>(r->psi), idx, re, im);
    else
        hipLaunchKernelGGL((k_set_one<float>), dim3(1), dim3(1), 0, r->stream,
                           static_cast<float2*>(r->psi), idx, float(re), float(im));
    return check_launch(r, "k_set_one");
}

int kdiag_full(qcs_register* r, const double* diag_dev) {
    QCS_CHUNKED(grid_for(r->A, kThreads), {
        if (r->prec == QCS_COMPLEX128)
            hipLaunchKernelGGL((k_diag_full<double>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<double2*>(r->psi), r->A, reinterpret_cast<const double2*>(diag_dev), boff);
        else
            hipLaunchKernelGGL((k_diag_full<float>), dim3(g), dim3(kThreads), 0, r->stream,
                               static_cast<float2*>(r->psi), r->A, reinterpret_cast<const double2*>(diag_dev), boff);
    })
    return check_launch(r, "k_diag_full");
}

}  // namespace qcs
