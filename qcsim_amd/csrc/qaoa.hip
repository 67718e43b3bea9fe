// qaoa.hip - batched QAOA on weighted Ising models (complex64 states, FP64
// energies), one persistent launch per stage, the stage sequence of an
// evaluation captured in a hipGraph.
//
// Semantics follow QAOAIsingSubalgorithm::Exec / Energy (IsingModel.h:557-567,
// 106-119, 203-206): |0> -> H on all qubits; per layer l the cost operator
// (per edge CNOT, Rz(2 gamma_l J), CNOT; per site Rz(gamma_l h), :491-531)
// then Rx(2 beta_l) on every qubit (:533-554, without the optional extra
// mixing). The cost operator is diagonal: CNOT Rz(2 gJ) CNOT = exp(-i g J z z)
// and Rz(g h) = exp(-i g h z / 2), so amplitude s gains
// exp(-i g C(s)), C(s) = sum_edges J z_i z_j + sum_sites h_i z_i / 2, z = +1
// for bit 0. The energy is the exact FP64 sum_s |psi_s|^2 E(s),
// E(s) = -(sum h z + sum J z z), which the reference estimates with 100000
// shots (IsingModel.h:569-578).
//
// Work decomposition for an n-qubit instance, tile bits T and group width G:
//   n <= T: one SMALL unit runs every layer and the energy with the whole
//           state in LDS (k_qaoa_tile<true>);
//   else passes over the state of two kinds:
//     tile pass (k_qaoa_tile<false>): the 2^(n-T) tiles of 2^T amplitudes,
//           in units of consecutive tiles, each tile in LDS: the cost
//           diagonal (on |+> generated in place in layer 0) unless a mixer
//           pass ran it, then Rx on qubits 0..T-1;
//     mixer pass (k_qaoa_mix<g, MODE>) on group j of the ng = ceil((n-T)/G)
//           groups of qubits T+jG .. (the last may be narrower), each thread
//           one 2^g-amplitude orbit in registers: Rx; or Rx, the next
//           layer's cost diagonal and its Rx (merged); or Rx and the energy
//           (the last pass).
//   The order of the passes is in instance_passes (p ng + 1 passes for
//   ng >= 2, p (ng + 1) for ng = 1).
// A launch pulls units from its own atomic ticket until none are left; the
// launches of one evaluation are captured in a hipGraph. Instances of more
// than 18 qubits are stored in a permuted order (arena_xor below).
//
// C(s) and E(s) are walked in Gray-code order over the bits a thread owns:
// one full evaluation (all sites and edges) per thread, then per step the
// change of one flipped spin q, 2 z_q h_q and 2 z_q sum_j J_qj z_j over its
// neighbours (per-instance adjacency lists).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace qcs {

struct QInst {
    uint64_t off;        // first amplitude in the arena
    uint32_t n, ne;      // qubits, edges
    uint32_t eoff, hoff; // into the edge / site arrays
    uint32_t aoff, boff; // into the adjacency offsets (n + 1) / entries (2 ne)
    uint32_t tbits;      // tile bits T (== n for SMALL instances)
    uint32_t pslot0;     // first energy-partial slot
    uint32_t npslots;    // number of energy partials
    uint32_t pad;
};

struct QUnit {
    uint32_t inst;
    uint32_t idx;    // tile (tile launch) or first 256-orbit chunk (mixer)
    uint32_t layer;
    uint32_t qlo;    // first qubit of the mixer group
    uint32_t slot;   // energy partial slot (mixer with energy)
    uint32_t ntile;  // consecutive tiles from idx (tile launches)
    uint32_t flags;  // tile launches: kTileRxOnly
};

// Tile unit flag: Rx(2 beta) on the tile qubits without the cost diagonal
// (the layer's cost already ran in a merged mixer pass).
constexpr uint32_t kTileRxOnly = 1;

constexpr uint32_t kMaxSites = 64, kMaxEdges = 256;

// Amplitude i of an instance is stored at position i ^ arena_xor(i): bits
// 4..9 of i XOR a 6-bit fold of its bits 18.. (two fields for up to 30
// qubits). The 64 rows of a mixer orbit on qubits 18..23 (2 MiB apart) then
// sit at different offsets of their 8 KiB blocks instead of one; an orbit on
// qubits below 18 sees one constant XOR, a tile (bits < 12) is permuted
// within itself (aligned runs of 64 stay runs of 64), and an instance of <= 18
// qubits is stored in order. The map is linear over XOR and self-inverse (the
// mask reads only bits >= 18 and writes only bits 4..9).
__host__ __device__ __forceinline__ uint64_t arena_xor(uint64_t i) {
    return (((i >> 18) ^ (i >> 24)) & 63u) << 4;
}
constexpr uint32_t kMixChunks = 8;  // 256-orbit chunks per mixer unit
constexpr uint32_t kTileRun = 8;    // tiles per tile unit (one instance) with the cost diagonal
constexpr uint32_t kTileRunRx = 64; // tiles per Rx-only tile unit (no per-tile terms to stage)

// Instance parameters staged in LDS by every unit.
struct IsingLds {
    double h[kMaxSites];
    double J[kMaxEdges];
    uint16_t aoff[kMaxSites + 1];
    uint8_t ei[kMaxEdges], ej[kMaxEdges];
    uint8_t aj[2 * kMaxEdges];
    uint8_t ae[2 * kMaxEdges];  // the joining edge (index into J) of each neighbour
};

struct IsingArgsDev {
    const double* h;
    const double* J;
    const uint8_t* ei;
    const uint8_t* ej;
    const uint16_t* aoff;
    const uint8_t* aj;
    const uint8_t* ae;
};

__device__ void load_params(IsingLds& P, const QInst& I, const IsingArgsDev& g) {
    for (uint32_t i = threadIdx.x; i < I.n; i += 256) P.h[i] = g.h[I.hoff + i];
    for (uint32_t i = threadIdx.x; i <= I.n; i += 256) P.aoff[i] = g.aoff[I.aoff + i];
    for (uint32_t k = threadIdx.x; k < I.ne; k += 256) {
        P.J[k] = g.J[I.eoff + k];
        P.ei[k] = g.ei[I.eoff + k];
        P.ej[k] = g.ej[I.eoff + k];
    }
    for (uint32_t k = threadIdx.x; k < 2 * I.ne; k += 256) {
        P.aj[k] = g.aj[I.boff + k];
        P.ae[k] = g.ae[I.boff + k];
    }
    __syncthreads();
}

// coupling of adjacency entry a
__device__ __forceinline__ double adj_J(const IsingLds& P, uint32_t a) { return P.J[P.ae[a]]; }

// hz = sum_i h_i z_i and jzz = sum_edges J z_i z_j of basis state s.
__device__ __forceinline__ void ising_full(uint64_t s, uint32_t n, uint32_t ne, const IsingLds& P, double& hz,
                                           double& jzz) {
    hz = 0;
    jzz = 0;
    for (uint32_t i = 0; i < n; ++i) hz += ((s >> i) & 1) ? -P.h[i] : P.h[i];
    for (uint32_t k = 0; k < ne; ++k) jzz += (((s >> P.ei[k]) ^ (s >> P.ej[k])) & 1) ? -P.J[k] : P.J[k];
}

// Update hz, jzz for flipping spin q of s (s is the state before the flip).
__device__ __forceinline__ void ising_flip(uint64_t s, uint32_t q, const IsingLds& P, double& hz, double& jzz) {
    double f = 0;
    for (uint32_t a = P.aoff[q]; a < P.aoff[q + 1]; ++a) f += ((s >> P.aj[a]) & 1) ? -adj_J(P, a) : adj_J(P, a);
    if ((s >> q) & 1) {  // z_q = -1
        hz += 2 * P.h[q];
        jzz += 2 * f;
    } else {
        hz -= 2 * P.h[q];
        jzz -= 2 * f;
    }
}

// C(s) = sum J z z + sum h z / 2 (the cost diagonal's exponent), E(s) =
// -(sum h z + sum J z z).
__device__ __forceinline__ double cost_of(double hz, double jzz) { return jzz + 0.5 * hz; }
__device__ __forceinline__ double energy_of(double hz, double jzz) { return -(hz + jzz); }

__device__ __forceinline__ float2 cost_phase(float2 a, double gamma, double cost) {
    double th = gamma * cost;
    th -= 6.283185307179586476925 * rint(th * 0.15915494309189533576888);  // to [-pi, pi]
    float sn, cs;
    sincosf(float(th), &sn, &cs);  // exp(-i th) = cs - i sn
    return make_float2(cs * a.x + sn * a.y, cs * a.y - sn * a.x);
}

// Rx(2 beta) = [[c, -i s], [-i s, c]] on the pair (a, b).
__device__ __forceinline__ void rx_pair(float2& a, float2& b, float c, float s) {
    const float2 x = a, y = b;
    a = make_float2(c * x.x + s * y.y, c * x.y - s * y.x);
    b = make_float2(s * x.y + c * y.x, -s * x.x + c * y.y);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return t;
}

// Cost diagonal on a tile of 2^tb amplitudes (global index base + i). Thread
// t owns elements t + 256 m (the tile bits 8..tb-1 of m), visited in Gray
// order so that each step flips one spin.
__device__ void lds_cost(float2* t, uint32_t tb, uint64_t base, double gamma, const QInst& I, const IsingLds& P) {
    const uint32_t C = 1u << tb;
    if (threadIdx.x < C) {
        const uint32_t per = C > 256 ? C >> 8 : 1;
        uint64_t s = base + threadIdx.x;
        double hz, jzz;
        ising_full(s, I.n, I.ne, P, hz, jzz);
        t[threadIdx.x] = cost_phase(t[threadIdx.x], gamma, cost_of(hz, jzz));
        for (uint32_t m = 1; m < per; ++m) {
            const uint32_t q = 8 + __builtin_ctz(m);
            ising_flip(s, q, P, hz, jzz);
            s ^= 1ull << q;
            const uint32_t i = uint32_t(s - base);
            t[i] = cost_phase(t[i], gamma, cost_of(hz, jzz));
        }
    }
    __syncthreads();
}

__device__ void lds_rx(float2* t, uint32_t tb, float c, float s) {
    const uint32_t C = 1u << tb;
    for (uint32_t q = 0; q < tb; ++q) {
        const uint32_t qb = 1u << q;
        for (uint32_t p = threadIdx.x; p < C / 2; p += 256) {
            const uint32_t i0 = ((p >> q) << (q + 1)) | (p & (qb - 1));
            rx_pair(t[i0], t[i0 | qb], c, s);
        }
        __syncthreads();
    }
}

// LDS position of tile element i for tiles of >= 256 amplitudes: one pad
// element after every 16, so that 16 lanes at a stride of 16 elements (the
// round on qubits 0..3) fall in distinct banks. The map is additive over
// index parts with disjoint bits when the lower part stays below 16 or the
// upper part has bits 0..3 clear: an access base + (k << lo) for an orbit
// base with bits lo..lo+3 clear is one base address per lane plus the
// compile-time offset pad(k << lo), an immediate of the LDS instruction.
__device__ __forceinline__ uint32_t swz(uint32_t i) { return i + (i >> 4); }
__device__ __forceinline__ uint32_t tix(uint32_t i, uint32_t tb) { return tb >= 8 ? swz(i) : i; }
// LDS elements of a tile of 2^tb amplitudes
__host__ __device__ constexpr uint32_t tile_lds_elems(uint32_t tb) {
    return (1u << tb) + (tb >= 8 ? (1u << (tb - 4)) : 0u);
}

template <int G>
__device__ __forceinline__ void rx_reg(float2* x, float c, float s) {
    constexpr int D = 1 << G;
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int k = 0; k < D; ++k)
            if (!((k >> q) & 1)) rx_pair(x[k], x[k | (1 << q)], c, s);
}

// Rx(2 beta) up to a real factor, for passes that rotate every qubit of a
// tile: Rx = c (I - i t X) with t = tan beta when |c| >= |s| (form 1), else
// Rx = s (u I - i X) with u = cot beta (form 2); |t|, |u| <= 1. Each output
// component is then one FMA instead of a product and an FMA; the factor c
// (or s) per qubit is applied once, as `scale` = c^nq (s^nq), folded into
// the cost phase of the tile.
struct RxForm {
    float k;      // t (form 1) or u (form 2)
    float scale;  // c^nq or s^nq
    int f1;       // 1: form 1
};

__device__ __forceinline__ RxForm rx_form(double c, double s, uint32_t nq) {
    const bool f1 = fabs(c) >= fabs(s);
    const double f = f1 ? c : s;
    double sc = 1;
    for (uint32_t i = 0; i < nq; ++i) sc *= f;
    return RxForm{float(f1 ? s / c : c / s), float(sc), f1 ? 1 : 0};
}

template <bool F1>
__device__ __forceinline__ void rxs_pair(float2& a, float2& b, float k) {
    const float2 x = a, y = b;
    if (F1) {  // a - i t b, b - i t a
        a = make_float2(fmaf(k, y.y, x.x), fmaf(-k, y.x, x.y));
        b = make_float2(fmaf(k, x.y, y.x), fmaf(-k, x.x, y.y));
    } else {   // u a - i b, u b - i a
        a = make_float2(fmaf(k, x.x, y.y), fmaf(k, x.y, -y.x));
        b = make_float2(fmaf(k, y.x, x.y), fmaf(k, y.y, -x.x));
    }
}

template <int G, bool F1>
__device__ __forceinline__ void rxs_reg_f(float2* x, float k) {
    constexpr int D = 1 << G;
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int i = 0; i < D; ++i)
            if (!((i >> q) & 1)) rxs_pair<F1>(x[i], x[i | (1 << q)], k);
}

template <int G>
__device__ __forceinline__ void rxs_reg(float2* x, const RxForm& r) {
    if (r.f1) rxs_reg_f<G, true>(x, r.k);
    else rxs_reg_f<G, false>(x, r.k);
}

enum : int { IN_PLUS = 0, IN_REG = 1, IN_LDS = 2 };

// Rx(2 beta) on the G tile qubits lo..lo+G-1, each thread (index tid) taking
// orbits of 2^G elements from the LDS tile into registers.
template <int G>
__device__ void rx_round(float2* t, uint32_t tid, uint32_t tb, uint32_t lo, const RxForm& rf) {
    constexpr int D = 1 << G;
    for (uint32_t o = tid; o < (1u << (tb - G)); o += 256) {
        const uint32_t lb = uint32_t(insert_run(o, lo, G));
        float2* const tp = t + swz(lb);  // lb has bits lo..lo+G-1 clear: swz is additive
        float2 y[D];
#pragma unroll
        for (int k = 0; k < D; ++k) y[k] = tp[swz(uint32_t(k) << lo)];
        rxs_reg<G>(y, rf);
#pragma unroll
        for (int k = 0; k < D; ++k) tp[swz(uint32_t(k) << lo)] = y[k];
    }
    __syncthreads();
}

// Tiles of 4096 amplitudes: wave w owns the 1024 elements with bits 10..11
// = w for the Rx rounds on qubits 4..7 and 0..3 and for the store, so those
// steps exchange data only within the wave and need no block barrier. In a
// round on qubits lo..lo+3 (lo = 0 or 4) lane bits 0..3 give the other four
// of bits 0..7 and lane bits 4..5 give bits 8..9.
__device__ __forceinline__ uint32_t wave_orbit_base(uint32_t tid, uint32_t lo) {
    const uint32_t lane = tid & 63, w = tid >> 6;
    return ((lane & 15) << (lo == 0 ? 4 : 0)) | ((lane >> 4) << 8) | (w << 10);
}

__device__ __forceinline__ void rx_wave(float2* t, uint32_t tid, uint32_t lo, const RxForm& rf) {
    float2* const tp = t + swz(wave_orbit_base(tid, lo));  // base bits lo..lo+3 clear
    float2 y[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) y[k] = tp[swz(uint32_t(k) << lo)];
    rxs_reg<4>(y, rf);
#pragma unroll
    for (int k = 0; k < 16; ++k) tp[swz(uint32_t(k) << lo)] = y[k];
}

// Orders this wave's LDS writes before its later LDS reads (the data a lane
// reads was written by another lane of the same wave).
__device__ __forceinline__ void lds_wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Terms of the cost C(s) = sum J z z + sum h z / 2 for a tile of 2^tb
// amplitudes whose bits tb.. are fixed (by `base`): K = the part of C on the
// fixed spins alone; F[i] (i < tb) = h_i / 2 + sum of J_ij z_j over the fixed
// neighbours j of i; ef[k] = exp(2 i gamma F[8 + k]); and, per instance and
// gamma, q[m] = exp(-i gamma Q(m)) with Q(m) = sum of J z z over the edges
// inside bits 8..tb-1 when those bits hold m.
struct TileCostLds {
    double F[kTileRun][12];
    double K[kTileRun];
    float2 q[16];
    float2 ef[kTileRun][4];
};

// Thread t's share of the cost phase that depends on the instance and gamma
// but not on the tile: c0 = sum of J z z over the edges inside bits 0..7 of
// t, plus sum_k G_k; eg[k] = exp(2 i gamma G_k), G_k = sum of J_ij z_j over
// the neighbours j < 8 of site i = 8 + k.
struct ThreadTerms {
    double c0;
    float2 eg[4];
};

// exp(-i th) for th in double, reduced to [-pi, pi] in double, then FP32.
__device__ __forceinline__ float2 expmi(double th) {
    th -= 6.283185307179586476925 * rint(th * 0.15915494309189533576888);
    float sn, cs;
    sincosf(float(th), &sn, &cs);
    return make_float2(cs, -sn);
}

__device__ __forceinline__ double zf(uint64_t v, uint32_t b) { return ((v >> b) & 1) ? -1.0 : 1.0; }

// The per-tile terms of the cost phase for the tiles tile0 .. tile0+ntile-1
// (ntile <= kTileRun) of 2^(8+M) amplitudes, all at once: F[j][i] and
// ef[j][i-8] by thread 32 j + i (i < tb), K[j] by thread 32 j + 31. The
// barrier that ends thread_terms orders them before their use.
template <int M>
__device__ void tile_terms(TileCostLds& TC, double gamma, const QInst& I, const IsingLds& P, uint64_t tile0,
                           uint32_t ntile) {
    constexpr uint32_t tb = 8 + M;
    const uint32_t j = threadIdx.x >> 5, i = threadIdx.x & 31;
    if (j >= ntile) return;
    const uint64_t base = (tile0 + j) << tb;
    if (i < tb) {
        double f = 0.5 * P.h[i];
        for (uint32_t a = P.aoff[i]; a < P.aoff[i + 1]; ++a)
            if (P.aj[a] >= tb) f += adj_J(P, a) * zf(base, P.aj[a]);
        TC.F[j][i] = f;
        if (i >= 8) TC.ef[j][i - 8] = expmi(-2.0 * f * gamma);
    } else if (i == 31) {
        double k = 0;
        for (uint32_t s = tb; s < I.n; ++s) k += 0.5 * P.h[s] * zf(base, s);
        for (uint32_t e = 0; e < I.ne; ++e)
            if (P.ei[e] >= tb && P.ej[e] >= tb) k += P.J[e] * zf(base, P.ei[e]) * zf(base, P.ej[e]);
        TC.K[j] = k;
    }
}

// Fills W for this thread and TC.q, for tiles of 2^(8+M) amplitudes; ends
// with a barrier.
template <int M>
__device__ void thread_terms(ThreadTerms& W, double gamma, const QInst& I, const IsingLds& P, TileCostLds& TC) {
    constexpr uint32_t tb = 8 + M;
    const uint32_t tid = threadIdx.x;
    if (tid >= 64 && tid < 64 + (1u << M)) {
        const uint64_t mv = uint64_t(tid - 64) << 8;
        double qv = 0;
        for (uint32_t e = 0; e < I.ne; ++e)
            if (P.ei[e] >= 8 && P.ei[e] < tb && P.ej[e] >= 8 && P.ej[e] < tb)
                qv += P.J[e] * zf(mv, P.ei[e]) * zf(mv, P.ej[e]);
        TC.q[tid - 64] = expmi(gamma * qv);
    }
    double c = 0;
    for (uint32_t i = 0; i < 8; ++i) {
        const double zi = zf(tid, i);
        for (uint32_t a = P.aoff[i]; a < P.aoff[i + 1]; ++a)
            if (P.aj[a] < i) c += adj_J(P, a) * zi * zf(tid, P.aj[a]);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const uint32_t i = 8 + k;
        double g = 0;
        for (uint32_t a = P.aoff[i]; a < P.aoff[i + 1]; ++a)
            if (P.aj[a] < 8) g += adj_J(P, a) * zf(tid, P.aj[a]);
        c += g;
        W.eg[k] = expmi(-2.0 * gamma * g);
    }
    W.c0 = c;
    __syncthreads();
}

// One layer (cost diagonal, then Rx(2 beta) on qubits 0..tb-1) on a tile of
// 2^tb amplitudes, 8 <= tb = 8 + M <= 12, whose element 0 has global index
// `base`; W and TC.q come from thread_terms<M> for this instance and gamma.
// Thread t takes elements t + 256 m (m < 2^M: the tile bits 8..tb-1) in Gray
// order over m and applies the cost phase; these 2^M elements are the orbit
// of qubits 8..tb-1, so their Rx follows in registers before the elements go
// to the (padded) LDS tile; two rounds then apply Rx on qubits 4..7 and
// 0..3 with orbits held in registers. The input is |+> (amplitude amp), the
// stored tile in global memory (element i at in[i ^ ci]) or the LDS tile;
// the result is left in the LDS tile.
//
// The cost phase splits along the tile: with z_i = 1 - 2 b_i,
//   C = K + A(t) + sum_{i in 8..tb-1} z_i F'_i(t) + Q(m),
// A(t) = sum_{i<8} z_i F[i] + sum of J z z over edges inside bits 0..7, and
// F'_i(t) = F[i] + G_{i-8}(t). So
//   exp(-i gamma C) = exp(-i gamma (K + A + sum F')) prod_{b_i = 1}
//                     exp(2 i gamma F'_i) q[m],
// with exp(2 i gamma F'_i) = ef[i-8] eg[i-8]: per tile one sincos per thread,
// and per amplitude three FP32 complex products (one along the Gray walk),
// instead of a sincos and an energy update per amplitude. The walk runs in
// FP32 so that the orbit, the walk and the Rx fit the 128 VGPRs that four
// waves per SIMD allow; its 2^M - 1 steps add at most about 2^M FP32
// roundings to a phase that is itself an FP32 sincos.
//
// `tid` is threadIdx.x passed through an empty asm statement by the caller,
// so that the compiler recomputes the LDS addresses per tile instead of
// holding them all in registers across a run of tiles.
template <int M, bool COST>
__device__ void tile_layer(float2* t, uint32_t tid, uint32_t jt, double gamma, const RxForm& rf, const QInst& I,
                           const TileCostLds* TC, const ThreadTerms& W, int in_mode,
                           const float2 (&xin)[1 << M], float amp) {
    float2 y[1 << M];
    if constexpr (!COST) {
        // Rx alone: the orbit scaled by the rotations' real factor
#pragma unroll
        for (uint32_t m = 0; m < (1u << M); ++m) {
            const float2 v = in_mode == IN_REG ? xin[m]
                                               : (in_mode == IN_LDS ? t[swz(tid) + swz(m << 8)] : make_float2(amp, 0.f));
            y[m] = make_float2(rf.scale * v.x, rf.scale * v.y);
        }
    } else {
    // per-thread terms, from the tile's terms (tile_terms, slot jt)
    double A = TC->K[jt] + W.c0;
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) A += zf(tid, i) * TC->F[jt][i];
    float2 e[M > 0 ? M : 1];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        A += TC->F[jt][8 + k];
        // exp(2 i gamma F'_k) = ef eg, multiplied and rescaled to modulus 1
        // (to O(delta^2) for |.|^2 = 1 + delta) in FP64, then rounded, so
        // that the walk's alternating products by it and its conjugate cancel
        // to FP32 rounding
        const float2 a = TC->ef[jt][k], b = W.eg[k];
        const double x = double(a.x) * b.x - double(a.y) * b.y, y = double(a.x) * b.y + double(a.y) * b.x;
        const double r = 1.5 - 0.5 * (x * x + y * y);
        e[k] = make_float2(float(x * r), float(y * r));
    }
    float2 ph = expmi(gamma * A);
    ph.x *= rf.scale;  // the real factor of the tb scaled rotations below
    ph.y *= rf.scale;
#pragma unroll
    for (uint32_t m = 0; m < (1u << M); ++m) {
        const uint32_t gm = m ^ (m >> 1);
        if (m) {
            const int k = __builtin_ctz(m);
            const float2 f = e[k];
            const float fy = ((gm >> k) & 1) ? f.y : -f.y;  // set: times e_k; cleared: divided by it
            ph = make_float2(ph.x * f.x - ph.y * fy, ph.x * fy + ph.y * f.x);
        }
        const float2 qm = TC->q[gm];
        const float pr = ph.x, pi = ph.y;
        const float wr = pr * qm.x - pi * qm.y, wi = pr * qm.y + pi * qm.x;
        const float2 v = in_mode == IN_REG ? xin[gm]
                                           : (in_mode == IN_LDS ? t[swz(tid) + swz(gm << 8)] : make_float2(amp, 0.f));
        y[gm] = make_float2(wr * v.x - wi * v.y, wr * v.y + wi * v.x);
    }
    }
    // y is the orbit of qubits 8..tb-1: their Rx before the LDS store
    if constexpr (M > 0) rxs_reg<M>(y, rf);
    // every wave has read what it needs of the tile in LDS (the previous
    // tile's store, or this layer's input) before it is overwritten
    __syncthreads();
#pragma unroll
    for (uint32_t m = 0; m < (1u << M); ++m) t[swz(tid) + swz(m << 8)] = y[m];
    __syncthreads();
}

// The rest of a layer on the tile in LDS: Rx on qubits 4..7 and 0..3.
template <int M>
__device__ void tile_rounds(float2* t, uint32_t tid, const RxForm& rf) {
    constexpr uint32_t tb = 8 + M;
    if constexpr (M == 4) {
        // wave-local rounds; the caller orders the result before other waves
        // read it
        rx_wave(t, tid, 4, rf);
        lds_wave_sync();
        rx_wave(t, tid, 0, rf);
        lds_wave_sync();
    } else {
        rx_round<4>(t, tid, tb, 4, rf);
        rx_round<4>(t, tid, tb, 0, rf);
    }
}

// A tile unit: one layer on tiles tile0 .. tile0+ntile-1 of an instance,
// 2^(8+M) amplitudes each, written back to psi. Software-pipelined: the next
// tile's global loads are issued as soon as this tile is in LDS (its input
// registers are then free), so that they are in flight during this tile's
// LDS rounds and stores and the next tile's per-thread terms. The first tile
// is peeled off the loop, so that the loop head is reached with the same
// memory operations outstanding on every path (the next tile's loads, then
// this tile's stores) and waits only for the loads.
template <int M, bool COST>
__device__ void tile_run(float2* t, float2* psi, uint64_t tile0, uint32_t ntile, double gamma, double c, double s,
                         const QInst& I, const IsingLds* P, TileCostLds* TC, bool first, float amp) {
    constexpr uint32_t tb = 8 + M, C = 1u << tb;
    const RxForm rf = rx_form(c, s, tb);
    ThreadTerms W{};
    if constexpr (COST) {
        tile_terms<M>(*TC, gamma, I, *P, tile0, ntile);
        thread_terms<M>(W, gamma, I, *P, *TC);
    }
    // stored positions: element i of the tile at base is io[i ^ ci] (io =
    // psi + base for tiles of >= 1024 amplitudes; a smaller tile trades
    // elements with the tile at base ^ (the mask's bits >= tb), which has
    // the same mask)
    auto io_of = [&](uint64_t base) {
        const uint64_t cx = arena_xor(base);
        return psi + (base ^ (cx & ~uint64_t(C - 1)));
    };
    auto ci_of = [&](uint64_t base) { return uint32_t(arena_xor(base)) & (C - 1); };
    float2 xin[1 << M];
    auto fetch = [&](uint64_t base) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid < 256u);
        const float2* io = io_of(base);
        const uint32_t ci = ci_of(base);
#pragma unroll
        for (uint32_t m = 0; m < (1u << M); ++m) xin[m] = ldamp<float, true>(io + ((tid + (m << 8)) ^ ci));
    };
    auto one_tile = [&](uint32_t j) {
        const uint64_t base = (tile0 + j) << tb;
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid < 256u);
        tile_layer<M, COST>(t, tid, j, gamma, rf, I, TC, W, first ? IN_PLUS : IN_REG, xin, amp);
        if (!first && j + 1 < ntile) fetch(base + C);
        tile_rounds<M>(t, tid, rf);
        float2* const io = io_of(base);
        const uint32_t ci = ci_of(base);
        if constexpr (M == 4) {
            // each wave stores the 1024 elements it rotated last
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r) {
                const uint32_t a = ((tid >> 6) << 10) | (tid & 63);
                stamp<float, true>(io + ((a | (r << 6)) ^ ci), t[swz(a) + swz(r << 6)]);
            }
        } else {
#pragma unroll
            for (uint32_t r = 0; r < C / 256; ++r) {
                stamp<float, true>(io + ((tid + 256 * r) ^ ci), t[swz(tid) + swz(256 * r)]);
            }
        }
    };
    if (!first) fetch(tile0 << tb);
    one_tile(0);
    for (uint32_t j = 1; j < ntile; ++j) one_tile(j);
}

// Every layer of a SMALL instance of 8 + M qubits, the state kept in the LDS
// tile.
template <int M>
__device__ void small_run(float2* t, const double* an, uint32_t p, const QInst& I, const IsingLds& P,
                          TileCostLds& TC, float amp) {
    for (uint32_t l = 0; l < p; ++l) {
        double sb, cb;
        sincos(an[2 * l + 1], &sb, &cb);
        ThreadTerms W;
        tile_terms<M>(TC, an[2 * l], I, P, 0, 1);
        thread_terms<M>(W, an[2 * l], I, P, TC);
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        __builtin_assume(tid < 256u);
        const RxForm rf = rx_form(cb, sb, 8 + M);
        float2 none[1 << M] = {};  // not read: the input is |+> or the LDS tile
        tile_layer<M, true>(t, tid, 0, an[2 * l], rf, I, &TC, W, l ? IN_LDS : IN_PLUS, none, amp);
        tile_rounds<M>(t, tid, rf);
    }
    __syncthreads();  // the energy pass reads every wave's elements
}

template <bool COST>
__device__ __forceinline__ void tile_run_tb(uint32_t tb, float2* t, float2* psi, uint64_t tile0, uint32_t ntile,
                                            double gamma, double c, double s, const QInst& I, const IsingLds* P,
                                            TileCostLds* TC, bool first, float amp) {
    switch (tb) {
        case 8: tile_run<0, COST>(t, psi, tile0, ntile, gamma, c, s, I, P, TC, first, amp); break;
        case 9: tile_run<1, COST>(t, psi, tile0, ntile, gamma, c, s, I, P, TC, first, amp); break;
        case 10: tile_run<2, COST>(t, psi, tile0, ntile, gamma, c, s, I, P, TC, first, amp); break;
        case 11: tile_run<3, COST>(t, psi, tile0, ntile, gamma, c, s, I, P, TC, first, amp); break;
        default: tile_run<4, COST>(t, psi, tile0, ntile, gamma, c, s, I, P, TC, first, amp); break;
    }
}

__device__ __forceinline__ void small_run_tb(uint32_t tb, float2* t, const double* an, uint32_t p, const QInst& I,
                             const IsingLds& P, TileCostLds& TC, float amp) {
    switch (tb) {
        case 8: small_run<0>(t, an, p, I, P, TC, amp); break;
        case 9: small_run<1>(t, an, p, I, P, TC, amp); break;
        case 10: small_run<2>(t, an, p, I, P, TC, amp); break;
        case 11: small_run<3>(t, an, p, I, P, TC, amp); break;
        default: small_run<4>(t, an, p, I, P, TC, amp); break;
    }
}

// Tile units (SMALL = false): one layer on a run of 2^T tiles. SMALL units:
// every layer and the energy of an instance that fits one tile. Three waves
// per SIMD (at most 168 VGPRs): the software-pipelined tile loop keeps the
// next tile's 2^M loads in registers across the LDS rounds, and at four
// waves (128 VGPRs) it spills inside the loop, where a scratch reload waits
// for every outstanding load and store (measured 1 % slower per batch than
// three waves; unpipelined, four and three waves measured alike).
template <bool SMALL>
__global__ __launch_bounds__(256, 3) void k_qaoa_tile(float2* __restrict__ arena, const QInst* __restrict__ insts,
                                                   const QUnit* __restrict__ units, uint32_t nunits,
                                                   uint32_t* __restrict__ ticket, const double* __restrict__ ang,
                                                   uint32_t p, IsingArgsDev ga, double* __restrict__ energy) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2* tile = reinterpret_cast<float2*>(smem_raw);
    __shared__ IsingLds P;
    __shared__ TileCostLds TC;
    __shared__ double red[4];
    __shared__ uint32_t su;
    for (;;) {
        if (threadIdx.x == 0) su = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t u = su;
        __syncthreads();
        if (u >= nunits) break;  // every wave of the block leaves together
        const QUnit U = units[u];
        const QInst I = insts[U.inst];
        load_params(P, I, ga);
        float2* psi = arena + I.off;
        const double* an = ang + size_t(U.inst) * p * 2;
        if (SMALL) {
            const uint32_t C = 1u << I.n;
            const float amp = float(1.0 / sqrt(double(C)));
            if (I.n >= 8) {
                small_run_tb(I.n, tile, an, p, I, P, TC, amp);
            } else {
                for (uint32_t i = threadIdx.x; i < C; i += 256) tile[i] = make_float2(amp, 0.f);
                __syncthreads();
                for (uint32_t l = 0; l < p; ++l) {
                    lds_cost(tile, I.n, 0, an[2 * l], I, P);
                    double sb, cb;
                    sincos(an[2 * l + 1], &sb, &cb);
                    lds_rx(tile, I.n, float(cb), float(sb));
                }
            }
            double acc = 0;
            if (threadIdx.x < C) {
                const uint32_t per = C > 256 ? C >> 8 : 1;
                uint64_t s = threadIdx.x;
                double hz, jzz;
                ising_full(s, I.n, I.ne, P, hz, jzz);
                for (uint32_t m = 0;; ++m) {
                    const float2 x = tile[tix(uint32_t(s), I.n)];
                    acc += (double(x.x) * x.x + double(x.y) * x.y) * energy_of(hz, jzz);
                    psi[s] = x;  // n <= 12: stored in order
                    if (m + 1 == per) break;
                    const uint32_t q = 8 + __builtin_ctz(m + 1);
                    ising_flip(s, q, P, hz, jzz);
                    s ^= 1ull << q;
                }
            }
            const double tot = block_sum(acc, red);
            if (threadIdx.x == 0) energy[U.inst] = tot;
        } else {
            const uint32_t tb = I.tbits, C = 1u << tb;
            const float amp = float(1.0 / sqrt(double(1ull << I.n)));
            double sb, cb;
            sincos(an[2 * U.layer + 1], &sb, &cb);
            // the first pass of an evaluation generates |+...+> in place
            const bool gen = U.layer == 0;
            if (tb >= 8) {
                tile_run_tb<true>(tb, tile, psi, U.idx, U.ntile, an[2 * U.layer], cb, sb, I, &P, &TC, gen, amp);
            } else {
                for (uint32_t j = 0; j < U.ntile; ++j) {
                    const uint64_t base = (uint64_t(U.idx) + j) << tb;
                    for (uint32_t i = threadIdx.x; i < C; i += 256)
                        tile[i] = gen ? make_float2(amp, 0.f)
                                      : ldamp<float, true>(psi + ((base + i) ^ arena_xor(base)));
                    __syncthreads();
                    lds_cost(tile, tb, base, an[2 * U.layer], I, P);
                    lds_rx(tile, tb, float(cb), float(sb));
                    for (uint32_t i = threadIdx.x; i < C; i += 256)
                        stamp<float, true>(psi + ((base + i) ^ arena_xor(base)), tile[i]);
                    __syncthreads();
                }
            }
        }
        __syncthreads();
    }
}

// Rx-only tile units (kTileRxOnly): Rx(2 beta) on the tile qubits of a run
// of tiles, the layer's cost diagonal having run in a mixer pass. Only the
// LDS tile is staged (no instance terms), and at most 128 VGPRs: four
// blocks per CU instead of the three that the cost kernel's 41 KiB of LDS
// and 168 registers allow.
__global__ __launch_bounds__(256, 4) void k_qaoa_tile_rx(float2* __restrict__ arena, const QInst* __restrict__ insts,
                                                      const QUnit* __restrict__ units, uint32_t nunits,
                                                      uint32_t* __restrict__ ticket, const double* __restrict__ ang,
                                                      uint32_t p) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2* tile = reinterpret_cast<float2*>(smem_raw);
    __shared__ uint32_t su;
    for (;;) {
        if (threadIdx.x == 0) su = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t u = su;
        __syncthreads();
        if (u >= nunits) break;  // every wave of the block leaves together
        const QUnit U = units[u];
        const QInst I = insts[U.inst];
        float2* psi = arena + I.off;
        const uint32_t tb = I.tbits, C = 1u << tb;
        double sb, cb;
        sincos(ang[(size_t(U.inst) * p + U.layer) * 2 + 1], &sb, &cb);
        if (tb >= 8) {
            tile_run_tb<false>(tb, tile, psi, U.idx, U.ntile, 0.0, cb, sb, I, nullptr, nullptr, false, 0.f);
        } else {
            for (uint32_t j = 0; j < U.ntile; ++j) {
                const uint64_t base = (uint64_t(U.idx) + j) << tb;
                for (uint32_t i = threadIdx.x; i < C; i += 256)
                    tile[i] = ldamp<float, true>(psi + ((base + i) ^ arena_xor(base)));
                __syncthreads();
                lds_rx(tile, tb, float(cb), float(sb));
                for (uint32_t i = threadIdx.x; i < C; i += 256)
                    stamp<float, true>(psi + ((base + i) ^ arena_xor(base)), tile[i]);
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// Mixer modes: Rx(2 beta_l) alone; the same and this unit's share of the
// energy (last pass of an evaluation); or a merged pass across the boundary
// of layers l and l + 1: Rx(2 beta_l), the cost diagonal of layer l + 1,
// Rx(2 beta_{l+1}), all on the group's orbit in registers. A merged pass is
// valid when the group's Rx is the last of layer l's rotations to run (the
// rotations of one layer act on different qubits and commute) and the first
// of layer l + 1's; the cost diagonal, applied with every amplitude's full
// index at hand, can run in any pass that sees each amplitude once.
enum : int { MIX_PLAIN = 0, MIX_ENERGY = 1, MIX_MERGE = 2 };

// Mixer units: kMixChunks chunks of 256 orbits of the qubits qlo ..
// qlo+g-1, in the unit's mode.
template <int g, int MODE>
__global__ __launch_bounds__(256) void k_qaoa_mix(float2* __restrict__ arena, const QInst* __restrict__ insts,
                                                  const QUnit* __restrict__ units, uint32_t nunits,
                                                  uint32_t* __restrict__ ticket, const double* __restrict__ ang,
                                                  uint32_t p, IsingArgsDev ga, double* __restrict__ partial) {
    constexpr int D = 1 << g;
    constexpr bool E = MODE == MIX_ENERGY, MG = MODE == MIX_MERGE;
    // low part of the merged pass's two-level walk over the group bits
    constexpr int GL = g < 3 ? g : 3, DL = 1 << GL, DH = D >> GL;
    __shared__ IsingLds P;
    __shared__ double qe[E ? D : 1];   // E: -sum J z z over the edges inside the group, per group value
    __shared__ float2 qph[MG ? D : 1];  // MG: exp(-i gamma' (Q(c) - Q(0))), Q = sum of J z z inside the group
    __shared__ double red[4];
    __shared__ uint32_t su;
    for (;;) {
        if (threadIdx.x == 0) su = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t u = su;
        __syncthreads();
        if (u >= nunits) break;
        const QUnit U = units[u];
        const QInst I = insts[U.inst];
        if (E || MG) load_params(P, I, ga);
        if (E) {
            for (uint32_t m = threadIdx.x; m < uint32_t(D); m += 256) {
                double v = 0;
                for (uint32_t e = 0; e < I.ne; ++e) {
                    const uint32_t a = P.ei[e], b = P.ej[e];
                    if (a >= U.qlo && a < U.qlo + g && b >= U.qlo && b < U.qlo + g)
                        v -= (((m >> (a - U.qlo)) ^ (m >> (b - U.qlo))) & 1) ? -P.J[e] : P.J[e];
                }
                qe[m] = v;
            }
            __syncthreads();
        }
        const double* an = ang + size_t(U.inst) * p * 2;
        // MG: gamma and beta of layer l + 1
        const uint32_t ln = U.layer + 1;
        const double gn = MG ? an[2 * ln] : 0.0;
        if (MG) {
            for (uint32_t m = threadIdx.x; m < uint32_t(D); m += 256) {
                double v = 0;  // Q(m) - Q(0): -2 J for each inside edge whose spins differ in m
                for (uint32_t e = 0; e < I.ne; ++e) {
                    const uint32_t a = P.ei[e], b = P.ej[e];
                    if (a >= U.qlo && a < U.qlo + g && b >= U.qlo && b < U.qlo + g &&
                        (((m >> (a - U.qlo)) ^ (m >> (b - U.qlo))) & 1))
                        v -= 2.0 * P.J[e];
                }
                qph[m] = expmi(gn * v);
            }
            __syncthreads();
        }
        float2* psi = arena + I.off;
        double sb, cb, sb2 = 0, cb2 = 1;
        sincos(an[2 * U.layer + 1], &sb, &cb);
        if (MG) sincos(an[2 * ln + 1], &sb2, &cb2);
        const float fc = float(cb), fs = float(sb), fc2 = float(cb2), fs2 = float(sb2);
        const uint64_t norb = (1ull << I.n) >> g;
        double acc = 0;
        // MG, E: hz and jzz of the previous orbit's base, updated spin by
        // spin to the next base (the chunks of a unit differ in one or two
        // spins) instead of a full evaluation per orbit
        double hz = 0, jzz = 0;
        uint64_t prev = 0;
#pragma unroll 1
        for (uint32_t r = 0; r < kMixChunks; ++r) {
            const uint64_t o = (uint64_t(U.idx) + r) * 256 + threadIdx.x;
            if (o >= norb) break;
            const uint64_t base = insert_run(o, U.qlo, g);
            // stored position of base | c << qlo: pb ^ row(c) (the two parts
            // are disjoint and the mask is linear)
            const uint64_t pb = base ^ arena_xor(base);
            auto row = [&](int c) {
                const uint64_t v = uint64_t(c) << U.qlo;
                return v ^ arena_xor(v);
            };
            float2 x[D];
#pragma unroll
            for (int c = 0; c < D; ++c) x[c] = ldamp<float, true>(psi + (pb ^ row(c)));
            // Terms that depend on the orbit's fixed spins alone, computed
            // while the loads are in flight. With z = 1 - 2 b, site
            // i = qlo + k and n_k = sum of J_ij z_j over the neighbours j of
            // i outside the group, setting bit k of c (z_i: +1 -> -1)
            //   MG: changes C = sum J z z + sum h z / 2 by -2 f_k,
            //       fk[k] = f_k = h_i / 2 + n_k, and the inside edges by
            //       Q(c) - Q(0), so C(base | c) = C(base) - 2 sum_{k: c_k = 1}
            //       f_k + Q(c) - Q(0);
            //   E:  changes E = -(sum h z + sum J z z) by fk[k] = 2 (h_i + n_k)
            //       outside the group; E(base | c) = E(base) - qe[0] +
            //       sum_{k: c_k = 1} fk[k] + qe[c].
            double tb0 = 0, fk[g];
            if (MG || E) {
                if (r == 0) {
                    ising_full(base, I.n, I.ne, P, hz, jzz);
                } else {
                    uint64_t s = prev;
                    for (uint64_t d = base ^ prev; d; d &= d - 1) {
                        const uint32_t q = uint32_t(__builtin_ctzll(d));
                        ising_flip(s, q, P, hz, jzz);
                        s ^= 1ull << q;
                    }
                }
                prev = base;
                tb0 = MG ? cost_of(hz, jzz) : energy_of(hz, jzz) - qe[0];
#pragma unroll
                for (int k = 0; k < g; ++k) {
                    const uint32_t i = U.qlo + k;
                    double nk = 0;
                    for (uint32_t a = P.aoff[i]; a < P.aoff[i + 1]; ++a) {
                        const uint32_t j = P.aj[a];
                        if (j < U.qlo || j >= U.qlo + g) nk += ((base >> j) & 1) ? -adj_J(P, a) : adj_J(P, a);
                    }
                    fk[k] = MG ? 0.5 * P.h[i] + nk : 2.0 * (P.h[i] + nk);
                }
            }
            rx_reg<g>(x, fc, fs);
            if (MG) {
                // exp(-i gamma' C(base | c)): the product over the set bits
                // of c runs as a Gray walk over the high bits of c and, from
                // each of its steps, one over the low GL bits (each walk
                // multiplies by ek or its conjugate as the bit is set or
                // cleared; at most 2^GL - 1 steps from a fresh start).
                float2 ek[g];
#pragma unroll
                for (int k = 0; k < g; ++k) ek[k] = expmi(-2.0 * gn * fk[k]);  // exp(2 i gamma' f_k)
                auto step = [&](float2 ph, int k, bool set) {
                    const float2 f = ek[k];
                    const float fy = set ? f.y : -f.y;
                    return make_float2(ph.x * f.x - ph.y * fy, ph.x * fy + ph.y * f.x);
                };
                float2 pho = expmi(gn * tb0);
#pragma unroll
                for (int mh = 0; mh < DH; ++mh) {
                    const int gh = mh ^ (mh >> 1);
                    if (mh) {
                        const int k = __builtin_ctz(uint32_t(mh));
                        pho = step(pho, GL + k, (gh >> k) & 1);
                    }
                    float2 ph = pho;
#pragma unroll
                    for (int ml = 0; ml < DL; ++ml) {
                        const int gl = ml ^ (ml >> 1);
                        if (ml) {
                            const int k = __builtin_ctz(uint32_t(ml));
                            ph = step(ph, k, (gl >> k) & 1);
                        }
                        const int c = (gh << GL) | gl;
                        const float2 qm = qph[c];
                        const float wr = ph.x * qm.x - ph.y * qm.y, wi = ph.x * qm.y + ph.y * qm.x;
                        const float2 v = x[c];
                        x[c] = make_float2(wr * v.x - wi * v.y, wr * v.y + wi * v.x);
                    }
                }
                rx_reg<g>(x, fc2, fs2);
            }
            if (E) {
                // sum over the orbit of |x_c|^2 E(base | c) along a Gray walk
                // over c (setting bit k adds fk[k], clearing subtracts it).
                // The table index passes through an opaque zero every 8
                // steps, so that the compiler cannot hoist all 2^g table
                // reads (2 VGPRs each) above the walk.
                double dsum = 0;
                uint32_t zq = 0;
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    const int c = m ^ (m >> 1);
                    if (m) {
                        const int k = __builtin_ctz(uint32_t(m));
                        dsum += ((c >> k) & 1) ? fk[k] : -fk[k];
                    }
                    if ((m & 7) == 0) asm volatile("" : "+v"(zq));
                    acc += (double(x[c].x) * x[c].x + double(x[c].y) * x[c].y) * (tb0 + dsum + qe[uint32_t(c) | zq]);
                }
            }
            // The store addresses are recomputed from a copy of `base` the
            // compiler cannot see through, so that it does not keep all 2^g
            // 64-bit load addresses live across the rotations (for g = 6 that
            // alone would be 128 VGPRs).
            uint64_t sbase = pb;
            asm volatile("" : "+v"(sbase));
#pragma unroll
            for (int c = 0; c < D; ++c) stamp<float, true>(psi + (sbase ^ row(c)), x[c]);
        }
        if (E) {
            const double tot = block_sum(acc, red);
            if (threadIdx.x == 0) partial[U.slot] = tot;
        }
        __syncthreads();
    }
}

__global__ void k_qaoa_reset(uint32_t* tickets, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) tickets[i] = 0;
}

// Fixed-order sum of each large instance's energy partials.
__global__ __launch_bounds__(256) void k_qaoa_energy(const QInst* __restrict__ insts, const uint32_t* __restrict__ large,
                                                     uint32_t nlarge, const double* __restrict__ partial,
                                                     double* __restrict__ energy) {
    __shared__ double red[4];
    for (uint32_t k = blockIdx.x; k < nlarge; k += gridDim.x) {
        const QInst I = insts[large[k]];
        double acc = 0;
        for (uint32_t i = threadIdx.x; i < I.npslots; i += 256) acc += partial[I.pslot0 + i];
        const double t = block_sum(acc, red);
        if (threadIdx.x == 0) energy[large[k]] = t;
    }
}

}  // namespace qcs

using namespace qcs;

struct qcs_qaoa {
    int device = 0;
    uint32_t ninst = 0, p = 0;
    std::vector<uint32_t> nq, ne;
    uint64_t total_amps = 0;
    float2* arena = nullptr;
    QInst* d_inst = nullptr;
    uint8_t *d_ei = nullptr, *d_ej = nullptr, *d_aj = nullptr;
    uint16_t* d_aoff = nullptr;
    double *d_J = nullptr, *d_h = nullptr;
    uint8_t* d_ae = nullptr;
    double *d_ang = nullptr, *d_energy = nullptr, *d_partial = nullptr;
    uint32_t* d_large = nullptr;
    uint32_t nlarge = 0;
    // launch plan: one persistent launch per entry, in order
    enum : uint8_t { L_SMALL = 0, L_TILE = 1, L_MIX = 2, L_TILE_RX = 3 };
    struct Launch {
        uint8_t kind, g, mode;  // mode: MIX_* for mixer launches
        uint32_t nunits;
        QUnit* units;
    };
    std::vector<Launch> plan;
    uint32_t nstages = 0;  // dependent steps of one evaluation
    uint32_t* d_tickets = nullptr;
    uint32_t tmax = 12, group = 6;
    int use_graph = 1;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
    std::vector<QInst> inst_h;

    IsingArgsDev args() const { return IsingArgsDev{d_h, d_J, d_ei, d_ej, d_aoff, d_aj, d_ae}; }
};

static void free_plan(qcs_qaoa* q) {
    for (auto& L : q->plan) (void)hipFree(L.units);
    q->plan.clear();
    if (q->d_tickets) (void)hipFree(q->d_tickets);
    q->d_tickets = nullptr;
    if (q->d_partial) (void)hipFree(q->d_partial);
    q->d_partial = nullptr;
    if (q->d_large) (void)hipFree(q->d_large);
    q->d_large = nullptr;
    if (q->exec) (void)hipGraphExecDestroy(q->exec);
    if (q->graph) (void)hipGraphDestroy(q->graph);
    q->exec = nullptr;
    q->graph = nullptr;
}

static int launch_sequence(qcs_qaoa* q) {
    const uint32_t nl = uint32_t(q->plan.size());
    hipLaunchKernelGGL(k_qaoa_reset, dim3(1), dim3(256), 0, q->stream, q->d_tickets, nl);
    const IsingArgsDev ga = q->args();
    for (uint32_t k = 0; k < nl; ++k) {
        const auto& L = q->plan[k];
        const unsigned grid = std::min<uint32_t>(L.nunits, 2048);
        uint32_t* tk = q->d_tickets + k;
        if (L.kind == qcs_qaoa::L_TILE_RX) {
            const size_t shm = sizeof(float2) * tile_lds_elems(q->tmax);
            hipLaunchKernelGGL(k_qaoa_tile_rx, dim3(grid), dim3(256), shm, q->stream, q->arena, q->d_inst, L.units,
                               L.nunits, tk, q->d_ang, q->p);
            continue;
        }
        if (L.kind == qcs_qaoa::L_SMALL || L.kind == qcs_qaoa::L_TILE) {
            uint32_t tb = q->tmax;
            if (L.kind == qcs_qaoa::L_SMALL) {
                tb = 0;
                for (uint32_t i = 0; i < q->ninst; ++i)
                    if (q->nq[i] <= q->tmax) tb = std::max(tb, q->nq[i]);
            }
            const size_t shm = sizeof(float2) * tile_lds_elems(tb);
            if (L.kind == qcs_qaoa::L_SMALL)
                hipLaunchKernelGGL((k_qaoa_tile<true>), dim3(grid), dim3(256), shm, q->stream, q->arena, q->d_inst,
                                   L.units, L.nunits, tk, q->d_ang, q->p, ga, q->d_energy);
            else
                hipLaunchKernelGGL((k_qaoa_tile<false>), dim3(grid), dim3(256), shm, q->stream, q->arena, q->d_inst,
                                   L.units, L.nunits, tk, q->d_ang, q->p, ga, q->d_energy);
            continue;
        }
#define QCS_MIX_MODE(GG, MM)                                                                                    \
    hipLaunchKernelGGL((k_qaoa_mix<GG, MM>), dim3(grid), dim3(256), 0, q->stream, q->arena, q->d_inst, L.units, \
                       L.nunits, tk, q->d_ang, q->p, ga, q->d_partial)
#define QCS_MIX(GG)                                                       \
    do {                                                                  \
        if (L.mode == MIX_ENERGY) QCS_MIX_MODE(GG, MIX_ENERGY);           \
        else if (L.mode == MIX_MERGE) QCS_MIX_MODE(GG, MIX_MERGE);        \
        else QCS_MIX_MODE(GG, MIX_PLAIN);                                 \
    } while (0)
        switch (L.g) {
            case 1: QCS_MIX(1); break;
            case 2: QCS_MIX(2); break;
            case 3: QCS_MIX(3); break;
            case 4: QCS_MIX(4); break;
            case 5: QCS_MIX(5); break;
            default: QCS_MIX(6); break;
        }
#undef QCS_MIX
#undef QCS_MIX_MODE
    }
    if (q->nlarge)
        hipLaunchKernelGGL(k_qaoa_energy, dim3(std::min<uint32_t>(q->nlarge, 1024)), dim3(256), 0, q->stream,
                           q->d_inst, q->d_large, q->nlarge, q->d_partial, q->d_energy);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "qaoa launch failed: %s", hipGetErrorString(e));
}

// One pass of a large instance over its amplitudes.
struct PassSpec {
    uint8_t kind;   // qcs_qaoa::L_TILE or L_MIX
    uint8_t g;      // mixer group width
    uint8_t mode;   // mixer: MIX_*; tile: 0 or kTileRxOnly
    uint32_t qlo;   // mixer: first qubit of the group
    uint32_t layer;
};

// The passes of one evaluation of an instance of n > T qubits whose qubits
// above the tile form ng mixer groups. ng = 1: per layer the tile pass (cost
// diagonal, Rx on the tile qubits) and the group's Rx. ng >= 2: the last two
// groups X = ng - 1 and Y = ng - 2 take turns (X first) at ending a layer
// with a merged pass (Rx of layer l, cost of layer l + 1, Rx of layer l + 1);
// the next layer's tile pass then runs without the cost and that layer
// skips the merged group. That is p ng + 1 passes instead of p (ng + 1): 9
// instead of 12 for two groups at depth 4. Every layer's cost runs before
// its rotations, and every rotation of a layer before the next layer's cost.
// (Starting instead with a mixer pass that generates |+...+>, applies layer
// 0's cost and its group's Rx, so that every tile pass only rotates, takes
// as many passes and measured 1.5 ms slower per 1024-instance batch.)
static std::vector<PassSpec> instance_passes(uint32_t n, uint32_t Tm, uint32_t G, uint32_t p) {
    const uint32_t ng = (n - Tm + G - 1) / G;
    auto width = [&](uint32_t j) { return uint8_t(j == ng - 1 ? (n - Tm) - G * (ng - 1) : G); };
    auto mix = [&](uint32_t j, uint8_t mode, uint32_t l) {
        return PassSpec{qcs_qaoa::L_MIX, width(j), mode, Tm + j * G, l};
    };
    std::vector<PassSpec> v;
    int merged = -1;  // the group whose Rx of this layer ran in the previous layer's merged pass
    for (uint32_t l = 0; l < p; ++l) {
        v.push_back(PassSpec{qcs_qaoa::L_TILE, 0, uint8_t(merged >= 0 ? kTileRxOnly : 0), 0, l});
        const uint32_t z = ng >= 2 ? (l % 2 == 0 ? ng - 1 : ng - 2) : 0;
        for (uint32_t j = 0; j < ng; ++j)
            if (int(j) != merged && j != z) v.push_back(mix(j, MIX_PLAIN, l));
        const bool last = l + 1 == p;
        v.push_back(mix(z, last ? MIX_ENERGY : (ng >= 2 ? MIX_MERGE : MIX_PLAIN), l));
        merged = (!last && ng >= 2) ? int(z) : -1;
    }
    return v;
}

// Build the launch plan for (tmax, group) and capture it.
static int build_plan(qcs_qaoa* q) {
    free_plan(q);
    const uint32_t G = q->group, Tm = q->tmax;
    struct Pending {
        uint8_t kind, g, mode;
        std::vector<QUnit> units;
    };
    std::vector<Pending> pend;
    std::vector<uint32_t> large;
    uint32_t pslot = 0;
    // SMALL instances: one launch, first (independent of the rest).
    for (uint32_t i = 0; i < q->ninst; ++i) {
        QInst& I = q->inst_h[i];
        if (q->nq[i] <= Tm) {
            I.tbits = q->nq[i];
            I.pslot0 = 0;
            I.npslots = 0;
            if (pend.empty()) pend.push_back(Pending{qcs_qaoa::L_SMALL, 0, 0, {}});
            pend.back().units.push_back(QUnit{i, 0, 0, 0, 0, 0, 0});
        }
    }
    for (uint32_t i = 0; i < q->ninst; ++i)
        if (q->nq[i] > Tm) large.push_back(i);
    std::vector<std::vector<PassSpec>> passes;
    size_t nstep = 0;
    for (uint32_t i : large) {
        passes.push_back(instance_passes(q->nq[i], Tm, G, q->p));
        nstep = std::max(nstep, passes.back().size());
        // energy slots: one per unit of the instance's last pass
        QInst& I = q->inst_h[i];
        const uint64_t chunks = (((1ull << q->nq[i]) >> passes.back().back().g) + 255) / 256;
        I.tbits = Tm;
        I.pslot0 = pslot;
        I.npslots = uint32_t((chunks + kMixChunks - 1) / kMixChunks);
        pslot += I.npslots;
    }
    // step s: pass s of every instance that has one, one launch per
    // (kind, width, mode) present
    for (size_t s = 0; s < nstep; ++s) {
        const size_t first = pend.size();
        for (size_t k = 0; k < large.size(); ++k) {
            if (s >= passes[k].size()) continue;
            const PassSpec& S = passes[k][s];
            const uint32_t i = large[k], n = q->nq[i];
            const bool tile = S.kind == qcs_qaoa::L_TILE;
            // Rx-only tile passes run on their own kernel
            const uint8_t kind = tile && (S.mode & kTileRxOnly) ? uint8_t(qcs_qaoa::L_TILE_RX) : S.kind;
            const uint8_t g = tile ? 0 : S.g, md = tile ? 0 : S.mode;
            size_t j = first;
            while (j < pend.size() && !(pend[j].kind == kind && pend[j].g == g && pend[j].mode == md)) ++j;
            if (j == pend.size()) pend.push_back(Pending{kind, g, md, {}});
            auto& P = pend[j];
            if (tile) {
                const uint32_t run = (S.mode & kTileRxOnly) ? kTileRunRx : kTileRun;
                for (uint32_t t = 0, nt = 1u << (n - Tm); t < nt; t += run)
                    P.units.push_back(QUnit{i, t, S.layer, 0, 0, std::min(run, nt - t), S.mode});
            } else {
                const bool en = S.mode == MIX_ENERGY;
                const uint64_t chunks = (((1ull << n) >> S.g) + 255) / 256;
                uint32_t slot = q->inst_h[i].pslot0;
                for (uint64_t c = 0; c < chunks; c += kMixChunks)
                    P.units.push_back(QUnit{i, uint32_t(c), S.layer, S.qlo, en ? slot++ : 0u, 0, 0});
            }
        }
    }
    q->nstages = uint32_t(pend.size());
    if (hipMemcpy(q->d_inst, q->inst_h.data(), sizeof(QInst) * q->ninst, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(QCS_EDEVICE, "instance upload failed");
    for (auto& P : pend) {
        if (P.units.empty()) continue;
        QUnit* d = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(QUnit) * P.units.size()) != hipSuccess)
            return set_error(QCS_ENOMEM, "unit table");
        q->plan.push_back(qcs_qaoa::Launch{P.kind, P.g, P.mode, uint32_t(P.units.size()), d});
        if (hipMemcpy(d, P.units.data(), sizeof(QUnit) * P.units.size(), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(QCS_EDEVICE, "unit upload failed");
    }
    if (hipMalloc(reinterpret_cast<void**>(&q->d_tickets), sizeof(uint32_t) * (q->plan.size() + 1)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&q->d_partial), sizeof(double) * (pslot + 1)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&q->d_large), sizeof(uint32_t) * (large.size() + 1)) != hipSuccess)
        return set_error(QCS_ENOMEM, "qaoa scratch");
    q->nlarge = uint32_t(large.size());
    if (!large.empty() &&
        hipMemcpy(q->d_large, large.data(), sizeof(uint32_t) * large.size(), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(QCS_EDEVICE, "upload failed");
    if (q->use_graph) {
        if (hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return set_error(QCS_EDEVICE, "stream capture failed");
        const int e = launch_sequence(q);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(q->stream, &g);
        if (e) return e;
        if (ce != hipSuccess) return set_error(QCS_EDEVICE, "end capture failed");
        q->graph = g;
        if (hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0) != hipSuccess)
            return set_error(QCS_EDEVICE, "graph instantiation failed");
    }
    return QCS_OK;
}

extern "C" {

int qcs_qaoa_destroy(qcs_qaoa* q) {
    if (!q) return QCS_OK;
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    free_plan(q);
    (void)hipFree(q->arena);
    (void)hipFree(q->d_inst);
    (void)hipFree(q->d_ei);
    (void)hipFree(q->d_ej);
    (void)hipFree(q->d_J);
    (void)hipFree(q->d_h);
    (void)hipFree(q->d_aoff);
    (void)hipFree(q->d_aj);
    (void)hipFree(q->d_ae);
    (void)hipFree(q->d_ang);
    (void)hipFree(q->d_energy);
    if (q->ev0) (void)hipEventDestroy(q->ev0);
    if (q->ev1) (void)hipEventDestroy(q->ev1);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
    return QCS_OK;
}

int qcs_qaoa_create(uint32_t ninst, const uint32_t* nqubits, const uint32_t* nedges, const uint32_t* ei,
                    const uint32_t* ej, const double* J, const double* h, uint32_t p, qcs_qaoa** out) {
    if (!out || !nqubits || !nedges || !h || ninst == 0 || p == 0) return set_error(QCS_EINVAL, "bad argument");
    *out = nullptr;
    auto* q = new qcs_qaoa;
    q->ninst = ninst;
    q->p = p;
    (void)hipGetDevice(&q->device);
    uint32_t eo = 0, ho = 0;
    q->inst_h.resize(ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t n = nqubits[i], m = nedges[i];
        if (n < 1 || n > 30 || m > kMaxEdges) {
            delete q;
            return set_error(QCS_EINVAL, "instance %u: need 1 <= n <= 30 and <= %u edges", i, kMaxEdges);
        }
        for (uint32_t k = 0; k < m; ++k)
            if (ei[eo + k] >= n || ej[eo + k] >= n || ei[eo + k] == ej[eo + k]) {
                delete q;
                return set_error(QCS_EINVAL, "instance %u: bad edge %u", i, k);
            }
        q->nq.push_back(n);
        q->ne.push_back(m);
        QInst& I = q->inst_h[i];
        std::memset(&I, 0, sizeof I);
        I.off = q->total_amps;
        I.n = n;
        I.ne = m;
        I.eoff = eo;
        I.hoff = ho;
        q->total_amps += 1ull << n;
        eo += m;
        ho += n;
    }
    std::vector<uint8_t> e8i(eo + 1), e8j(eo + 1);
    for (uint32_t k = 0; k < eo; ++k) {
        e8i[k] = uint8_t(ei[k]);
        e8j[k] = uint8_t(ej[k]);
    }
    // adjacency lists (CSR per instance): neighbours of each site with the
    // instance's index of the joining edge
    std::vector<uint16_t> aoff(ho + ninst + 1, 0);
    std::vector<uint8_t> aj(2 * eo + 1, 0);
    std::vector<uint8_t> ae(2 * eo + 1, 0);
    uint32_t ao = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        QInst& I = q->inst_h[i];
        const uint32_t n = I.n, m = I.ne, e0 = I.eoff;
        I.aoff = ao;
        I.boff = 2 * e0;
        std::vector<uint32_t> deg(n + 1, 0);
        for (uint32_t k = 0; k < m; ++k) {
            ++deg[ei[e0 + k]];
            ++deg[ej[e0 + k]];
        }
        uint32_t acc = 0;
        for (uint32_t v = 0; v < n; ++v) {
            aoff[ao + v] = uint16_t(acc);
            acc += deg[v];
        }
        aoff[ao + n] = uint16_t(acc);
        std::vector<uint32_t> fill(n, 0);
        for (uint32_t k = 0; k < m; ++k) {
            const uint32_t a = ei[e0 + k], b = ej[e0 + k];
            const uint32_t pa = I.boff + aoff[ao + a] + fill[a]++, pb = I.boff + aoff[ao + b] + fill[b]++;
            aj[pa] = uint8_t(b);
            ae[pa] = uint8_t(k);
            aj[pb] = uint8_t(a);
            ae[pb] = uint8_t(k);
        }
        ao += n + 1;
    }
    auto al = [](void** p, size_t b) { return hipMalloc(p, b ? b : 8) == hipSuccess; };
    if (!al(reinterpret_cast<void**>(&q->arena), sizeof(float2) * q->total_amps) ||
        !al(reinterpret_cast<void**>(&q->d_inst), sizeof(QInst) * ninst) ||
        !al(reinterpret_cast<void**>(&q->d_ei), eo + 1) || !al(reinterpret_cast<void**>(&q->d_ej), eo + 1) ||
        !al(reinterpret_cast<void**>(&q->d_J), sizeof(double) * (eo + 1)) ||
        !al(reinterpret_cast<void**>(&q->d_h), sizeof(double) * (ho + 1)) ||
        !al(reinterpret_cast<void**>(&q->d_aoff), sizeof(uint16_t) * aoff.size()) ||
        !al(reinterpret_cast<void**>(&q->d_aj), aj.size()) ||
        !al(reinterpret_cast<void**>(&q->d_ae), ae.size()) ||
        !al(reinterpret_cast<void**>(&q->d_ang), sizeof(double) * 2 * p * ninst) ||
        !al(reinterpret_cast<void**>(&q->d_energy), sizeof(double) * ninst)) {
        (void)hipGetLastError();
        qcs_qaoa_destroy(q);
        return set_error(QCS_ENOMEM, "qaoa allocation failed (%llu amplitudes)", (unsigned long long)q->total_amps);
    }
    if (hipMemcpy(q->d_ei, e8i.data(), eo + 1, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(q->d_ej, e8j.data(), eo + 1, hipMemcpyHostToDevice) != hipSuccess ||
        (eo && hipMemcpy(q->d_J, J, sizeof(double) * eo, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(q->d_h, h, sizeof(double) * ho, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(q->d_aoff, aoff.data(), sizeof(uint16_t) * aoff.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(q->d_aj, aj.data(), aj.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(q->d_ae, ae.data(), ae.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&q->ev0) != hipSuccess || hipEventCreate(&q->ev1) != hipSuccess) {
        qcs_qaoa_destroy(q);
        return set_error(QCS_EDEVICE, "qaoa setup failed");
    }
    const int e = build_plan(q);
    if (e) {
        qcs_qaoa_destroy(q);
        return e;
    }
    *out = q;
    return QCS_OK;
}

int qcs_qaoa_configure(qcs_qaoa* q, uint32_t tile_bits, uint32_t group, int use_graph) {
    if (!q) return set_error(QCS_EINVAL, "null batch");
    if (tile_bits < 6 || tile_bits > 12) return set_error(QCS_EINVAL, "tile bits must be 6..12");
    if (group < 4 || group > 6) return set_error(QCS_EINVAL, "group width must be 4..6");
    if (hipStreamSynchronize(q->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
    q->tmax = tile_bits;
    q->group = group;
    q->use_graph = use_graph;
    return build_plan(q);
}

int qcs_qaoa_evaluate(qcs_qaoa* q, const double* gamma, const double* beta, double* energy) {
    if (!q || !gamma || !beta || !energy) return set_error(QCS_EINVAL, "null argument");
    std::vector<double> ang(size_t(2) * q->p * q->ninst);
    for (uint32_t i = 0; i < q->ninst; ++i)
        for (uint32_t l = 0; l < q->p; ++l) {
            ang[(size_t(i) * q->p + l) * 2] = gamma[size_t(i) * q->p + l];
            ang[(size_t(i) * q->p + l) * 2 + 1] = beta[size_t(i) * q->p + l];
        }
    if (hipMemcpyAsync(q->d_ang, ang.data(), sizeof(double) * ang.size(), hipMemcpyHostToDevice, q->stream) !=
        hipSuccess)
        return set_error(QCS_EDEVICE, "angle upload failed");
    (void)hipEventRecord(q->ev0, q->stream);
    int e = QCS_OK;
    if (q->use_graph && q->exec) {
        if (hipGraphLaunch(q->exec, q->stream) != hipSuccess) e = set_error(QCS_EDEVICE, "graph launch failed");
    } else {
        e = launch_sequence(q);
    }
    (void)hipEventRecord(q->ev1, q->stream);
    if (e) return e;
    if (hipMemcpyAsync(energy, q->d_energy, sizeof(double) * q->ninst, hipMemcpyDeviceToHost, q->stream) !=
            hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "qaoa evaluation failed");
    (void)hipEventElapsedTime(&q->last_ms, q->ev0, q->ev1);
    return QCS_OK;
}

int qcs_qaoa_last_ms(qcs_qaoa* q, float* ms) {
    if (!q || !ms) return set_error(QCS_EINVAL, "null argument");
    *ms = q->last_ms;
    return QCS_OK;
}

int qcs_qaoa_get_state(qcs_qaoa* q, uint32_t inst, double* amps) {
    if (!q || !amps || inst >= q->ninst) return set_error(QCS_EINVAL, "bad argument");
    const uint64_t A = 1ull << q->nq[inst];
    std::vector<float2> tmp(A);
    if (hipMemcpyAsync(tmp.data(), q->arena + q->inst_h[inst].off, sizeof(float2) * A, hipMemcpyDeviceToHost,
                       q->stream) != hipSuccess ||
        hipStreamSynchronize(q->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "state download failed");
    for (uint64_t i = 0; i < A; ++i) {
        const float2 v = tmp[i ^ arena_xor(i)];
        amps[2 * i] = v.x;
        amps[2 * i + 1] = v.y;
    }
    return QCS_OK;
}

int qcs_qaoa_plan(uint32_t n, uint32_t tile_bits, uint32_t group, uint32_t p, uint32_t* out, uint32_t cap,
                  uint32_t* npasses) {
    if (!npasses || (cap && !out) || n == 0 || n > 40 || p == 0) return set_error(QCS_EINVAL, "bad argument");
    if (tile_bits < 6 || tile_bits > 12) return set_error(QCS_EINVAL, "tile bits must be 6..12");
    if (group < 4 || group > 6) return set_error(QCS_EINVAL, "group width must be 4..6");
    if (n <= tile_bits) {
        if (cap) {
            const uint32_t w[5] = {0, 0, 0, 0, 0};
            std::memcpy(out, w, sizeof w);
        }
        *npasses = 1;
        return QCS_OK;
    }
    const std::vector<PassSpec> v = instance_passes(n, tile_bits, group, p);
    for (size_t k = 0; k < v.size() && k < cap; ++k) {
        const PassSpec& S = v[k];
        const uint32_t w[5] = {S.kind == qcs_qaoa::L_TILE ? 1u : 2u, S.g, S.mode, S.qlo, S.layer};
        std::memcpy(out + 5 * k, w, sizeof w);
    }
    *npasses = uint32_t(v.size());
    return QCS_OK;
}

int qcs_qaoa_info(qcs_qaoa* q, uint32_t* stages, uint64_t* amplitudes, uint32_t* tile_bits, uint32_t* group) {
    if (!q) return set_error(QCS_EINVAL, "null batch");
    if (stages) *stages = uint32_t(q->plan.size());
    if (amplitudes) *amplitudes = q->total_amps;
    if (tile_bits) *tile_bits = q->tmax;
    if (group) *group = q->group;
    return QCS_OK;
}

}  // extern "C"
