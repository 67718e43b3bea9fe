// qft.hip - quantum Fourier transform on a contiguous qubit range [sq, eq].
//
// The reference builds QFT/IQFT from Hadamards, controlled phase shifts and a
// SWAP network (QuantumFourierTransform.h:35-87, QubitsSwapper.h:23-33): at
// n = 30 that is 480 full passes over the state. Here the same operator is
// applied in about (n - 12)/6 + 2 passes (5 at n = 30):
//
//   The controlled phase shifts between target `cur` and every control below
//   it (phase pi/2, pi/4, ... for distance 1, 2, ...) multiply to one
//   diagonal D_cur: amplitudes with bit cur set get exp(i pi L / 2^(cur-sq)),
//   L = bits sq..cur-1 of the index. The forward sequence is then
//   H(eq), D_eq, H(eq-1), D_(eq-1), ..., D_(sq+1), H(sq) and the inverse is
//   H(sq), D*_(sq+1), H(sq+1), ..., D*_eq, H(eq) (after the swaps).
//
//   k_qft_group applies that sequence for up to 4 consecutive targets at once
//   with the 2^4 amplitudes of each orbit in registers, k_qft_gpair for 5 or
//   6 with each orbit split over a lane pair; k_qft_lds applies it for all
//   targets below 12 inside 4096-amplitude chunks staged in LDS, in rounds
//   of up to 3 targets held in registers; the
//   SWAP network is one bit-reversal pass (k_bitrev_tile, LDS-tiled, or
//   k_bitrev_pairs when runs of >= 8 amplitudes stay contiguous).
//
// Phases come from three 4096-entry tables: exp(i pi x / 2^36) for
// x = x2 2^24 + x1 2^12 + x0 is T2[x2] T1[x1] T0[x0], computed on the host in
// long double; the kernels need no transcendental functions.
// qft_sequence() issues the reference's exact gate sequence instead.
#include <cmath>
#include <cstdlib>
#include <complex>
#include <mutex>
#include <vector>

#include "common.h"

namespace qcs {

constexpr int kTwBits = 12;
constexpr int kTwLen = 1 << kTwBits;
constexpr int kPhaseBits = 36;  // phases are multiples of pi / 2^36
constexpr int kLdsLog = 12;     // k_qft_lds chunk: 4096 amplitudes
constexpr int kLdsThreads = 512; // k_qft_lds block: one 8-amplitude orbit per thread and round
constexpr int kLdsRound = 3;    // targets per k_qft_lds round (8 amplitudes per thread)
constexpr int kGroupMax = 6;   // widest register group; roots exp(i pi j / 2^kGroupMax)

// ---- twiddle tables (process-wide, one device copy per device) -------------
static std::mutex g_tw_mu;
static double2* g_tw[64] = {nullptr};

static double2* twiddles(int device) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (device < 0 || device >= 64) return nullptr;
    if (g_tw[device]) return g_tw[device];
    std::vector<double2> h(3 * kTwLen);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int lvl = 0; lvl < 3; ++lvl) {
        // level 2: pi j / 2^12, level 1: pi j / 2^24, level 0: pi j / 2^36
        const long double scale = pi / std::ldexp(1.0L, kTwBits * (3 - lvl));
        for (int j = 0; j < kTwLen; ++j) {
            const long double a = scale * j;
            h[lvl * kTwLen + j] = make_double2(double(std::cos(a)), double(std::sin(a)));
        }
    }
    double2* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(double2) * h.size()) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_tw[device] = d;
    return d;
}

// exp(i pi L / 2^k) for integer 0 <= L < 2^k, k <= 36, in double.
__device__ __forceinline__ double2 phase_pi(const double2* __restrict__ tw, uint64_t L, uint32_t k) {
    const uint64_t x = L << (kPhaseBits - k);
    const double2 a = tw[2 * kTwLen + ((x >> 24) & (kTwLen - 1))];
    const uint64_t lo = x & ((1ull << 24) - 1);
    if (lo == 0) return a;
    const double2 b = tw[kTwLen + (lo >> 12)];
    const double2 c = tw[lo & (kTwLen - 1)];
    double2 ab;
    ab.x = a.x * b.x - a.y * b.y;
    ab.y = a.x * b.y + a.y * b.x;
    double2 r;
    r.x = ab.x * c.x - ab.y * c.y;
    r.y = ab.x * c.y + ab.y * c.x;
    return r;
}

template <typename T>
__device__ __forceinline__ cv<T> mul_ph(cv<T> v, double2 p, bool conj) {
    const T pr = T(p.x), pi = conj ? T(-p.y) : T(p.y);
    cv<T> r;
    r.x = pr * v.x - pi * v.y;
    r.y = pr * v.y + pi * v.x;
    return r;
}

// Unnormalised Hadamard butterfly (a + b, a - b): each pass applies the
// 2^(-1/2) of all its Hadamards once, to every amplitude it stores
// (hscale), instead of once per butterfly.
template <typename T>
__device__ __forceinline__ void hadamard(cv<T>& a, cv<T>& b) {
    const cv<T> x = a, y = b;
    a.x = x.x + y.x;
    a.y = x.y + y.y;
    b.x = x.x - y.x;
    b.y = x.y - y.y;
}

// 2^(-h/2): the normalisation of h Hadamards.
template <typename T>
__device__ __forceinline__ T hscale(uint32_t h) {
    const T p = T(ldexp(1.0, -int(h / 2)));
    return (h & 1) ? p * T(0.70710678118654752440) : p;
}

template <typename T>
__device__ __forceinline__ cv<T> scaled(cv<T> v, T f) {
    v.x *= f;
    v.y *= f;
    return v;
}

struct QftArgs {
    uint32_t sq, eq, lo, g;  // range and this pass's targets [lo, lo+g-1]
};

// exp(i pi j / 32), j < 32: the phases between targets of one group, passed
// as a kernel argument so that every lookup is a uniform constant.
struct QftRoots {
    double2 w[1 << kGroupMax];
};

static QftRoots group_roots() {
    QftRoots q;
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int j = 0; j < (1 << kGroupMax); ++j) {
        const long double a = pi * j / (1 << kGroupMax);
        q.w[j] = make_double2(double(std::cos(a)), double(std::sin(a)));
    }
    return q;
}

// D_(t+1) (or its conjugate) on local element c of an orbit whose first
// index is `base`; the group occupies bits [lo, lo+G).
template <typename T, int G>
__device__ __forceinline__ void apply_d(cv<T>* x, uint64_t base, uint32_t t, const QftArgs& a,
                                        const double2* __restrict__ tw, const QftRoots& R, bool conj) {
    constexpr int D = 1 << G;
    const uint32_t k = t + 1 - a.sq;             // phase denominator exponent
    const uint32_t nlow = a.lo - a.sq;            // fixed bits sq..lo-1 below the group
    const uint64_t Lf = nlow ? ((base >> a.sq) & ((1ull << nlow) - 1)) : 0;
    const double2 F = phase_pi(tw, Lf, k);         // exp(i pi Lf / 2^k)
    const uint32_t tl = t - a.lo;                  // local bit of target t
    const bool top_fixed = (t + 1 >= a.lo + G);    // bit t+1 is outside the group
    if (top_fixed && !((base >> (t + 1)) & 1)) return;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        if (!top_fixed && !((c >> (tl + 1)) & 1)) continue;
        const uint64_t Gt = uint64_t(c) & ((2ull << tl) - 1);  // bits lo..t
        // exp(i pi (Gt 2^nlow + Lf) / 2^k) = exp(i pi Gt / 2^(t+1-lo)) * F,
        // and exp(i pi Gt / 2^(tl+1)) = R.w[Gt << (kGroupMax - 1 - tl)]
        const double2 p = R.w[Gt << (kGroupMax - 1 - tl)];
        double2 f;
        f.x = p.x * F.x - p.y * F.y;
        f.y = p.x * F.y + p.y * F.x;
        x[c] = mul_ph<T>(x[c], f, conj);
    }
}

template <typename T, int G>
__device__ __forceinline__ void apply_h(cv<T>* x, uint32_t tl) {
    constexpr int D = 1 << G;
#pragma unroll
    for (int c = 0; c < D; ++c)
        if (!((c >> tl) & 1)) hadamard<T>(x[c], x[c | (1 << tl)]);
}

template <typename T, int G, bool INV>
__global__ __launch_bounds__(256) void k_qft_group(cv<T>* __restrict__ psi, uint64_t norb, QftArgs a,
                                                   const double2* __restrict__ tw, const QftRoots R) {
    constexpr int D = 1 << G;
    const uint64_t o = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (o >= norb) return;
    const uint64_t base = insert_run(o, a.lo, G);
    // orbit element c sits at base + c 2^lo (bits lo..lo+G-1 of base are 0)
    cv<T>* const p0 = psi + base;
    const uint64_t stride = 1ull << a.lo;
    cv<T> x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = ldamp<T, kStagedNtLoads>(p0 + c * stride);
    if (!INV) {
#pragma unroll
        for (int tl = G - 1; tl >= 0; --tl) {
            const uint32_t t = a.lo + tl;
            if (t < a.eq) apply_d<T, G>(x, base, t, a, tw, R, false);
            apply_h<T, G>(x, tl);
            // keep each target's phase loads next to their use (register
            // pressure of the 2^G-amplitude orbit)
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int tl = 0; tl < G; ++tl) {
            const uint32_t t = a.lo + tl;
            apply_h<T, G>(x, tl);
            if (t < a.eq) apply_d<T, G>(x, base, t, a, tw, R, true);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    cv<T>* sp = p0;  // store addresses from an opaque copy (see k_qft_gpair)
    asm volatile("" : "+v"(sp));
#pragma unroll
    for (int c = 0; c < D; ++c) stamp<T, true>(sp + c * stride, scaled<T>(x[c], hscale<T>(G)));
}

// G = 5 or 6 targets per pass with each 2^G-amplitude orbit split over a
// lane pair: lane half h holds orbit elements c + 2^(G-1) h (c < 2^(G-1)), so
// a thread keeps 16 or 32 amplitudes in registers; the target whose bit is
// the pair bit (local bit G - 1) exchanges values with the partner lane.
template <typename T, int G>
__device__ __forceinline__ void apply_d_half(cv<T>* x, uint64_t base, uint32_t t, uint32_t half, const QftArgs& a,
                                             const double2* __restrict__ tw, const QftRoots& R, bool conj) {
    constexpr int H = 1 << (G - 1);
    const uint32_t k = t + 1 - a.sq;
    const uint32_t nlow = a.lo - a.sq;
    const uint64_t Lf = nlow ? ((base >> a.sq) & ((1ull << nlow) - 1)) : 0;
    const double2 F = phase_pi(tw, Lf, k);
    const uint32_t tl = t - a.lo;
    const bool top_fixed = (t + 1 >= a.lo + G);
    if (top_fixed && !((base >> (t + 1)) & 1)) return;
#pragma unroll
    for (int c = 0; c < H; ++c) {
        const uint32_t cl = uint32_t(c) | (half << (G - 1));
        if (!top_fixed && !((cl >> (tl + 1)) & 1)) continue;
        const uint32_t Gt = cl & ((2u << tl) - 1);
        // x F, then times the root (a wave-uniform constant): no per-element
        // phase is formed, which keeps 32 amplitudes per lane within budget
        x[c] = mul_ph<T>(mul_ph<T>(x[c], F, conj), R.w[Gt << (kGroupMax - 1 - tl)], conj);
    }
}

// The value of lane l ^ 1: a DPP row move (quad_perm 1, 0, 3, 2) per 32-bit
// half, in the VALU instead of an LDS-crossbar permute.
__device__ __forceinline__ float lane_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_xor1(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b)), 0xB1, 0xF, 0xF, false));
    const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), 0xB1, 0xF, 0xF, false));
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

template <typename T, int H>
__device__ __forceinline__ void h_pair(cv<T>* x, uint32_t half) {
#pragma unroll
    for (int c = 0; c < H; ++c) {
        const T yx = lane_xor1(x[c].x), yy = lane_xor1(x[c].y);
        if (half) {
            x[c].x = yx - x[c].x;
            x[c].y = yy - x[c].y;
        } else {
            x[c].x = x[c].x + yx;
            x[c].y = x[c].y + yy;
        }
    }
}

template <typename T, int G, bool INV>
__global__ __launch_bounds__(256) void k_qft_gpair(cv<T>* __restrict__ psi, uint64_t norb, QftArgs a,
                                                   const double2* __restrict__ tw, const QftRoots R) {
    constexpr int H = 1 << (G - 1);
    const uint64_t gt = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    const uint64_t o = gt >> 1;
    const uint32_t half = threadIdx.x & 1;
    if (o >= norb) return;  // both lanes of a pair leave together
    const uint64_t base = insert_run(o, a.lo, G);
    cv<T>* const p0 = psi + base + (uint64_t(half) << (a.lo + G - 1));
    const uint64_t stride = 1ull << a.lo;
    cv<T> x[H];
#pragma unroll
    for (int c = 0; c < H; ++c) x[c] = ldamp<T, kStagedNtLoads>(p0 + c * stride);
    if (!INV) {
#pragma unroll
        for (int tl = G - 1; tl >= 0; --tl) {
            const uint32_t t = a.lo + tl;
            if (t < a.eq) apply_d_half<T, G>(x, base, t, half, a, tw, R, false);
            if (tl == G - 1)
                h_pair<T, H>(x, half);
            else
                apply_h<T, G - 1>(x, tl);
            // keep each target's phase loads next to their use (register
            // pressure of the 2^(G-1) amplitudes per lane)
            if constexpr (G == 6) __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int tl = 0; tl < G; ++tl) {
            const uint32_t t = a.lo + tl;
            if (tl == G - 1)
                h_pair<T, H>(x, half);
            else
                apply_h<T, G - 1>(x, tl);
            if (t < a.eq) apply_d_half<T, G>(x, base, t, half, a, tw, R, true);
            if constexpr (G == 6) __builtin_amdgcn_sched_barrier(0);
        }
    }
    // store addresses from a copy of p0 the compiler cannot see through, so
    // that it does not keep all 2^(G-1) 64-bit load addresses live across
    // the transform
    cv<T>* sp = p0;
    asm volatile("" : "+v"(sp));
#pragma unroll
    for (int c = 0; c < H; ++c) stamp<T, true>(sp + c * stride, scaled<T>(x[c], hscale<T>(G)));
}

// LDS position of chunk element i: bits 0..3 XOR bits 4..7, so that the
// 16 lanes of one access phase hit different banks for every round layout.
__device__ __forceinline__ uint32_t swz(uint32_t i) { return i ^ ((i >> 4) & 15u); }

// x * exp(i pi J / 64) (conjugate with CONJ) for a root index J known at
// compile time: J = 0 is the identity, J = 32 is a multiplication by i.
template <typename T, int J, bool CONJ>
__device__ __forceinline__ cv<T> mul_root(cv<T> x, const QftRoots& R) {
    static_assert(J >= 0 && J < (1 << kGroupMax), "root index");
    if constexpr (J == 0) {
        return x;
    } else if constexpr (J == (1 << (kGroupMax - 1))) {
        cv<T> r;
        r.x = CONJ ? x.y : -x.y;
        r.y = CONJ ? -x.x : x.x;
        return r;
    } else {
        return mul_ph<T>(x, R.w[J], CONJ);
    }
}

// Intra-group phase of target tl on element c (the part of D_(lo+tl+1) that
// depends on the group bits lo..lo+tl): exp(i pi Gt / 2^(tl+1)),
// Gt = c & (2^(tl+1) - 1), for elements whose bit tl+1 is set.
template <typename T, int G, int TL, bool CONJ>
__device__ __forceinline__ void apply_p(cv<T>* x, const QftRoots& R) {
#pragma unroll
    for (int c = 0; c < (1 << G); ++c) {
        if (!((c >> (TL + 1)) & 1)) continue;
        constexpr int shift = kGroupMax - 1 - TL;
        const int Gt = c & ((2 << TL) - 1);
        switch (Gt << shift) {  // the index is a compile-time constant per unrolled c
#define QCS_ROOT_CASE(J) \
    case J: x[c] = mul_root<T, J, CONJ>(x[c], R); break;
            QCS_ROOT_CASE(0) QCS_ROOT_CASE(1) QCS_ROOT_CASE(2) QCS_ROOT_CASE(3) QCS_ROOT_CASE(4) QCS_ROOT_CASE(5)
            QCS_ROOT_CASE(6) QCS_ROOT_CASE(7) QCS_ROOT_CASE(8) QCS_ROOT_CASE(9) QCS_ROOT_CASE(10) QCS_ROOT_CASE(11)
            QCS_ROOT_CASE(12) QCS_ROOT_CASE(13) QCS_ROOT_CASE(14) QCS_ROOT_CASE(15) QCS_ROOT_CASE(16)
            QCS_ROOT_CASE(17) QCS_ROOT_CASE(18) QCS_ROOT_CASE(19) QCS_ROOT_CASE(20) QCS_ROOT_CASE(21)
            QCS_ROOT_CASE(22) QCS_ROOT_CASE(23) QCS_ROOT_CASE(24) QCS_ROOT_CASE(25) QCS_ROOT_CASE(26)
            QCS_ROOT_CASE(27) QCS_ROOT_CASE(28) QCS_ROOT_CASE(29) QCS_ROOT_CASE(30) QCS_ROOT_CASE(31)
            QCS_ROOT_CASE(32) QCS_ROOT_CASE(33) QCS_ROOT_CASE(34) QCS_ROOT_CASE(35) QCS_ROOT_CASE(36)
            QCS_ROOT_CASE(37) QCS_ROOT_CASE(38) QCS_ROOT_CASE(39) QCS_ROOT_CASE(40) QCS_ROOT_CASE(41)
            QCS_ROOT_CASE(42) QCS_ROOT_CASE(43) QCS_ROOT_CASE(44) QCS_ROOT_CASE(45) QCS_ROOT_CASE(46)
            QCS_ROOT_CASE(47) QCS_ROOT_CASE(48) QCS_ROOT_CASE(49) QCS_ROOT_CASE(50) QCS_ROOT_CASE(51)
            QCS_ROOT_CASE(52) QCS_ROOT_CASE(53) QCS_ROOT_CASE(54) QCS_ROOT_CASE(55) QCS_ROOT_CASE(56)
            QCS_ROOT_CASE(57) QCS_ROOT_CASE(58) QCS_ROOT_CASE(59) QCS_ROOT_CASE(60) QCS_ROOT_CASE(61)
            QCS_ROOT_CASE(62) QCS_ROOT_CASE(63)
#undef QCS_ROOT_CASE
            default: break;
        }
    }
}

// The top target's intra-group phase (its bit tl + 1 = G is above the group
// and set for every element of the orbit when it is set at all):
// exp(i pi c / 2^G) on every element c.
template <typename T, int G, bool CONJ>
__device__ __forceinline__ void apply_p_top(cv<T>* x, const QftRoots& R) {
#pragma unroll
    for (int c = 0; c < (1 << G); ++c) {
        switch (c << (kGroupMax - G)) {
#define QCS_ROOT_CASE(J) \
    case J: x[c] = mul_root<T, J, CONJ>(x[c], R); break;
            QCS_ROOT_CASE(0) QCS_ROOT_CASE(1) QCS_ROOT_CASE(2) QCS_ROOT_CASE(3) QCS_ROOT_CASE(4) QCS_ROOT_CASE(5)
            QCS_ROOT_CASE(6) QCS_ROOT_CASE(7) QCS_ROOT_CASE(8) QCS_ROOT_CASE(9) QCS_ROOT_CASE(10) QCS_ROOT_CASE(11)
            QCS_ROOT_CASE(12) QCS_ROOT_CASE(13) QCS_ROOT_CASE(14) QCS_ROOT_CASE(15) QCS_ROOT_CASE(16)
            QCS_ROOT_CASE(17) QCS_ROOT_CASE(18) QCS_ROOT_CASE(19) QCS_ROOT_CASE(20) QCS_ROOT_CASE(21)
            QCS_ROOT_CASE(22) QCS_ROOT_CASE(23) QCS_ROOT_CASE(24) QCS_ROOT_CASE(25) QCS_ROOT_CASE(26)
            QCS_ROOT_CASE(27) QCS_ROOT_CASE(28) QCS_ROOT_CASE(29) QCS_ROOT_CASE(30) QCS_ROOT_CASE(31)
            QCS_ROOT_CASE(32) QCS_ROOT_CASE(33) QCS_ROOT_CASE(34) QCS_ROOT_CASE(35) QCS_ROOT_CASE(36)
            QCS_ROOT_CASE(37) QCS_ROOT_CASE(38) QCS_ROOT_CASE(39) QCS_ROOT_CASE(40) QCS_ROOT_CASE(41)
            QCS_ROOT_CASE(42) QCS_ROOT_CASE(43) QCS_ROOT_CASE(44) QCS_ROOT_CASE(45) QCS_ROOT_CASE(46)
            QCS_ROOT_CASE(47) QCS_ROOT_CASE(48) QCS_ROOT_CASE(49) QCS_ROOT_CASE(50) QCS_ROOT_CASE(51)
            QCS_ROOT_CASE(52) QCS_ROOT_CASE(53) QCS_ROOT_CASE(54) QCS_ROOT_CASE(55) QCS_ROOT_CASE(56)
            QCS_ROOT_CASE(57) QCS_ROOT_CASE(58) QCS_ROOT_CASE(59) QCS_ROOT_CASE(60) QCS_ROOT_CASE(61)
            QCS_ROOT_CASE(62) QCS_ROOT_CASE(63)
#undef QCS_ROOT_CASE
            default: break;
        }
    }
}

template <typename T>
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One round of k_qft_lds: targets a.lo .. a.lo+G-1, each thread taking
// orbits of 2^G chunk elements into registers.
//
// The phase of D_(t+1) on an element splits into the part from the group
// bits lo..t (apply_p) and F_t = exp(i pi Lf / 2^(t+1-sq)) from the bits
// sq..lo-1 below the group, which the round does not change. F_t multiplies
// the elements whose bit t+1 is set, and every later operation of the round
// (forward: Hadamards and phases on bits <= t) keeps that set of elements
// and commutes with a factor that depends only on bits outside lo..t; so
// all F_t of the round can be applied at its end as one factor per element:
// F_t = F^(2^(top-t)) with F = F_top, hence the product is F^m with
// m = sum over the round's phased targets t with bit t+1 set of 2^(top-t).
// In the inverse the same argument moves the conjugate factors to the
// start of the round. Per orbit a table Q[j] = F^(2j + b) (b = the element-
// independent bit above the group) is built with a few complex products;
// each element then takes one complex product instead of one per target.
// The round on one orbit in registers: x[c] is the element base + c 2^lo.
template <typename T, int G, bool INV>
__device__ __forceinline__ void round_orbit(cv<T>* x, uint64_t base, const QftArgs& a,
                                            const double2* __restrict__ tw, const QftRoots& R) {
    constexpr int D = 1 << G;
    const uint32_t top = a.lo + G - 1;
    const bool has_top = top < a.eq;  // the top target has a phase (all lower ones do)
    const uint32_t nlow = a.lo - a.sq;
    const uint32_t ktop = top + 1 - a.sq;
    {
        const bool bt = has_top && ((base >> (top + 1)) & 1);
        const uint64_t Lf = nlow ? ((base >> a.sq) & ((1ull << nlow) - 1)) : 0;
        // Q[j] = F^(2j + bt), j < D / 2
        double2 Q[D / 2];
        if (Lf) {
            const double2 F = phase_pi(tw, Lf, ktop);
            const double2 F2 = cmul2<T>(F, F);
            Q[0] = bt ? F : make_double2(1.0, 0.0);
#pragma unroll
            for (int j = 1; j < D / 2; ++j) Q[j] = cmul2<T>(Q[j - 1], F2);
        }
        auto apply_f = [&]() {
            if (!Lf) return;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                int mc = 0;  // the even part of m: bits 1..G-1 of c, weights 2^(G-1-tl)
#pragma unroll
                for (int tl = 0; tl + 1 < G; ++tl) mc += ((c >> (tl + 1)) & 1) << (G - 1 - tl);
                x[c] = mul_ph<T>(x[c], Q[mc >> 1], INV);
            }
        };
        if (!INV) {
            if (bt) apply_p_top<T, G, false>(x, R);
            apply_h<T, G>(x, G - 1);
            if constexpr (G >= 2) {
                apply_p<T, G, G - 2, false>(x, R);
                apply_h<T, G>(x, G - 2);
            }
            if constexpr (G >= 3) {
                apply_p<T, G, G - 3, false>(x, R);
                apply_h<T, G>(x, G - 3);
            }
            apply_f();
        } else {
            apply_f();
            if constexpr (G >= 3) {
                apply_h<T, G>(x, G - 3);
                apply_p<T, G, G - 3, true>(x, R);
            }
            if constexpr (G >= 2) {
                apply_h<T, G>(x, G - 2);
                apply_p<T, G, G - 2, true>(x, R);
            }
            apply_h<T, G>(x, G - 1);
            if (bt) apply_p_top<T, G, true>(x, R);
        }
    }
}

template <typename T, int G, bool INV>
__device__ void lds_round(cv<T>* s, uint64_t cbase, uint32_t clog, const QftArgs& a, const double2* __restrict__ tw,
                          const QftRoots& R) {
    constexpr int D = 1 << G;
    const uint32_t norb = 1u << (clog - G);
    for (uint32_t o = threadIdx.x; o < norb; o += kLdsThreads) {
        const uint32_t lb = uint32_t(insert_run(o, a.lo, G));
        cv<T> x[D];
#pragma unroll
        for (int c = 0; c < D; ++c) x[c] = s[swz(lb + (uint32_t(c) << a.lo))];
        round_orbit<T, G, INV>(x, cbase | lb, a, tw, R);
#pragma unroll
        for (int c = 0; c < D; ++c) s[swz(lb + (uint32_t(c) << a.lo))] = x[c];
    }
}

template <typename T, bool INV>
__device__ void lds_round_g(cv<T>* s, uint64_t cbase, uint32_t clog, const QftArgs& a, const double2* tw,
                            const QftRoots& R) {
    static_assert(kLdsRound <= 3, "rounds of at most 3 targets");
    switch (a.g) {
        case 1: lds_round<T, 1, INV>(s, cbase, clog, a, tw, R); break;
        case 2: lds_round<T, 2, INV>(s, cbase, clog, a, tw, R); break;
        default: lds_round<T, 3, INV>(s, cbase, clog, a, tw, R); break;
    }
}

// Targets sq..hi (hi < clog <= kLdsLog) inside contiguous chunks of 2^clog
// amplitudes staged in LDS, in rounds of up to kLdsRound consecutive targets
// (from the top down for the QFT, from sq up for its inverse); one LDS round
// trip per round instead of one per target. One chunk per block (one-shot
// grid, in launches of at most kMaxBlocks blocks from chunk boff).
template <typename T, bool INV>
__global__ __launch_bounds__(kLdsThreads) void k_qft_lds(cv<T>* __restrict__ psi, uint32_t clog, uint64_t boff, QftArgs a,
                                                 uint32_t hi, const double2* __restrict__ tw, const QftRoots R) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cv<T>* s = reinterpret_cast<cv<T>*>(smem_raw);
    constexpr int EPT = (1 << kLdsLog) / kLdsThreads;  // chunk elements per thread
    const uint32_t C = 1u << clog;
    const uint64_t cbase = (boff + blockIdx.x) << clog;
    // Forward, full chunk, the top round on the chunk's top 3 bits: a
    // thread's 8 loaded elements tid + 512 k are one orbit of those bits, so
    // that round runs in registers before the tile goes to LDS (one LDS round
    // trip fewer).
    static_assert(EPT == 8 && kLdsRound == 3, "the loaded elements are one 3-target orbit");
    const bool reg_top = !INV && clog == uint32_t(kLdsLog) && hi == clog - 1 && hi >= a.sq + 2;
    {
        cv<T> v[EPT];
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const uint32_t i = threadIdx.x + uint32_t(kLdsThreads) * k;
            if (i < C) v[k] = ldamp<T, kStagedNtLoads>(psi + cbase + i);
        }
        if (reg_top) {
            QftArgs r = a;
            r.lo = hi - 2;
            r.g = 3;
            round_orbit<T, 3, false>(v, cbase | threadIdx.x, r, tw, R);
        }
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const uint32_t i = threadIdx.x + uint32_t(kLdsThreads) * k;
            if (i < C) s[swz(i)] = v[k];
        }
        __syncthreads();
    }
    if (!INV) {
        for (int top = int(hi) - (reg_top ? kLdsRound : 0); top >= int(a.sq); top -= kLdsRound) {
            QftArgs r = a;
            r.lo = uint32_t(top - (kLdsRound - 1) > int(a.sq) ? top - (kLdsRound - 1) : int(a.sq));
            r.g = uint32_t(top) - r.lo + 1;
            lds_round_g<T, false>(s, cbase, clog, r, tw, R);
            __syncthreads();
        }
    } else {
        for (uint32_t lo = a.sq; lo <= hi; lo += kLdsRound) {
            QftArgs r = a;
            r.lo = lo;
            r.g = hi - lo + 1 < uint32_t(kLdsRound) ? hi - lo + 1 : uint32_t(kLdsRound);
            lds_round_g<T, true>(s, cbase, clog, r, tw, R);
            __syncthreads();
        }
    }
    const T hs = hscale<T>(hi - a.sq + 1);
    for (uint32_t i = threadIdx.x; i < C; i += kLdsThreads) stamp<T, true>(psi + cbase + i, scaled<T>(s[swz(i)], hs));
}

// ---- bit reversal of the range [sq, eq] ---------------------------------------
__device__ __forceinline__ uint64_t rev_bits(uint64_t v, uint32_t nb) {
    return nb ? (__brevll(v) >> (64 - nb)) : 0;
}

// Pair swaps; used when sq >= 3 so that 8+ consecutive amplitudes move
// together (whole 128 B lines for complex128).
template <typename T>
__global__ __launch_bounds__(256) void k_bitrev_pairs(cv<T>* __restrict__ psi, uint64_t A, uint32_t sq, uint32_t m) {
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    const uint64_t rmask = ((1ull << m) - 1) << sq;
    for (uint64_t s = uint64_t(blockIdx.x) * 256 + threadIdx.x; s < A; s += stride) {
        const uint64_t r = (s & rmask) >> sq;
        const uint64_t rr = rev_bits(r, m);
        if (rr <= r) continue;
        const uint64_t p = (s & ~rmask) | (rr << sq);
        const cv<T> x = psi[s], y = psi[p];
        psi[s] = y;
        psi[p] = x;
    }
}

// Tiled: Lo = bits [sq, sq+b), Hi = bits (eq-b, eq], Mid = the rest of the
// range. A tile fixes the bits outside [0, sq+b) U Hi; its rows are the
// 2^(sq+b) contiguous amplitudes of bits [0, sq+b). Tile (Mid) maps onto tile
// (rev Mid) with (hi, lo) -> (rev lo, rev hi); one block owns both tiles of a
// pair, so reads complete before writes.
struct BitrevArgs {
    uint32_t sq, m, b;     // range start, range length, tile bits
    uint32_t rowlog;       // sq + b
    uint32_t midlo, midn;  // Mid bits: [sq+b, sq+b+midn)
    uint32_t hh;           // midn / 2
};

// Tile pairs per value of the bits above eq: with Mid = [A | c | B] (A, B of
// hh bits, c the middle bit when midn is odd), rev(Mid) = [rev B | c | rev A],
// so each unordered pair is exactly one (A, c, B) with A <= x = rev(B): for
// each c, sum over x of (x + 1) = 2^hh (2^hh + 1) / 2 pairs.
__host__ __device__ __forceinline__ uint64_t bitrev_pairs_per_outer(const BitrevArgs& a) {
    return (((uint64_t(1) << a.hh) * ((uint64_t(1) << a.hh) + 1)) / 2) << (a.midn & 1);
}

// One block per tile pair (one-shot grid, in launches of at most kMaxBlocks
// blocks from block boff).
template <typename T>
__global__ __launch_bounds__(256) void k_bitrev_tile(cv<T>* __restrict__ psi, uint64_t boff, BitrevArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cv<T>* s = reinterpret_cast<cv<T>*>(smem_raw);
    constexpr int EPT = 4;  // tile elements per thread (tiles of <= 1024)
    // block -> (outer, c, x, A); T(x) = x (x + 1) / 2 <= rem < T(x + 1)
    const uint64_t blk = boff + blockIdx.x;
    const uint32_t cb = a.midn & 1;
    const uint64_t per_c = bitrev_pairs_per_outer(a) >> cb;
    const uint64_t outer = blk / (per_c << cb);
    uint64_t rem = blk - outer * (per_c << cb);
    const uint64_t c = cb ? rem / per_c : 0;
    rem -= c * per_c;
    uint64_t x = uint64_t((sqrt(8.0 * double(rem) + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > rem) --x;
    while ((x + 1) * (x + 2) / 2 <= rem) ++x;
    const uint64_t A = rem - x * (x + 1) / 2;
    const uint64_t mid = (A << (a.hh + cb)) | (c << a.hh) | rev_bits(x, a.hh);
    const uint64_t mid_r = rev_bits(mid, a.midn);
    const bool two = mid_r != mid;
    const uint32_t hi_lo = a.sq + a.m - a.b;  // first Hi bit
    const uint64_t base0 = (outer << (a.sq + a.m)) | (mid << a.rowlog);
    const uint64_t base1 = (outer << (a.sq + a.m)) | (mid_r << a.rowlog);
    // LDS rows are padded by one element so that the transposed reads of
    // consecutive lanes (stride of one row) fall in different banks.
    const uint32_t R = 1u << a.rowlog, H = 1u << a.b, E = R * H, RP = R + 1, EP = RP * H;
    const uint32_t lowmask = (1u << a.sq) - 1;
    cv<T> p0[EPT], p1[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const uint32_t e = threadIdx.x + 256u * k;
        if (e < E) {
            const uint64_t off = (uint64_t(e >> a.rowlog) << hi_lo) | (e & (R - 1));
            p0[k] = ldamp<T, kStagedNtLoads>(psi + (base0 | off));
            if (two) p1[k] = ldamp<T, kStagedNtLoads>(psi + (base1 | off));
        }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const uint32_t e = threadIdx.x + 256u * k;
        if (e < E) {
            const uint32_t r = e & (R - 1), h = e >> a.rowlog;
            s[h * RP + r] = p0[k];
            if (two) s[EP + h * RP + r] = p1[k];
        }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < E; e += 256) {
        // destination (h', r') reads source (h, r) with lo' = rev(h), h' = rev(lo)
        const uint32_t rd = e & (R - 1), hd = e >> a.rowlog;
        const uint32_t lo_d = rd >> a.sq, out_low = rd & lowmask;
        const uint32_t lo_s = uint32_t(rev_bits(hd, a.b)), h_s = uint32_t(rev_bits(lo_d, a.b));
        const uint32_t src = h_s * RP + ((lo_s << a.sq) | out_low);
        // tile0's destination data comes from tile1 (or itself) and vice versa
        stamp<T, true>(psi + (base0 | (uint64_t(hd) << hi_lo) | rd), two ? s[EP + src] : s[src]);
        if (two) stamp<T, true>(psi + (base1 | (uint64_t(hd) << hi_lo) | rd), s[src]);
    }
}

// Whole range + low bits in one contiguous chunk (small ranges).
template <typename T>
__global__ __launch_bounds__(256) void k_bitrev_chunk(cv<T>* __restrict__ psi, uint32_t sq, uint32_t m) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cv<T>* s = reinterpret_cast<cv<T>*>(smem_raw);
    const uint32_t clog = sq + m, C = 1u << clog;
    const uint64_t cbase = uint64_t(blockIdx.x) << clog;
    for (uint32_t i = threadIdx.x; i < C; i += 256) s[i] = psi[cbase + i];
    __syncthreads();
    const uint32_t lowmask = (1u << sq) - 1;
    for (uint32_t i = threadIdx.x; i < C; i += 256) {
        const uint32_t src = (uint32_t(rev_bits(i >> sq, m)) << sq) | (i & lowmask);
        psi[cbase + i] = s[src];
    }
}

// ---- host side -------------------------------------------------------------------

template <typename T>
static int launch_group(qcs_register* r, const double2* tw, uint32_t sq, uint32_t eq, uint32_t lo, uint32_t g,
                        bool inv) {
    const uint64_t norb = r->A >> g;
    const unsigned grid = unsigned((norb + 255) / 256);
    const QftArgs a{sq, eq, lo, g};
    static const QftRoots roots = group_roots();
    auto* psi = static_cast<cv<T>*>(r->psi);
#define QCS_G(GG)                                                                                             \
    do {                                                                                                      \
        if (inv)                                                                                              \
            hipLaunchKernelGGL((k_qft_group<T, GG, true>), dim3(grid), dim3(256), 0, r->stream, psi, norb, a, \
                               tw, roots);                                                                    \
        else                                                                                                  \
            hipLaunchKernelGGL((k_qft_group<T, GG, false>), dim3(grid), dim3(256), 0, r->stream, psi, norb, a, \
                               tw, roots);                                                                    \
    } while (0)
    switch (g) {
        case 1: QCS_G(1); break;
        case 2: QCS_G(2); break;
        case 3: QCS_G(3); break;
        case 4: QCS_G(4); break;
        default: {
            // 5 or 6 targets: orbits split over lane pairs
            const unsigned gridp = unsigned((2 * norb + 255) / 256);
#define QCS_GP(GG, II)                                                                                           \
    hipLaunchKernelGGL((k_qft_gpair<T, GG, II>), dim3(gridp), dim3(256), 0, r->stream, psi, norb, a, tw, roots)
            if (g == 5) {
                if (inv) QCS_GP(5, true);
                else QCS_GP(5, false);
            } else {
                if (inv) QCS_GP(6, true);
                else QCS_GP(6, false);
            }
#undef QCS_GP
            break;
        }
    }
#undef QCS_G
    return check_launch(r, "k_qft_group");
}

template <typename T>
static int launch_lds(qcs_register* r, const double2* tw, uint32_t sq, uint32_t eq, uint32_t hi, uint32_t clog,
                      bool inv) {
    const QftArgs a{sq, eq, sq, hi - sq + 1};
    const uint64_t nchunks = r->A >> clog;
    const size_t shm = sizeof(cv<T>) << clog;
    auto* psi = static_cast<cv<T>*>(r->psi);
    static const QftRoots roots = group_roots();
    for (uint64_t boff = 0; boff < nchunks; boff += kMaxBlocks) {
        const unsigned g = unsigned(std::min<uint64_t>(kMaxBlocks, nchunks - boff));
        if (inv)
            hipLaunchKernelGGL((k_qft_lds<T, true>), dim3(g), dim3(kLdsThreads), shm, r->stream, psi, clog, boff, a, hi, tw,
                               roots);
        else
            hipLaunchKernelGGL((k_qft_lds<T, false>), dim3(g), dim3(kLdsThreads), shm, r->stream, psi, clog, boff, a, hi, tw,
                               roots);
    }
    return check_launch(r, "k_qft_lds");
}

template <typename T>
static int bitrev(qcs_register* r, uint32_t sq, uint32_t eq) {
    const uint32_t m = eq - sq + 1;
    if (m < 2) return QCS_OK;
    auto* psi = static_cast<cv<T>*>(r->psi);
    if (sq + m <= 11) {  // whole range in one LDS chunk (<= 2048 amplitudes)
        const uint32_t clog = sq + m;
        hipLaunchKernelGGL((k_bitrev_chunk<T>), dim3(unsigned(r->A >> clog)), dim3(256), sizeof(cv<T>) << clog,
                           r->stream, psi, sq, m);
        return check_launch(r, "k_bitrev_chunk");
    }
    if (sq >= 3) {
        hipLaunchKernelGGL((k_bitrev_pairs<T>), dim3(16384), dim3(256), 0, r->stream, psi, r->A, sq, m);
        return check_launch(r, "k_bitrev_pairs");
    }
    BitrevArgs a;
    a.sq = sq;
    a.m = m;
    a.b = 5 - sq;  // rows of 32 contiguous amplitudes
    if (2 * a.b > m) a.b = m / 2;
    a.rowlog = sq + a.b;
    a.midlo = a.rowlog;
    a.midn = m - 2 * a.b;
    a.hh = a.midn / 2;
    const uint64_t nblocks = (r->A >> (sq + m)) * bitrev_pairs_per_outer(a);
    const size_t EP = ((size_t(1) << a.rowlog) + 1) << a.b;  // padded tile
    for (uint64_t boff = 0; boff < nblocks; boff += kMaxBlocks) {
        const unsigned g = unsigned(std::min<uint64_t>(kMaxBlocks, nblocks - boff));
        hipLaunchKernelGGL((k_bitrev_tile<T>), dim3(g), dim3(256), 2 * EP * sizeof(cv<T>), r->stream, psi, boff, a);
    }
    return check_launch(r, "k_bitrev_tile");
}

template <typename T>
static int qft_fused(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    const double2* tw = twiddles(r->device);
    if (!tw) return set_error(QCS_ENOMEM, "twiddle table allocation failed");
    // LDS pass covers targets [sq, lds_hi] when sq lies inside a 4096-chunk
    const uint32_t clog = r->n < uint32_t(kLdsLog) ? r->n : uint32_t(kLdsLog);
    const bool use_lds = sq < clog && (eq - sq + 1) >= 2;
    const uint32_t lds_hi = use_lds ? (eq < clog - 1 ? eq : clog - 1) : sq - 1;
    // register groups over (lds_hi, eq], in blocks of <= gmax from the top
    // (QCS_QFT_GROUP overrides the widest group, 1..6)
    static const int gmax = [] {
        const char* v = std::getenv("QCS_QFT_GROUP");
        const int g = v ? std::atoi(v) : kGroupMax;
        return g >= 1 && g <= kGroupMax ? g : kGroupMax;
    }();
    std::vector<std::pair<uint32_t, uint32_t>> groups;  // (lo, g)
    const uint32_t first = use_lds ? lds_hi + 1 : sq;
    for (int64_t hi = eq; hi >= int64_t(first);) {
        const uint32_t g = uint32_t(std::min<int64_t>(gmax, hi - int64_t(first) + 1));
        groups.push_back({uint32_t(hi) - g + 1, g});
        hi -= g;
    }
    int e = QCS_OK;
    if (!inverse) {
        for (const auto& gr : groups)
            if ((e = launch_group<T>(r, tw, sq, eq, gr.first, gr.second, false))) return e;
        if (use_lds && (e = launch_lds<T>(r, tw, sq, eq, lds_hi, clog, false))) return e;
        if (do_swap && (e = bitrev<T>(r, sq, eq))) return e;
    } else {
        if (do_swap && (e = bitrev<T>(r, sq, eq))) return e;
        if (use_lds && (e = launch_lds<T>(r, tw, sq, eq, lds_hi, clog, true))) return e;
        for (auto it = groups.rbegin(); it != groups.rend(); ++it)
            if ((e = launch_group<T>(r, tw, sq, eq, it->first, it->second, true))) return e;
    }
    return QCS_OK;
}

int qft_sequence(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    const double hs = 1.0 / std::sqrt(2.0);
    const hcplx H[4] = {{hs, 0}, {hs, 0}, {hs, 0}, {-hs, 0}};
    auto had = [&](uint32_t q) { return k1q(r, 0, q, nullptr, 0, H); };
    auto cphase = [&](uint32_t cur, uint32_t ctrl, double phase) {
        const std::complex<double> p = std::polar(1.0, phase);
        const uint32_t bits[2] = {cur, ctrl};
        return kphase(r, bits, 2, hcplx{p.real(), p.imag()});
    };
    auto swaps = [&]() {
        for (uint32_t a = sq, b = eq; a < b; ++a, --b) {
            const int e = kswap(r, a, b, nullptr, 0, hcplx{1, 0}, false);
            if (e) return e;
        }
        return 0;
    };
    const int s = int(sq), e = int(eq);
    int err = 0;
    if (!inverse) {
        if ((err = had(eq))) return err;
        for (int cur = e; cur > s; --cur) {
            double phase = M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                if ((err = cphase(cur, ctrl, phase))) return err;
                phase *= 0.5;
            }
            if ((err = had(cur - 1))) return err;
        }
        if (do_swap && (err = swaps())) return err;
    } else {
        if (do_swap && (err = swaps())) return err;
        for (int cur = s + 1; cur <= e; ++cur) {
            if ((err = had(cur - 1))) return err;
            double phase = -M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                if ((err = cphase(cur, ctrl, phase))) return err;
                phase *= 0.5;
            }
        }
        if ((err = had(eq))) return err;
    }
    return QCS_OK;
}

int qft(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    if (r->n > uint32_t(kPhaseBits)) return qft_sequence(r, sq, eq, do_swap, inverse);
    return r->prec == QCS_COMPLEX128 ? qft_fused<double>(r, sq, eq, do_swap, inverse)
                                     : qft_fused<float>(r, sq, eq, do_swap, inverse);
}

}  // namespace qcs
