// qft.hip - quantum Fourier transform on a contiguous qubit range [sq, eq].
//
// The reference builds QFT/IQFT from Hadamards, controlled phase shifts and a
// SWAP network (QuantumFourierTransform.h:35-87, QubitsSwapper.h:23-33): at
// n = 30 that is 480 full passes over the state. Here the same operator is
// applied in about n/5 + 2 passes:
//
//   The controlled phase shifts between target `cur` and every control below
//   it (phase pi/2, pi/4, ... for distance 1, 2, ...) multiply to one
//   diagonal D_cur: amplitudes with bit cur set get exp(i pi L / 2^(cur-sq)),
//   L = bits sq..cur-1 of the index. The forward sequence is then
//   H(eq), D_eq, H(eq-1), D_(eq-1), ..., D_(sq+1), H(sq) and the inverse is
//   H(sq), D*_(sq+1), H(sq+1), ..., D*_eq, H(eq) (after the swaps).
//
//   k_qft_group applies that sequence for up to 5 consecutive targets at once
//   with the 2^5 amplitudes of each orbit in registers; k_qft_lds applies it
//   for all targets below 12 inside 4096-amplitude chunks staged in LDS; the
//   SWAP network is one bit-reversal pass (k_bitrev_tile, LDS-tiled, or
//   k_bitrev_pairs when runs of >= 8 amplitudes stay contiguous).
//
// Phases come from three 4096-entry tables: exp(i pi x / 2^36) for
// x = x2 2^24 + x1 2^12 + x0 is T2[x2] T1[x1] T0[x0], computed on the host in
// long double; the kernels need no transcendental functions.
// qft_sequence() issues the reference's exact gate sequence instead.
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

#include "common.h"

namespace qcs {

constexpr int kTwBits = 12;
constexpr int kTwLen = 1 << kTwBits;
constexpr int kPhaseBits = 36;  // phases are multiples of pi / 2^36
constexpr int kLdsLog = 12;     // k_qft_lds chunk: 4096 amplitudes
constexpr int kGroupMax = 5;

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

template <typename T>
__device__ __forceinline__ void hadamard(cv<T>& a, cv<T>& b) {
    const T s = T(0.70710678118654752440);
    const cv<T> x = a, y = b;
    a.x = (x.x + y.x) * s;
    a.y = (x.y + y.y) * s;
    b.x = (x.x - y.x) * s;
    b.y = (x.y - y.y) * s;
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
    cv<T> x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = ldamp<T, true>(psi + (base | (uint64_t(c) << a.lo)));
    if (!INV) {
#pragma unroll
        for (int tl = G - 1; tl >= 0; --tl) {
            const uint32_t t = a.lo + tl;
            if (t < a.eq) apply_d<T, G>(x, base, t, a, tw, R, false);
            apply_h<T, G>(x, tl);
        }
    } else {
#pragma unroll
        for (int tl = 0; tl < G; ++tl) {
            const uint32_t t = a.lo + tl;
            apply_h<T, G>(x, tl);
            if (t < a.eq) apply_d<T, G>(x, base, t, a, tw, R, true);
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) stamp<T, true>(psi + (base | (uint64_t(c) << a.lo)), x[c]);
}

// Targets sq..hi (hi < clog) inside contiguous chunks of 2^clog amplitudes.
template <typename T, bool INV>
__global__ __launch_bounds__(256) void k_qft_lds(cv<T>* __restrict__ psi, uint32_t clog, QftArgs a, uint32_t hi,
                                                 const double2* __restrict__ tw) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cv<T>* s = reinterpret_cast<cv<T>*>(smem_raw);
    const uint32_t C = 1u << clog;
    const uint64_t cbase = uint64_t(blockIdx.x) << clog;
    for (uint32_t i = threadIdx.x; i < C; i += 256) s[i] = ldamp<T, true>(psi + cbase + i);
    __syncthreads();
    const uint32_t nst = hi - a.sq + 1;
    for (uint32_t step = 0; step < nst; ++step) {
        const uint32_t t = INV ? a.sq + step : hi - step;
        const uint32_t tb = 1u << t;
        const bool has_d = t < a.eq;
        const uint32_t k = t + 1 - a.sq;
        for (uint32_t p = threadIdx.x; p < C / 2; p += 256) {
            const uint32_t i0 = ((p >> t) << (t + 1)) | (p & (tb - 1));  // bit t = 0
            const uint32_t i1 = i0 | tb;
            cv<T> u = s[i0], v = s[i1];
            // bit t+1 of the full index (inside the chunk or the chunk base)
            const uint64_t full = cbase | i0;
            const bool dset = has_d && ((full >> (t + 1)) & 1);
            const uint64_t L0 = (full >> a.sq) & ((1ull << (t - a.sq)) - 1);  // bits sq..t-1
            if (!INV) {
                if (dset) {
                    u = mul_ph<T>(u, phase_pi(tw, L0, k), false);
                    v = mul_ph<T>(v, phase_pi(tw, L0 | (1ull << (t - a.sq)), k), false);
                }
                hadamard<T>(u, v);
            } else {
                hadamard<T>(u, v);
                if (dset) {
                    u = mul_ph<T>(u, phase_pi(tw, L0, k), true);
                    v = mul_ph<T>(v, phase_pi(tw, L0 | (1ull << (t - a.sq)), k), true);
                }
            }
            s[i0] = u;
            s[i1] = v;
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < C; i += 256) stamp<T, true>(psi + cbase + i, s[i]);
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
};

template <typename T>
__global__ __launch_bounds__(256) void k_bitrev_tile(cv<T>* __restrict__ psi, uint64_t ntiles, BitrevArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    cv<T>* s = reinterpret_cast<cv<T>*>(smem_raw);
    const uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    // tile id -> (outside-high bits, Mid); the tile index enumerates all bits
    // above rowlog except the Hi bits.
    const uint32_t hi_lo = a.sq + a.m - a.b;      // first Hi bit
    const uint64_t below_hi = tile & ((1ull << (hi_lo - a.rowlog)) - 1);  // Mid bits
    const uint64_t above = tile >> (hi_lo - a.rowlog);                     // bits above eq
    const uint64_t mid = below_hi;
    const uint64_t mid_r = rev_bits(mid, a.midn);
    if (mid_r < mid) return;  // the partner block handles this pair
    const uint32_t R = 1u << a.rowlog, H = 1u << a.b, E = R * H;
    auto tile_base = [&](uint64_t md) {
        return (above << (a.sq + a.m)) | (md << a.rowlog);
    };
    const uint64_t base0 = tile_base(mid), base1 = tile_base(mid_r);
    const bool two = mid_r != mid;
    // element e = (h, r): r = e % R (bits [0, rowlog)), h = e / R (Hi bits)
    for (uint32_t e = threadIdx.x; e < E; e += 256) {
        const uint32_t r = e & (R - 1), h = e >> a.rowlog;
        s[e] = psi[base0 | (uint64_t(h) << hi_lo) | r];
        if (two) s[E + e] = psi[base1 | (uint64_t(h) << hi_lo) | r];
    }
    __syncthreads();
    const uint32_t lowmask = (1u << a.sq) - 1;
    for (uint32_t e = threadIdx.x; e < E; e += 256) {
        // destination (h', r') reads source (h, r) with lo' = rev(h), h' = rev(lo)
        const uint32_t rd = e & (R - 1), hd = e >> a.rowlog;
        const uint32_t lo_d = rd >> a.sq, out_low = rd & lowmask;
        const uint32_t lo_s = uint32_t(rev_bits(hd, a.b)), h_s = uint32_t(rev_bits(lo_d, a.b));
        const uint32_t src = (h_s << a.rowlog) | (lo_s << a.sq) | out_low;
        // tile0's destination data comes from tile1 (or itself) and vice versa
        psi[base0 | (uint64_t(hd) << hi_lo) | rd] = two ? s[E + src] : s[src];
        if (two) psi[base1 | (uint64_t(hd) << hi_lo) | rd] = s[src];
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
        default: QCS_G(5); break;
    }
#undef QCS_G
    return check_launch(r, "k_qft_group");
}

template <typename T>
static int launch_lds(qcs_register* r, const double2* tw, uint32_t sq, uint32_t eq, uint32_t hi, uint32_t clog,
                      bool inv) {
    const QftArgs a{sq, eq, sq, hi - sq + 1};
    const unsigned grid = unsigned(r->A >> clog);
    const size_t shm = sizeof(cv<T>) << clog;
    auto* psi = static_cast<cv<T>*>(r->psi);
    if (inv)
        hipLaunchKernelGGL((k_qft_lds<T, true>), dim3(grid), dim3(256), shm, r->stream, psi, clog, a, hi, tw);
    else
        hipLaunchKernelGGL((k_qft_lds<T, false>), dim3(grid), dim3(256), shm, r->stream, psi, clog, a, hi, tw);
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
    const uint64_t ntiles = r->A >> (a.rowlog + a.b);
    const size_t E = size_t(1) << (a.rowlog + a.b);
    hipLaunchKernelGGL((k_bitrev_tile<T>), dim3(unsigned(ntiles)), dim3(256), 2 * E * sizeof(cv<T>), r->stream, psi,
                       ntiles, a);
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
    // register groups over (lds_hi, eq], in blocks of <= 5 from the top
    std::vector<std::pair<uint32_t, uint32_t>> groups;  // (lo, g)
    const uint32_t first = use_lds ? lds_hi + 1 : sq;
    for (int64_t hi = eq; hi >= int64_t(first);) {
        const uint32_t g = uint32_t(std::min<int64_t>(kGroupMax, hi - int64_t(first) + 1));
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
