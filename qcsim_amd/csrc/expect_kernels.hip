// expect_kernels.hip - Pauli-string expectations, and host vectors streamed
// through the device: inner products with a host vector, diagonal and dense
// operators on the whole register, complex64 conversions.
//
// Pauli strings. A string P = i^{|x & z|} X^x Z^z (per qubit: X if only its
// x bit is set, Z if only its z bit, Y = iXZ if both, identity if neither)
// has (P psi)_s = i^{ny} (-1)^{popcount((s^x) & z)} psi_{s^x}, ny = |x & z|.
// The reference measures such a string after a basis change and averages
// 10000 sampled parities (VQE.h:116-142, with the change of basis at
// :33-44); it leaves this exact form as a TODO (QubitRegister.h:649). Here
// the terms that share one x mask are evaluated in one read-only pass:
//   x = 0:  <P> = sum_s |psi_s|^2 sigma(s), sigma(s) = (-1)^{popcount(s & z)}
//   x != 0: the pair (s, t = s ^ x) with bit h = msb(x) of s clear adds
//           sigma(s) (kappa a + conj a), a = conj(psi_s) psi_t and
//           kappa = (-1)^{ny}, since sigma(t) = kappa sigma(s); so each term
//           sums sigma(s) Re a (ny even) or sigma(s) Im a (ny odd), and
//           <P> = 2 (-1)^{ny/2} S (even) or 2 (-1)^{(ny-1)/2} S (odd).
// Each amplitude is read once per x mask; all sums are FP64 with a fixed
// grid and a fixed combination order.
#include <algorithm>
#include <map>
#include <vector>

#include "common.h"

namespace qcs {

constexpr int kExThreads = 256;

constexpr uint32_t kFlipBits = 16;  // iteration bits: 2^35 amplitudes over a grid of 2^19 work-items

// One pass's terms, passed by value (kernel arguments, read by scalar loads).
// Iteration k of a work-item visits s = s_thread | s_iter(g(k)), g the Gray
// code, so step k (k >= 1) flips exactly iteration bit b = ctz(k): state bit
// bit[b], and the sign of term j flips by w[b][j] (1 << 31 when z_j holds
// that bit: the sign bit of the double +-1.0).
struct PauliFlip {
    uint32_t w[kFlipBits][kPauliTerms];
    uint64_t bit[kFlipBits];
    uint64_t z[kPauliTerms];
};

// One pass for the terms j < nt that share x (PAIR: x != 0, whose highest
// bit is hbit). The grid stride is a power of two, so an index splits into
// disjoint bit sets: s = s_thread | s_iter, s_thread from the work-item's
// index and s_iter from the iteration alone (insert_run maps disjoint bit
// sets to disjoint bit sets). Hence sigma_j(s) = sigma_j(s_thread)
// sigma_j(s_iter): the second factor is the same for the whole wave, kept as
// the high word of a double +-1.0 on the scalar unit and updated with one XOR
// per term and iteration along the Gray walk; the vector unit adds v (+-) per
// term with FP64 FMAs (x = 0: |psi_s|^2; x != 0: conj(psi_s) psi_{s^x}, real
// and imaginary parts, of which the host keeps the one its term needs); the
// first factor is applied once per work-item at the end.
// Product term of one iteration: |psi_s|^2 (x = 0), or conj(psi_s) psi_{s^x}
// as (re, im).
template <typename T, bool PAIR>
__device__ __forceinline__ void pauli_product(const cv<T>* __restrict__ psi, uint64_t s, uint64_t x, double& re,
                                              double& im) {
    if constexpr (PAIR) {
        const cv<T> u = ldamp<T, true>(psi + s);
        const cv<T> w = ldamp<T, true>(psi + (s ^ x));
        re = double(u.x) * double(w.x) + double(u.y) * double(w.y);
        im = double(u.x) * double(w.y) - double(u.y) * double(w.x);
    } else {
        const cv<T> v = ldamp<T, true>(psi + s);
        re = double(v.x) * double(v.x) + double(v.y) * double(v.y);
        im = 0;
    }
}

template <typename T, int G, bool PAIR>
__global__ __launch_bounds__(kExThreads) void k_pauli(const cv<T>* __restrict__ psi, uint64_t npairs, uint64_t x,
                                                      uint32_t hbit, uint32_t nt, uint32_t niter, uint32_t odd_mask,
                                                      PauliFlip F, double* __restrict__ partial) {
    // iterations per loop trip, their loads issued together: 4 single loads
    // (x = 0) or 2 pairs
    constexpr uint32_t U = PAIR ? 2 : 4;
    __shared__ double lds[4][G];
    double ar[G], ai[PAIR ? G : 1];
#pragma unroll
    for (int j = 0; j < G; ++j) {
        ar[j] = 0;
        if constexpr (PAIR) ai[j] = 0;
    }
    uint64_t sg[G];  // the per-iteration signs (+-1.0) as bit patterns
#pragma unroll
    for (int j = 0; j < G; ++j) sg[j] = 0x3FF0000000000000ull;
    const uint64_t t = uint64_t(blockIdx.x) * kExThreads + threadIdx.x;
    // a work-item beyond a state smaller than one grid reads index 0 and
    // adds nothing
    const bool live = t < npairs;
    const double wt = live ? 1.0 : 0.0;
    const uint64_t st = live ? (PAIR ? insert_run(t, hbit, 1) : t) : 0;
    auto acc = [&](double re, double im, int j) {
        const double sgn = __longlong_as_double(static_cast<long long>(sg[j]));
        ar[j] = fma(re, sgn, ar[j]);
        if constexpr (PAIR) ai[j] = fma(im, sgn, ai[j]);
    };
    uint64_t sk = 0;
    if (niter >= U) {
        for (uint32_t k = 0; k < niter; k += U) {
            if (k) {  // Gray step k flips iteration bit ctz(k) >= log2(U)
                const uint32_t b = __builtin_ctz(k);
                sk ^= F.bit[b];
#pragma unroll
                for (int j = 0; j < G; ++j) sg[j] ^= uint64_t(F.w[b][j]) << 32;
            }
            // steps k+1 .. k+U-1 flip bits 0 / 1 0 1 / 0
            uint64_t s[U];
            s[0] = st | sk;
            s[1] = s[0] ^ F.bit[0];
            if constexpr (U == 4) {
                s[2] = s[1] ^ F.bit[1];
                s[3] = s[2] ^ F.bit[0];
            }
            double re[U], im[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) pauli_product<T, PAIR>(psi, s[u], x, re[u], im[u]);
#pragma unroll
            for (int j = 0; j < G; ++j) {
                acc(re[0], im[0], j);
                sg[j] ^= uint64_t(F.w[0][j]) << 32;
                acc(re[1], im[1], j);
                if constexpr (U == 4) {
                    sg[j] ^= uint64_t(F.w[1][j]) << 32;
                    acc(re[2], im[2], j);
                    sg[j] ^= uint64_t(F.w[0][j]) << 32;
                    acc(re[3], im[3], j);
                }
            }
            // sk and sg of the trip's last step, k + U - 1 (from step k:
            // bit 1 for U = 4, the two flips of bit 0 cancelling; bit 0 for
            // U = 2)
            sk ^= U == 4 ? F.bit[1] : F.bit[0];
        }
    } else {
        for (uint32_t k = 0; k < niter; ++k) {
            if (k) {
                const uint32_t b = __builtin_ctz(k);
                sk ^= F.bit[b];
#pragma unroll
                for (int j = 0; j < G; ++j) sg[j] ^= uint64_t(F.w[b][j]) << 32;
            }
            double re, im;
            pauli_product<T, PAIR>(psi, st | sk, x, re, im);
#pragma unroll
            for (int j = 0; j < G; ++j) acc(re, im, j);
        }
    }
    // per term: the part the host needs, the work-item's sign, the wave's sum
    const uint32_t w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        double v = ar[j];
        if constexpr (PAIR)
            if ((odd_mask >> j) & 1) v = ai[j];
        v *= wt;
        if (__popcll(st & F.z[j]) & 1) v = -v;
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) lds[w][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < nt)
        partial[uint64_t(blockIdx.x) * nt + threadIdx.x] =
            ((lds[0][threadIdx.x] + lds[1][threadIdx.x]) + lds[2][threadIdx.x]) + lds[3][threadIdx.x];
}

// sum over blocks of the per-block partials of each term, fixed order
__global__ __launch_bounds__(kExThreads) void k_pauli_final(const double* __restrict__ partial, uint32_t nb,
                                                            uint32_t nt, double* __restrict__ out) {
    __shared__ double lds[4];
    for (uint32_t j = 0; j < nt; ++j) {
        double acc = 0;
        for (uint32_t b = threadIdx.x; b < nb; b += kExThreads) acc += partial[uint64_t(b) * nt + j];
        const double t = block_sum256(acc, lds);
        if (threadIdx.x == 0) out[j] = t;
    }
}

// <psi[0..len) | v[0..len)> partial sums (conj(psi) v), FP64.
template <typename T>
__global__ __launch_bounds__(kExThreads) void k_inner_vec(const cv<T>* __restrict__ psi,
                                                          const double2* __restrict__ v, uint64_t len,
                                                          double* __restrict__ partial) {
    __shared__ double lds[4];
    double re = 0, im = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kExThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kExThreads + threadIdx.x; i < len; i += stride) {
        const cv<T> a = ldamp<T, true>(psi + i);
        const double2 b = ldamp<double, true>(v + i);
        re += double(a.x) * b.x + double(a.y) * b.y;
        im += double(a.x) * b.y - double(a.y) * b.x;
    }
    const double tr = block_sum256(re, lds);
    const double ti = block_sum256(im, lds);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = tr;
        partial[2 * blockIdx.x + 1] = ti;
    }
}

__global__ __launch_bounds__(kExThreads) void k_inner_final(const double* __restrict__ partial, uint32_t nb,
                                                            double* __restrict__ out) {
    __shared__ double lds[4];
    for (int c = 0; c < 2; ++c) {
        double acc = 0;
        for (uint32_t b = threadIdx.x; b < nb; b += kExThreads) acc += partial[2 * b + c];
        const double t = block_sum256(acc, lds);
        if (threadIdx.x == 0) out[c] = t;
    }
}

// psi[i] *= d[i] for i < len (one amplitude per lane)
template <typename T>
__global__ __launch_bounds__(kExThreads) void k_diag_part(cv<T>* __restrict__ psi, const double2* __restrict__ d,
                                                          uint64_t len) {
    const uint64_t i = uint64_t(blockIdx.x) * kExThreads + threadIdx.x;
    if (i >= len) return;
    const double2 f = ldamp<double, true>(d + i);
    const cv<T> a = ldamp<T, true>(psi + i);
    cv<T> o;
    o.x = T(f.x * double(a.x) - f.y * double(a.y));
    o.y = T(f.x * double(a.y) + f.y * double(a.x));
    stamp<T, true>(psi + i, o);
}

// y[row0 + r] = sum_j M[r][j] x[j] for the rows r < nrows of a row-major
// panel: one wave per row, lanes over the columns (consecutive lanes on
// consecutive matrix entries), FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(kExThreads) void k_gemv_panel(const double2* __restrict__ M, const cv<T>* __restrict__ x,
                                                           cv<T>* __restrict__ y, uint64_t cols, uint64_t row0,
                                                           uint64_t nrows) {
    const uint64_t r = uint64_t(blockIdx.x) * (kExThreads / 64) + (threadIdx.x >> 6);
    if (r >= nrows) return;  // whole waves leave together
    const uint32_t lane = threadIdx.x & 63;
    const double2* row = M + r * cols;
    double re = 0, im = 0;
    for (uint64_t j = lane; j < cols; j += 64) {
        const double2 m = ldamp<double, true>(row + j);
        const cv<T> v = x[j];
        re += m.x * double(v.x) - m.y * double(v.y);
        im += m.x * double(v.y) + m.y * double(v.x);
    }
    re = wave_sum(re);
    im = wave_sum(im);
    if (lane == 0) {
        cv<T> o;
        o.x = T(re);
        o.y = T(im);
        y[row0 + r] = o;
    }
}

__global__ void k_d2f_part(const double2* __restrict__ in, float2* __restrict__ out, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = make_float2(float(in[i].x), float(in[i].y));
}
__global__ void k_f2d_part(const float2* __restrict__ in, double2* __restrict__ out, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = make_double2(double(in[i].x), double(in[i].y));
}

// ---------------------------------------------------------------------------
// Host side.

int ensure_xfer(qcs_register* r, size_t bytes) {
    if (bytes > kXferBytes) return set_error(QCS_EINVAL, "transfer chunk of %zu bytes too large", bytes);
    if (r->xfer && r->xfer_bytes >= bytes) return QCS_OK;
    if (hipStreamSynchronize(r->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
    if (r->xfer) (void)hipFree(r->xfer);
    r->xfer = nullptr;
    r->xfer_bytes = 0;
    if (hipMalloc(&r->xfer, bytes) != hipSuccess) {
        (void)hipGetLastError();
        r->xfer = nullptr;
        return set_error(QCS_ENOMEM, "transfer buffer of %zu bytes", bytes);
    }
    r->xfer_bytes = bytes;
    return QCS_OK;
}

// Amplitudes per transfer chunk (16 B each on the host side).
static uint64_t xfer_amps(const qcs_register* r) {
    const uint64_t cap = kXferBytes / 16;
    return r->A < cap ? r->A : cap;
}

static uint32_t grid256(uint64_t n) { return uint32_t((n + 255) / 256); }

template <typename T>
static int pauli_t(qcs_register* r, uint32_t nterms, const uint64_t* xmask, const uint64_t* zmask, double* out) {
    const uint64_t full = r->A - 1;
    // terms grouped by x mask, in order of first appearance
    std::vector<uint64_t> xs;
    std::map<uint64_t, std::vector<uint32_t>> groups;
    for (uint32_t t = 0; t < nterms; ++t) {
        if ((xmask[t] | zmask[t]) & ~full) return set_error(QCS_EINVAL, "term %u: mask beyond the register", t);
        auto it = groups.find(xmask[t]);
        if (it == groups.end()) {
            xs.push_back(xmask[t]);
            groups[xmask[t]] = {t};
        } else {
            it->second.push_back(t);
        }
    }
    double* res = r->red + kRedResult;
    for (const uint64_t x : xs) {
        const std::vector<uint32_t>& terms = groups[x];
        const uint64_t npairs = x ? r->A / 2 : r->A;
        const uint32_t hbit = x ? uint32_t(63 - __builtin_clzll(x)) : 0;
        for (size_t t0 = 0; t0 < terms.size(); t0 += kPauliTerms) {
            const uint32_t nt = uint32_t(std::min<size_t>(kPauliTerms, terms.size() - t0));
            uint32_t odd = 0;
            // a power-of-two grid (the sign split needs a power-of-two stride)
            uint64_t nb = 1;
            while (nb < uint64_t(kPauliBlocks) && nb * kExThreads * 8 < npairs) nb <<= 1;
            const uint64_t stride = nb * kExThreads;
            const uint32_t lstride = uint32_t(__builtin_ctzll(stride));
            const uint64_t niter = npairs > stride ? npairs / stride : 1;
            const uint32_t ibits = uint32_t(__builtin_ctzll(niter));
            if (ibits > kFlipBits) return set_error(QCS_EINVAL, "state too large for the Pauli pass");
            PauliFlip F{};
            for (uint32_t j = 0; j < nt; ++j) {
                const uint64_t z = zmask[terms[t0 + j]];
                F.z[j] = z;
                if (__builtin_popcountll(x & z) & 1) odd |= 1u << j;
            }
            for (uint32_t b = 0; b < ibits; ++b) {
                const uint32_t kb = lstride + b, pos = (x && kb >= hbit) ? kb + 1 : kb;
                F.bit[b] = 1ull << pos;
                for (uint32_t j = 0; j < nt; ++j) F.w[b][j] = uint32_t((F.z[j] >> pos) & 1) << 31;
            }
            const cv<T>* psi = static_cast<const cv<T>*>(r->psi);
            const dim3 grid{unsigned(nb)}, block{unsigned(kExThreads)};
            const uint32_t ni = uint32_t(niter);
            if (x == 0) {
                if (nt <= 8)
                    hipLaunchKernelGGL((k_pauli<T, 8, false>), grid, block, 0, r->stream, psi, npairs, x, hbit, nt, ni,
                                       odd, F, r->red);
                else
                    hipLaunchKernelGGL((k_pauli<T, kPauliTerms, false>), grid, block, 0, r->stream, psi, npairs, x,
                                       hbit, nt, ni, odd, F, r->red);
            } else {
                if (nt <= 8)
                    hipLaunchKernelGGL((k_pauli<T, 8, true>), grid, block, 0, r->stream, psi, npairs, x, hbit, nt, ni,
                                       odd, F, r->red);
                else
                    hipLaunchKernelGGL((k_pauli<T, kPauliTerms, true>), grid, block, 0, r->stream, psi, npairs, x,
                                       hbit, nt, ni, odd, F, r->red);
            }
            if (const int e = check_launch(r, "k_pauli")) return e;
            hipLaunchKernelGGL(k_pauli_final, dim3(1), dim3(kExThreads), 0, r->stream, r->red, uint32_t(nb), nt, res);
            if (const int e = check_launch(r, "k_pauli_final")) return e;
            double sums[kPauliTerms];
            if (const int e = fetch(r, res, int(nt), sums)) return e;
            for (uint32_t j = 0; j < nt; ++j) {
                const uint32_t t = terms[t0 + j];
                const uint32_t ny = uint32_t(__builtin_popcountll(x & zmask[t]));
                if (x == 0) {
                    out[t] = sums[j];
                } else {
                    const uint32_t k = (ny & 1) ? (ny - 1) / 2 : ny / 2;
                    out[t] = (k & 1 ? -2.0 : 2.0) * sums[j];
                }
            }
        }
    }
    return QCS_OK;
}

int pauli_expectation(qcs_register* r, uint32_t nterms, const uint64_t* xmask, const uint64_t* zmask, double* out) {
    return r->prec == QCS_COMPLEX128 ? pauli_t<double>(r, nterms, xmask, zmask, out)
                                     : pauli_t<float>(r, nterms, xmask, zmask, out);
}

template <typename T>
static int inner_host_t(qcs_register* r, const double* amps, double* re_im) {
    const uint64_t chunk = xfer_amps(r);
    if (const int e = ensure_xfer(r, chunk * 16)) return e;
    double sre = 0, sim = 0;  // chunk sums added in chunk order
    for (uint64_t off = 0; off < r->A; off += chunk) {
        const uint64_t len = std::min(chunk, r->A - off);
        if (hipMemcpyAsync(r->xfer, amps + 2 * off, len * 16, hipMemcpyHostToDevice, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "upload failed");
        const uint32_t nb = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(kRedBlocks, (len + 1023) / 1024)));
        hipLaunchKernelGGL((k_inner_vec<T>), dim3(nb), dim3(kExThreads), 0, r->stream,
                           static_cast<const cv<T>*>(r->psi) + off, static_cast<const double2*>(r->xfer), len, r->red);
        if (const int e = check_launch(r, "k_inner_vec")) return e;
        hipLaunchKernelGGL(k_inner_final, dim3(1), dim3(kExThreads), 0, r->stream, r->red, nb, r->red + kRedResult);
        if (const int e = check_launch(r, "k_inner_final")) return e;
        double z[2];
        if (const int e = fetch(r, r->red + kRedResult, 2, z)) return e;
        sre += z[0];
        sim += z[1];
    }
    re_im[0] = sre;
    re_im[1] = sim;
    return QCS_OK;
}

int inner_host(qcs_register* r, const double* amps, double* re_im) {
    return r->prec == QCS_COMPLEX128 ? inner_host_t<double>(r, amps, re_im) : inner_host_t<float>(r, amps, re_im);
}

template <typename T>
static int diag_host_t(qcs_register* r, const double* diag) {
    const uint64_t chunk = xfer_amps(r);
    if (const int e = ensure_xfer(r, chunk * 16)) return e;
    for (uint64_t off = 0; off < r->A; off += chunk) {
        const uint64_t len = std::min(chunk, r->A - off);
        // pageable source: the copy is ordered after the previous chunk's
        // kernel on this stream, so the buffer is free again
        if (hipMemcpyAsync(r->xfer, diag + 2 * off, len * 16, hipMemcpyHostToDevice, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "diagonal upload failed");
        hipLaunchKernelGGL((k_diag_part<T>), dim3(grid256(len)), dim3(kExThreads), 0, r->stream,
                           static_cast<cv<T>*>(r->psi) + off, static_cast<const double2*>(r->xfer), len);
        if (const int e = check_launch(r, "k_diag_part")) return e;
    }
    // the caller may release `diag` once the call returns
    return hipStreamSynchronize(r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "diagonal sync failed");
}

int apply_diagonal_host(qcs_register* r, const double* diag) {
    return r->prec == QCS_COMPLEX128 ? diag_host_t<double>(r, diag) : diag_host_t<float>(r, diag);
}

// psi <- M psi for a dense row-major 2^n x 2^n matrix in host memory:
// the result goes to the front of the transfer buffer, the matrix streams
// through the rest of it in panels of whole rows, and the result is copied
// over the state at the end.
template <typename T>
static int matrix_host_t(qcs_register* r, const double* mat) {
    const uint64_t A = r->A;
    const size_t ybytes = ((A * sizeof(cv<T>)) + 255) & ~size_t(255);
    const size_t rowbytes = A * 16;
    if (ybytes + rowbytes > kXferBytes) return set_error(QCS_EINVAL, "operator matrix too large for the register");
    const uint64_t rows = std::min<uint64_t>(A, (kXferBytes - ybytes) / rowbytes);
    if (const int e = ensure_xfer(r, ybytes + rows * rowbytes)) return e;
    auto* y = static_cast<cv<T>*>(r->xfer);
    auto* panel = reinterpret_cast<double2*>(static_cast<char*>(r->xfer) + ybytes);
    for (uint64_t r0 = 0; r0 < A; r0 += rows) {
        const uint64_t nr = std::min(rows, A - r0);
        if (hipMemcpyAsync(panel, mat + 2 * r0 * A, nr * rowbytes, hipMemcpyHostToDevice, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "matrix upload failed");
        const uint64_t nblk = (nr + 3) / 4;
        hipLaunchKernelGGL((k_gemv_panel<T>), dim3(unsigned(nblk)), dim3(kExThreads), 0, r->stream, panel,
                           static_cast<const cv<T>*>(r->psi), y, A, r0, nr);
        if (const int e = check_launch(r, "k_gemv_panel")) return e;
    }
    if (hipMemcpyAsync(r->psi, y, A * sizeof(cv<T>), hipMemcpyDeviceToDevice, r->stream) != hipSuccess ||
        hipStreamSynchronize(r->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "operator result copy failed");
    return QCS_OK;
}

int apply_matrix_host(qcs_register* r, const double* mat) {
    return r->prec == QCS_COMPLEX128 ? matrix_host_t<double>(r, mat) : matrix_host_t<float>(r, mat);
}

// complex64 registers: FP64 host amplitudes converted on the device, chunk
// by chunk through the transfer buffer
int convert_upload(qcs_register* r, const double* amps) {
    const uint64_t chunk = xfer_amps(r);
    if (const int e = ensure_xfer(r, chunk * 16)) return e;
    for (uint64_t off = 0; off < r->A; off += chunk) {
        const uint64_t len = std::min(chunk, r->A - off);
        if (hipMemcpyAsync(r->xfer, amps + 2 * off, len * 16, hipMemcpyHostToDevice, r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "upload failed");
        hipLaunchKernelGGL(k_d2f_part, dim3(grid256(len)), dim3(256), 0, r->stream,
                           static_cast<const double2*>(r->xfer), static_cast<float2*>(r->psi) + off, len);
        if (const int e = check_launch(r, "k_d2f_part")) return e;
    }
    return hipStreamSynchronize(r->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "sync failed");
}

int convert_download(qcs_register* r, double* amps) {
    const uint64_t chunk = xfer_amps(r);
    if (const int e = ensure_xfer(r, chunk * 16)) return e;
    for (uint64_t off = 0; off < r->A; off += chunk) {
        const uint64_t len = std::min(chunk, r->A - off);
        hipLaunchKernelGGL(k_f2d_part, dim3(grid256(len)), dim3(256), 0, r->stream,
                           static_cast<const float2*>(r->psi) + off, static_cast<double2*>(r->xfer), len);
        if (const int e = check_launch(r, "k_f2d_part")) return e;
        if (hipMemcpyAsync(amps + 2 * off, r->xfer, len * 16, hipMemcpyDeviceToHost, r->stream) != hipSuccess ||
            hipStreamSynchronize(r->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "download failed");
    }
    return QCS_OK;
}

}  // namespace qcs
