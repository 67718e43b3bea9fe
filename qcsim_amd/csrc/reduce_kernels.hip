// reduce_kernels.hip - FP64 reductions, CDF search and collapse.
//
// All sums are accumulated in FP64 whatever the storage precision, with a
// fixed grid and a fixed combination order, so results are deterministic
// run to run. The CDF search that the reference runs as one serial FP64 scan
// (QubitRegister.h:171-190; QubitRegisterCalculator.h:955-970) is split into
//   1. per-chunk inclusive scans of |psi|^2 (2048 amplitudes per block),
//   2. one-block scan of the chunk totals into a chunk-level CDF,
//   3. per shot: binary search of the chunk CDF, then the chunk's local scan
//      is recomputed with the same code (bit-identical) to find the first
//      index with prob <= cdf (or prob < cdf for the truncation threshold).
// GPU and CPU prefix sums differ only by rounding; the tests report the
// distance of each uniform to the nearest CDF boundary.
#include <cfloat>
#include <vector>

#include <algorithm>

#include "common.h"

namespace qcs {

// Sum of |psi_s|^2 over s = insert_run(o, first, len) | val.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_sum_sq(const cv<T>* __restrict__ psi, uint64_t norb,
                                                        uint32_t first, uint32_t len, uint64_t val,
                                                        double* __restrict__ partial) {
    __shared__ double lds[4];
    double acc = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
#pragma unroll 4
    for (uint64_t o = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; o < norb; o += stride)
        acc += norm2<T>(ldamp<T, true>(psi + (insert_run(o, first, len) | val)));
    const double t = block_sum256(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Large sums in one pass: block b adds the 2048 matching amplitudes
// 2048 b .. 2048 b + 2047 (8 per thread, 256 apart: consecutive lanes on
// consecutive amplitudes), then k_sum_partials adds the block sums with the
// fixed grid and k_final_sum the grid's. Fixed order, so deterministic.
constexpr uint32_t kSumBlock = 2048;

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_sum_sq_blocks(const cv<T>* __restrict__ psi, uint64_t norb,
                                                               uint32_t first, uint32_t len, uint64_t val,
                                                               double* __restrict__ partial, uint64_t b0) {
    __shared__ double lds[4];
    const uint64_t blk = b0 + blockIdx.x;
    const uint64_t o0 = blk * kSumBlock + threadIdx.x;
    double acc = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSumBlock / kRedThreads; ++k) {
        const uint64_t o = o0 + uint64_t(k) * kRedThreads;
        if (o < norb) acc += norm2<T>(ldamp<T, true>(psi + (insert_run(o, first, len) | val)));
    }
    const double t = block_sum256(acc, lds);
    if (threadIdx.x == 0) partial[blk] = t;
}

__global__ __launch_bounds__(kRedThreads) void k_sum_partials(const double* __restrict__ in, uint64_t n,
                                                              double* __restrict__ partial) {
    __shared__ double lds[4];
    double acc = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) acc += in[i];
    const double t = block_sum256(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Fixed-order sum of `n` partials (n <= kRedBlocks) by one block.
__global__ __launch_bounds__(kRedThreads) void k_final_sum(const double* __restrict__ partial, int n,
                                                           int stride_elems, double* __restrict__ out) {
    __shared__ double lds[4];
    for (int comp = 0; comp < stride_elems; ++comp) {
        double acc = 0;
        for (int i = threadIdx.x; i < n; i += kRedThreads) acc += partial[i * stride_elems + comp];
        const double t = block_sum256(acc, lds);
        if (threadIdx.x == 0) out[comp] = t;
    }
}

template <typename Ta, typename Tb>
__global__ __launch_bounds__(kRedThreads) void k_inner(const cv<Ta>* __restrict__ a, const cv<Tb>* __restrict__ b,
                                                       uint64_t A, double* __restrict__ partial) {
    __shared__ double lds[4];
    double re = 0, im = 0, mx = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
#pragma unroll 4
    for (uint64_t s = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; s < A; s += stride) {
        const cv<Ta> x = ldamp<Ta, true>(a + s);
        const cv<Tb> y = ldamp<Tb, true>(b + s);
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        re += xr * yr + xi * yi;  // conj(x) * y
        im += xr * yi - xi * yr;
        const double dr = xr - yr, di = xi - yi;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
    const double tr = block_sum256(re, lds);
    const double ti = block_sum256(im, lds);
    double m = wave_max(mx);
    __shared__ double lm[4];
    if ((threadIdx.x & 63) == 0) lm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[blockIdx.x * 3 + 0] = tr;
        partial[blockIdx.x * 3 + 1] = ti;
        partial[blockIdx.x * 3 + 2] = fmax(fmax(lm[0], lm[1]), fmax(lm[2], lm[3]));
    }
}

__global__ __launch_bounds__(64) void k_final_max(const double* __restrict__ partial, int n, double* __restrict__ out) {
    double m = 0;
    for (int i = threadIdx.x; i < n; i += 64) m = fmax(m, partial[i * 3 + 2]);
    m = wave_max(m);
    if (threadIdx.x == 0) out[2] = m;
}

// ---------------------------------------------------------------------------
// CDF.

struct ChunkGeom {
    uint64_t clen;  // amplitudes per chunk
    uint32_t per;   // amplitudes per thread (clen / 256, at least 1)
};

// Inclusive scan of |psi|^2 over one chunk. Thread t owns elements
// [t*per, (t+1)*per); returns the exclusive prefix of thread t within the
// chunk and the chunk total (to all threads). Identical code runs in the
// chunk pass and in the search pass, so both see bit-identical values.
// The norms are first staged in LDS by coalesced loads (consecutive lanes on
// consecutive amplitudes; a thread's own run of `per` amplitudes would put
// its lanes `per` amplitudes apart), one pad slot per 8 so that the
// per-thread runs fall in different banks; each thread then adds its run
// in index order, as before.
constexpr uint32_t kChunkMax = 1u << kChunkLog;
constexpr uint32_t kNrmLds = kChunkMax + kChunkMax / 8;

__device__ __forceinline__ uint32_t nrm_slot(uint64_t e) { return uint32_t(e + (e >> 3)); }

template <typename T>
__device__ __forceinline__ double chunk_thread_prefix(const cv<T>* __restrict__ psi, uint64_t cbase,
                                                      ChunkGeom g, double* lds, double* nrm, double* total) {
    for (uint64_t i = threadIdx.x; i < g.clen; i += kScanThreads) nrm[nrm_slot(i)] = norm2<T>(psi[cbase + i]);
    __syncthreads();
    double mine = 0;
    const uint64_t beg = uint64_t(threadIdx.x) * g.per;
    for (uint32_t i = 0; i < g.per; ++i) {
        const uint64_t e = beg + i;
        if (e < g.clen) mine += nrm[nrm_slot(e)];
    }
    // inclusive wave scan
    double v = mine;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    double excl = __shfl_up(v, 1, 64);
    if (lane == 0) excl = 0;
    const int w = threadIdx.x >> 6;
    if (lane == 63) lds[w] = v;
    __syncthreads();
    double wpre = 0;
    for (int k = 0; k < w; ++k) wpre += lds[k];
    *total = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
    return wpre + excl;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_chunk_sums(const cv<T>* __restrict__ psi, ChunkGeom g,
                                                             double* __restrict__ chunk, uint64_t c0) {
    __shared__ double lds[4];
    __shared__ double nrm[kNrmLds];
    double total;
    const uint64_t c = c0 + blockIdx.x;
    chunk_thread_prefix<T>(psi, c * g.clen, g, lds, nrm, &total);
    if (threadIdx.x == 0) chunk[c] = total;
}

// Chunk totals are scanned in groups of kScanGroup by many blocks (each
// thread adds a run of 8 in index order, then the block scans the thread
// sums), the group totals by one block, and the group offsets are added
// back: the same fixed order on every run.
constexpr uint32_t kScanGroup = 2048;

__global__ __launch_bounds__(kScanThreads) void k_scan_group(double* __restrict__ chunk, uint64_t nchunks,
                                                             double* __restrict__ grp) {
    __shared__ double lds[4];
    constexpr uint32_t per = kScanGroup / kScanThreads;
    const uint64_t beg = uint64_t(blockIdx.x) * kScanGroup + uint64_t(threadIdx.x) * per;
    double v[per];
    double run = 0;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) {
        const uint64_t e = beg + i;
        run += e < nchunks ? chunk[e] : 0.0;
        v[i] = run;
    }
    double x = run;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double n = __shfl_up(x, o, 64);
        if (lane >= o) x += n;
    }
    double off = __shfl_up(x, 1, 64);
    if (lane == 0) off = 0;
    const int w = threadIdx.x >> 6;
    if (lane == 63) lds[w] = x;
    __syncthreads();
    for (int k = 0; k < w; ++k) off += lds[k];
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) {
        const uint64_t e = beg + i;
        if (e < nchunks) chunk[e] = off + v[i];
    }
    if (threadIdx.x == kScanThreads - 1) grp[blockIdx.x] = off + run;
}

__global__ __launch_bounds__(kScanThreads) void k_add_group_offsets(double* __restrict__ chunk, uint64_t nchunks,
                                                                    const double* __restrict__ grp) {
    const uint64_t e = uint64_t(blockIdx.x) * kScanThreads + threadIdx.x + kScanGroup;  // group 0 adds 0
    if (e < nchunks) chunk[e] += grp[e / kScanGroup - 1];
}

// In-place inclusive scan of up to a few thousand values (the group totals)
// by one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan_chunks(double* __restrict__ chunk, uint64_t nchunks) {
    __shared__ double lds[16];
    const uint64_t per = (nchunks + 1023) / 1024;
    const uint64_t beg = uint64_t(threadIdx.x) * per;
    double run = 0;
    for (uint64_t i = 0; i < per; ++i) {
        const uint64_t e = beg + i;
        if (e < nchunks) {
            run += chunk[e];
            chunk[e] = run;
        }
    }
    double v = run;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    double off = __shfl_up(v, 1, 64);
    if (lane == 0) off = 0;
    const int w = threadIdx.x >> 6;
    if (lane == 63) lds[w] = v;
    __syncthreads();
    for (int k = 0; k < w; ++k) off += lds[k];
    for (uint64_t i = 0; i < per; ++i) {
        const uint64_t e = beg + i;
        if (e < nchunks) chunk[e] += off;
    }
}

// One block per shot. out[s] = first index i with prob <= cdf(i) (strict:
// prob < cdf(i)), or `dflt` when the total stays below prob.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_search(const cv<T>* __restrict__ psi, ChunkGeom g,
                                                         const double* __restrict__ chunk, uint64_t nchunks,
                                                         const double* __restrict__ probs, int one_minus,
                                                         int strict, uint64_t dflt, uint64_t* __restrict__ out) {
    __shared__ double lds[4];
    __shared__ double nrm[kNrmLds];
    __shared__ uint64_t sc;
    __shared__ unsigned long long found;
    const double prob = one_minus ? 1.0 - probs[blockIdx.x] : probs[blockIdx.x];
    if (threadIdx.x == 0) {
        // first c with chunk[c] >= prob (strict: > prob)
        uint64_t lo = 0, hi = nchunks;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            const double v = chunk[mid];
            const bool ok = strict ? (v > prob) : (v >= prob);
            if (ok) hi = mid;
            else lo = mid + 1;
        }
        sc = lo;
        found = ~0ull;
    }
    __syncthreads();
    const uint64_t c = sc;
    if (c >= nchunks) {
        if (threadIdx.x == 0) out[blockIdx.x] = dflt;
        return;
    }
    double total;
    const uint64_t cbase = c * g.clen;
    const double pre = chunk_thread_prefix<T>(psi, cbase, g, lds, nrm, &total);
    double acc = (c > 0 ? chunk[c - 1] : 0.0) + pre;
    const uint64_t beg = uint64_t(threadIdx.x) * g.per;
    double run = 0;
    for (uint32_t i = 0; i < g.per; ++i) {
        const uint64_t e = beg + i;
        if (e >= g.clen) break;
        run += nrm[nrm_slot(e)];
        const double v = acc + run;
        const bool ok = strict ? (v > prob) : (v >= prob);
        if (ok) {
            atomicMin(&found, (unsigned long long)(cbase + e));
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long f = found;
        out[blockIdx.x] = (f == ~0ull) ? (cbase + g.clen - 1) : uint64_t(f);
    }
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_collapse(cv<T>* __restrict__ psi, uint64_t A, uint64_t mask,
                                                          uint64_t val, T scale, uint64_t boff) {
    // one amplitude per lane, one pass (no grid-stride loop)
    const uint64_t s = (uint64_t(blockIdx.x) + boff) * kRedThreads + threadIdx.x;
    if (s >= A) return;
    cv<T> v;
    if ((s & mask) == val) {
        v = ldamp<T, true>(psi + s);
        v.x *= scale;
        v.y *= scale;
    } else {
        v.x = 0;
        v.y = 0;
    }
    stamp<T, true>(psi + s, v);
}

// Exact Ising energy sum_s |psi_s|^2 E(s), E(s) = -(sum h z + sum J z z).
struct IsingArgs {
    uint32_t n, ne;
};

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_ising(const cv<T>* __restrict__ psi, uint64_t A, IsingArgs ia,
                                                       const double* __restrict__ hv,
                                                       const uint32_t* __restrict__ ei,
                                                       const uint32_t* __restrict__ ej,
                                                       const double* __restrict__ J,
                                                       double* __restrict__ partial) {
    __shared__ double lds[4];
    __shared__ double sh[64], sJ[256];
    __shared__ uint32_t si[256], sj[256];
    for (uint32_t i = threadIdx.x; i < ia.n; i += kRedThreads) sh[i] = hv[i];
    for (uint32_t k = threadIdx.x; k < ia.ne; k += kRedThreads) {
        sJ[k] = J[k];
        si[k] = ei[k];
        sj[k] = ej[k];
    }
    __syncthreads();
    double acc = 0;
    const uint64_t stride = uint64_t(gridDim.x) * kRedThreads;
    for (uint64_t s = uint64_t(blockIdx.x) * kRedThreads + threadIdx.x; s < A; s += stride) {
        double e = 0;
        for (uint32_t i = 0; i < ia.n; ++i) e += (s >> i & 1) ? -sh[i] : sh[i];
        for (uint32_t k = 0; k < ia.ne; ++k) {
            const uint64_t par = ((s >> si[k]) ^ (s >> sj[k])) & 1;
            e += par ? -sJ[k] : sJ[k];
        }
        acc += norm2<T>(psi[s]) * (-e);
    }
    const double t = block_sum256(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Host side.

static int red_blocks(uint64_t work) {
    uint64_t b = (work + kRedThreads * 4 - 1) / (kRedThreads * 4);
    if (b > uint64_t(kRedBlocks)) b = kRedBlocks;
    if (b < 1) b = 1;
    return int(b);
}

int fetch(qcs_register* r, const double* dev, int n, double* host) {
    if (hipMemcpyAsync(r->hpin, dev, sizeof(double) * n, hipMemcpyDeviceToHost, r->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "hipMemcpyAsync D2H failed");
    if (hipStreamSynchronize(r->stream) != hipSuccess) return set_error(QCS_EDEVICE, "hipStreamSynchronize failed");
    for (int i = 0; i < n; ++i) host[i] = r->hpin[i];
    return QCS_OK;
}

template <typename T>
static int rsum_sq_t(qcs_register* r, uint32_t first, uint32_t len, uint64_t val, double* out) {
    const uint64_t norb = r->A >> len;
    int nb = 0, e = 0;
    if (norb >= (uint64_t(1) << 20)) {
        // block sums go to the CDF buffer (2^n / 2048 doubles, scratch
        // between searches), then the fixed grid adds them
        const uint64_t nblk = (norb + kSumBlock - 1) / kSumBlock;
        for (uint64_t b0 = 0; b0 < nblk; b0 += (1ull << 22)) {
            const unsigned g = unsigned(std::min<uint64_t>(1ull << 22, nblk - b0));
            hipLaunchKernelGGL((k_sum_sq_blocks<T>), dim3(g), dim3(kRedThreads), 0, r->stream,
                               static_cast<const cv<T>*>(r->psi), norb, first, len, val, r->chunk, b0);
            if ((e = check_launch(r, "k_sum_sq_blocks"))) return e;
        }
        nb = red_blocks(nblk);
        hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(kRedThreads), 0, r->stream, r->chunk, nblk, r->red);
        if ((e = check_launch(r, "k_sum_partials"))) return e;
    } else {
        nb = red_blocks(norb);
        hipLaunchKernelGGL((k_sum_sq<T>), dim3(nb), dim3(kRedThreads), 0, r->stream,
                           static_cast<const cv<T>*>(r->psi), norb, first, len, val, r->red);
        if ((e = check_launch(r, "k_sum_sq"))) return e;
    }
    double* res = r->red + kRedResult;
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, r->stream, r->red, nb, 1, res);
    e = check_launch(r, "k_final_sum");
    if (e) return e;
    return fetch(r, res, 1, out);
}

int rsum_sq(qcs_register* r, uint32_t first, uint32_t len, uint64_t val, double* out) {
    return r->prec == QCS_COMPLEX128 ? rsum_sq_t<double>(r, first, len, val, out)
                                     : rsum_sq_t<float>(r, first, len, val, out);
}

template <typename Ta, typename Tb>
static int rinner_t(qcs_register* a, qcs_register* b, double* re_im, double* max_abs) {
    const int nb = red_blocks(a->A);
    hipLaunchKernelGGL((k_inner<Ta, Tb>), dim3(nb), dim3(kRedThreads), 0, a->stream,
                       static_cast<const cv<Ta>*>(a->psi), static_cast<const cv<Tb>*>(b->psi), a->A, a->red);
    int e = check_launch(a, "k_inner");
    if (e) return e;
    double* res = a->red + kRedResult;
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, a->stream, a->red, nb, 3, res);
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(64), 0, a->stream, a->red, nb, res);
    e = check_launch(a, "k_final_sum");
    if (e) return e;
    double tmp[3];
    e = fetch(a, res, 3, tmp);
    if (e) return e;
    re_im[0] = tmp[0];
    re_im[1] = tmp[1];
    if (max_abs) *max_abs = tmp[2];
    return QCS_OK;
}

int rinner(qcs_register* a, qcs_register* b, double* re_im, double* max_abs) {
    // b's stream must be drained before a's stream reads it.
    if (hipStreamSynchronize(b->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
    if (a->prec == QCS_COMPLEX128) {
        return b->prec == QCS_COMPLEX128 ? rinner_t<double, double>(a, b, re_im, max_abs)
                                         : rinner_t<double, float>(a, b, re_im, max_abs);
    }
    return b->prec == QCS_COMPLEX128 ? rinner_t<float, double>(a, b, re_im, max_abs)
                                     : rinner_t<float, float>(a, b, re_im, max_abs);
}

static ChunkGeom chunk_geom(const qcs_register* r) {
    ChunkGeom g;
    g.clen = r->A < (1ull << kChunkLog) ? r->A : (1ull << kChunkLog);
    g.per = uint32_t((g.clen + kScanThreads - 1) / kScanThreads);
    return g;
}

template <typename T>
static int cdf_search_t(qcs_register* r, uint64_t nshots, const double* vals_host, int one_minus, int strict,
                        uint64_t dflt, uint64_t* out_host, bool build, const ShotDev* dev) {
    const ChunkGeom g = chunk_geom(r);
    const uint64_t nchunks = r->A / g.clen;
    const auto* psi = static_cast<const cv<T>*>(r->psi);
    // at most 2^22 chunks (2^30 work-items) per launch: a grid's size in
    // work-items must stay below 2^32
    int e = 0;
    for (uint64_t c0 = 0; build && c0 < nchunks; c0 += (1ull << 22)) {
        const uint64_t cnt = std::min<uint64_t>(1ull << 22, nchunks - c0);
        hipLaunchKernelGGL((k_chunk_sums<T>), dim3(unsigned(cnt)), dim3(kScanThreads), 0, r->stream, psi, g,
                           r->chunk, c0);
        e = check_launch(r, "k_chunk_sums");
        if (e) return e;
    }
    // group totals live after the chunk totals (the buffer holds both)
    const uint64_t ngroups = (nchunks + kScanGroup - 1) / kScanGroup;
    double* grp = r->chunk + nchunks;
    if (build) {
        hipLaunchKernelGGL(k_scan_group, dim3(unsigned(ngroups)), dim3(kScanThreads), 0, r->stream, r->chunk,
                           nchunks, grp);
        e = check_launch(r, "k_scan_group");
        if (e) return e;
    }
    if (build && ngroups > 1) {
        hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(1024), 0, r->stream, grp, ngroups);
        e = check_launch(r, "k_scan_chunks");
        if (e) return e;
        const uint64_t nadd = nchunks - kScanGroup;
        hipLaunchKernelGGL(k_add_group_offsets, dim3(unsigned((nadd + kScanThreads - 1) / kScanThreads)),
                           dim3(kScanThreads), 0, r->stream, r->chunk, nchunks, grp);
        e = check_launch(r, "k_add_group_offsets");
        if (e) return e;
    }
    // per-shot buffers: the caller's, or stream-ordered allocations
    double* dvals = nullptr;
    uint64_t* douts = nullptr;
    if (dev) {
        if (nshots > dev->cap) return set_error(QCS_EINVAL, "%llu shots exceed the buffer of %llu",
                                                (unsigned long long)nshots, (unsigned long long)dev->cap);
        dvals = dev->in;
        douts = dev->out;
    } else if (hipMallocAsync(reinterpret_cast<void**>(&dvals), sizeof(double) * nshots, r->stream) != hipSuccess ||
               hipMallocAsync(reinterpret_cast<void**>(&douts), sizeof(uint64_t) * nshots, r->stream) != hipSuccess) {
        return set_error(QCS_ENOMEM, "shot buffers");
    }
    if (hipMemcpyAsync(dvals, vals_host, sizeof(double) * nshots, hipMemcpyHostToDevice, r->stream) != hipSuccess)
        e = set_error(QCS_EDEVICE, "shot upload failed");
    for (uint64_t s0 = 0; s0 < nshots && !e; s0 += 65535) {
        const uint64_t cnt = nshots - s0 < 65535 ? nshots - s0 : 65535;
        hipLaunchKernelGGL((k_search<T>), dim3(unsigned(cnt)), dim3(kScanThreads), 0, r->stream, psi, g, r->chunk,
                           nchunks, dvals + s0, one_minus, strict, dflt, douts + s0);
        e = check_launch(r, "k_search");
    }
    if (!e && hipMemcpyAsync(out_host, douts, sizeof(uint64_t) * nshots, hipMemcpyDeviceToHost, r->stream) !=
                  hipSuccess)
        e = set_error(QCS_EDEVICE, "shot download failed");
    if (!dev) {
        (void)hipFreeAsync(dvals, r->stream);
        (void)hipFreeAsync(douts, r->stream);
    }
    if (hipStreamSynchronize(r->stream) != hipSuccess && !e) e = set_error(QCS_EDEVICE, "cdf search sync failed");
    return e;
}

int cdf_search(qcs_register* r, uint64_t nshots, const double* vals_host, int one_minus, int strict, uint64_t dflt,
               uint64_t* out_host, int build_cdf, const ShotDev* dev) {
    if (nshots == 0) return QCS_OK;
    return r->prec == QCS_COMPLEX128
               ? cdf_search_t<double>(r, nshots, vals_host, one_minus, strict, dflt, out_host, build_cdf, dev)
               : cdf_search_t<float>(r, nshots, vals_host, one_minus, strict, dflt, out_host, build_cdf, dev);
}

int kcollapse(qcs_register* r, uint64_t mask, uint64_t val, double scale) {
    // at most 2^22 blocks (2^30 work-items) per launch: a grid's size in
    // work-items must stay below 2^32
    const uint64_t nblk = (r->A + kRedThreads - 1) / kRedThreads;
    for (uint64_t boff = 0; boff < nblk; boff += (1ull << 22)) {
        const unsigned g = unsigned(std::min<uint64_t>(1ull << 22, nblk - boff));
        if (r->prec == QCS_COMPLEX128)
            hipLaunchKernelGGL((k_collapse<double>), dim3(g), dim3(kRedThreads), 0, r->stream,
                               static_cast<double2*>(r->psi), r->A, mask, val, scale, boff);
        else
            hipLaunchKernelGGL((k_collapse<float>), dim3(g), dim3(kRedThreads), 0, r->stream,
                               static_cast<float2*>(r->psi), r->A, mask, val, float(scale), boff);
        if (const int e = check_launch(r, "k_collapse")) return e;
    }
    return QCS_OK;
}

template <typename T>
static int rising_t(qcs_register* r, uint32_t ne, const uint32_t* ei, const uint32_t* ej, const double* J,
                    const double* h, double* energy) {
    if (r->n > 64 || ne > 256) return set_error(QCS_EINVAL, "ising: at most 64 sites and 256 edges");
    // pack inputs into one device block
    const size_t bytes = sizeof(double) * (r->n + ne) + sizeof(uint32_t) * 2 * ne + 64;
    char* d = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&d), bytes, r->stream) != hipSuccess)
        return set_error(QCS_ENOMEM, "ising args");
    std::vector<char> hbuf(bytes);
    double* hh = reinterpret_cast<double*>(hbuf.data());
    for (uint32_t i = 0; i < r->n; ++i) hh[i] = h[i];
    for (uint32_t k = 0; k < ne; ++k) hh[r->n + k] = J[k];
    uint32_t* hi = reinterpret_cast<uint32_t*>(hbuf.data() + sizeof(double) * (r->n + ne));
    for (uint32_t k = 0; k < ne; ++k) {
        hi[k] = ei[k];
        hi[ne + k] = ej[k];
    }
    if (hipMemcpyAsync(d, hbuf.data(), bytes, hipMemcpyHostToDevice, r->stream) != hipSuccess) {
        (void)hipFreeAsync(d, r->stream);
        return set_error(QCS_EDEVICE, "ising args upload");
    }
    const double* dh = reinterpret_cast<const double*>(d);
    const uint32_t* dij = reinterpret_cast<const uint32_t*>(d + sizeof(double) * (r->n + ne));
    const int nb = red_blocks(r->A);
    IsingArgs ia{r->n, ne};
    hipLaunchKernelGGL((k_ising<T>), dim3(nb), dim3(kRedThreads), 0, r->stream, static_cast<const cv<T>*>(r->psi),
                       r->A, ia, dh, dij, dij + ne, dh + r->n, r->red);
    int e = check_launch(r, "k_ising");
    double* res = r->red + kRedResult;
    if (!e) {
        hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kRedThreads), 0, r->stream, r->red, nb, 1, res);
        e = check_launch(r, "k_final_sum");
    }
    (void)hipFreeAsync(d, r->stream);
    if (e) return e;
    return fetch(r, res, 1, energy);
}

int rising(qcs_register* r, uint32_t ne, const uint32_t* ei, const uint32_t* ej, const double* J, const double* h,
           double* energy) {
    return r->prec == QCS_COMPLEX128 ? rising_t<double>(r, ne, ei, ej, J, h, energy)
                                     : rising_t<float>(r, ne, ei, ej, J, h, energy);
}

}  // namespace qcs
