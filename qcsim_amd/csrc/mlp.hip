// mlp.hip - MLP surrogate for QAOA energies: forward, backward and Adam in
// hand-written HIP kernels. No counterpart in the reference.
//
// Network: in -> H -> H -> ... -> 1 (L hidden layers of width H, ReLU), BF16
// weights and activations, FP32 accumulation; FP32 master weights, FP32
// gradients, Adam. Loss: mean squared error. BF16 is stored as the top 16
// bits of an IEEE float, rounded to nearest even (bf16_from), so a CPU
// restatement reproduces every rounding.
//
// The GEMMs run on the vector ALUs through a 64x64x16 LDS-tiled kernel
// (each thread 4x4 outputs): the matrix cores are kept for the fused gate
// blocks of the statevector engine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"

namespace qcs {

typedef uint16_t bf16;

__host__ __device__ inline bf16 bf16_from(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16(u >> 16);
}
__host__ __device__ inline float bf16_to(bf16 h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

template <typename T> __device__ inline float ldf(const T* p) { return float(*p); }
template <> __device__ inline float ldf<bf16>(const bf16* p) { return bf16_to(*p); }

enum { EPI_RELU_BF16 = 0, EPI_LINEAR_F32 = 1, EPI_MASK_F32 = 2, EPI_F32 = 3 };

// C[M x N] = op(A)[M x K] * op(B)[K x N] with an epilogue.
// A is M x K row-major (lda) or, with TA, K x M row-major; B likewise.
template <typename TA, typename TB, bool TA_, bool TB_, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const TA* __restrict__ A, int lda,
                                              const TB* __restrict__ B, int ldb, float* __restrict__ Cf,
                                              bf16* __restrict__ Cb, int ldc, const float* __restrict__ bias,
                                              const bf16* __restrict__ mask, int ldm) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = threadIdx.x; e < 1024; e += 256) {
            // A tile: 64 rows (m) x 16 cols (k)
            const int mm = e >> 4, kk = e & 15;
            const int gm = m0 + mm, gk = k0 + kk;
            float v = 0.f;
            if (gm < M && gk < K) v = TA_ ? ldf<TA>(A + size_t(gk) * lda + gm) : ldf<TA>(A + size_t(gm) * lda + gk);
            As[kk][mm] = v;
            // B tile: 16 rows (k) x 64 cols (n)
            const int kb = e >> 6, nb = e & 63;
            const int gkb = k0 + kb, gn = n0 + nb;
            float w = 0.f;
            if (gkb < K && gn < N) w = TB_ ? ldf<TB>(B + size_t(gn) * ldb + gkb) : ldf<TB>(B + size_t(gkb) * ldb + gn);
            Bs[kb][nb] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            float v = acc[i][j];
            if (EPI == EPI_RELU_BF16) {
                v += bias[gn];
                Cb[size_t(gm) * ldc + gn] = bf16_from(v > 0.f ? v : 0.f);
            } else if (EPI == EPI_LINEAR_F32) {
                Cf[size_t(gm) * ldc + gn] = v + bias[gn];
            } else if (EPI == EPI_MASK_F32) {
                Cf[size_t(gm) * ldc + gn] = (bf16_to(mask[size_t(gm) * ldm + gn]) > 0.f) ? v : 0.f;
            } else {
                Cf[size_t(gm) * ldc + gn] = v;
            }
        }
    }
}

// db[n] = sum_m dZ[m, n] (fixed order per column).
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += dZ[size_t(m) * N + n];
    db[n] = s;
}

// Output layer gradient for MSE: dY = 2 (y - t) / B; per-sample loss terms.
__global__ void k_mse(int Bn, const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                      float* __restrict__ lossv) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Bn) return;
    const float d = y[i] - t[i];
    dy[i] = 2.f * d / float(Bn);
    lossv[i] = d * d;
}

__global__ void k_sum_f32(int n, const float* __restrict__ v, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Adam on FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(int n, float* __restrict__ w, bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = bf16_from(wi);
}

__global__ void k_to_bf16(int n, const float* __restrict__ x, bf16* __restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = bf16_from(x[i]);
}

// Gather a batch of rows (by index) from the device dataset, inputs to BF16.
__global__ void k_gather(int Bn, int D, const uint32_t* __restrict__ idx, const float* __restrict__ X,
                         const float* __restrict__ Y, bf16* __restrict__ xb, float* __restrict__ yb) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= Bn * D) return;
    const int r = e / D, c = e % D;
    const uint32_t src = idx[r];
    xb[e] = bf16_from(X[size_t(src) * D + c]);
    if (c == 0) yb[r] = Y[src];
}

}  // namespace qcs

using namespace qcs;

struct qcs_mlp {
    int in = 0, H = 0, L = 0;
    std::vector<int> dims;             // in, H, ..., H, 1
    std::vector<size_t> woff, boff;    // offsets into the flat parameter vector
    size_t nparams = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 master / grad / Adam
    bf16* wb = nullptr;                                            // BF16 weights (biases unused in bf16)
    // activations for the current batch
    int cap = 0;
    bf16* x0 = nullptr;
    std::vector<bf16*> act;  // outputs of hidden layers (bf16)
    float* out = nullptr;     // network output (fp32)
    float *dz0 = nullptr, *dz1 = nullptr, *yt = nullptr, *lossv = nullptr, *lsum = nullptr;
    // dataset
    float *dX = nullptr, *dY = nullptr;
    uint32_t* didx = nullptr;
    uint64_t nrows = 0;
    hipStream_t stream = nullptr;
    uint64_t step = 0;
    // host staging of the bf16 inputs; refilled only after the stream has
    // drained, so its previous upload has completed
    std::vector<bf16> hx;
};

static int mlp_reserve(qcs_mlp* n, int B) {
    if (B <= n->cap) return QCS_OK;
    auto fr = [](void* p) {
        if (p) (void)hipFree(p);
    };
    fr(n->x0);
    for (auto* a : n->act) fr(a);
    fr(n->out);
    fr(n->dz0);
    fr(n->dz1);
    fr(n->yt);
    fr(n->lossv);
    fr(n->lsum);
    n->act.assign(n->L, nullptr);
    bool ok = hipMalloc(reinterpret_cast<void**>(&n->x0), sizeof(bf16) * B * n->in) == hipSuccess;
    for (int l = 0; l < n->L; ++l)
        ok = ok && hipMalloc(reinterpret_cast<void**>(&n->act[l]), sizeof(bf16) * B * n->H) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->out), sizeof(float) * B) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->dz0), sizeof(float) * B * n->H) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->dz1), sizeof(float) * B * n->H) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->yt), sizeof(float) * B) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->lossv), sizeof(float) * B) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&n->lsum), sizeof(float) * 4) == hipSuccess;
    if (!ok) return set_error(QCS_ENOMEM, "mlp activations");
    n->cap = B;
    return QCS_OK;
}

static dim3 ggrid(int M, int N) { return dim3(unsigned((N + 63) / 64), unsigned((M + 63) / 64)); }

// Forward for B rows already in n->x0 (bf16).
static int mlp_forward(qcs_mlp* n, int B) {
    const bf16* x = n->x0;
    int din = n->in;
    for (int l = 0; l < n->L; ++l) {
        hipLaunchKernelGGL((k_gemm<bf16, bf16, false, false, EPI_RELU_BF16>), ggrid(B, n->H), dim3(256), 0, n->stream, B,
                           n->H, din, x, din, n->wb + n->woff[l], n->H, nullptr, n->act[l], n->H, n->w + n->boff[l],
                           nullptr, 0);
        x = n->act[l];
        din = n->H;
    }
    hipLaunchKernelGGL((k_gemm<bf16, bf16, false, false, EPI_LINEAR_F32>), ggrid(B, 1), dim3(256), 0, n->stream, B, 1,
                       din, x, din, n->wb + n->woff[n->L], 1, n->out, nullptr, 1, n->w + n->boff[n->L], nullptr, 0);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "mlp forward: %s", hipGetErrorString(e));
}

// Backward (MSE against n->yt) into n->g; loss into n->lsum[0].
static int mlp_backward(qcs_mlp* n, int B) {
    hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, n->stream, B, n->out, n->yt, n->dz0, n->lossv);
    hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(256), 0, n->stream, B, n->lossv, n->lsum);
    // dz0 holds dY (B x 1) for the output layer
    float* dz = n->dz0;
    float* other = n->dz1;
    for (int l = n->L; l >= 0; --l) {
        const int dout = n->dims[l + 1], din = n->dims[l];
        const bf16* xin = l == 0 ? n->x0 : n->act[l - 1];
        // dW = x^T dz : din x dout
        hipLaunchKernelGGL((k_gemm<bf16, float, true, false, EPI_F32>), ggrid(din, dout), dim3(256), 0, n->stream, din,
                           dout, B, xin, din, dz, dout, n->g + n->woff[l], nullptr, dout, nullptr, nullptr, 0);
        hipLaunchKernelGGL(k_colsum, dim3((dout + 255) / 256), dim3(256), 0, n->stream, B, dout, dz, n->g + n->boff[l]);
        if (l == 0) break;
        // dx = dz W^T, masked by ReLU'(x) (x = act[l-1] > 0): B x din
        hipLaunchKernelGGL((k_gemm<float, bf16, false, true, EPI_MASK_F32>), ggrid(B, din), dim3(256), 0, n->stream, B,
                           din, dout, dz, dout, n->wb + n->woff[l], dout, other, nullptr, din, nullptr, n->act[l - 1], din);
        std::swap(dz, other);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "mlp backward: %s", hipGetErrorString(e));
}

static int mlp_adam(qcs_mlp* n, float lr) {
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, float(n->step)), c2 = 1.f - std::pow(b2, float(n->step));
    hipLaunchKernelGGL(k_adam, dim3(unsigned((n->nparams + 255) / 256)), dim3(256), 0, n->stream, int(n->nparams), n->w,
                       n->wb, n->g, n->m, n->v, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError() == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "adam launch failed");
}

extern "C" {

int qcs_mlp_destroy(qcs_mlp* n) {
    if (!n) return QCS_OK;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    void* ps[] = {n->w, n->g, n->m, n->v, n->wb, n->x0, n->out, n->dz0, n->dz1, n->yt, n->lossv, n->lsum, n->dX, n->dY, n->didx};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (auto* a : n->act)
        if (a) (void)hipFree(a);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return QCS_OK;
}

int qcs_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t layers, uint64_t seed, qcs_mlp** out) {
    if (!out || in_dim == 0 || hidden == 0 || layers == 0 || layers > 8) return set_error(QCS_EINVAL, "bad MLP shape");
    auto* n = new qcs_mlp;
    n->in = int(in_dim);
    n->H = int(hidden);
    n->L = int(layers);
    n->dims.push_back(n->in);
    for (int l = 0; l < n->L; ++l) n->dims.push_back(n->H);
    n->dims.push_back(1);
    size_t off = 0;
    for (int l = 0; l <= n->L; ++l) {
        n->woff.push_back(off);
        off += size_t(n->dims[l]) * n->dims[l + 1];
        n->boff.push_back(off);
        off += size_t(n->dims[l + 1]);
    }
    n->nparams = off;
    // He-uniform initialisation, zero biases
    std::vector<float> w(off, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l <= n->L; ++l) {
        const float lim = std::sqrt(6.f / float(n->dims[l]));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (size_t i = 0; i < size_t(n->dims[l]) * n->dims[l + 1]; ++i) w[n->woff[l] + i] = U(rng);
    }
    bool ok = hipMalloc(reinterpret_cast<void**>(&n->w), sizeof(float) * off) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&n->g), sizeof(float) * off) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&n->m), sizeof(float) * off) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&n->v), sizeof(float) * off) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&n->wb), sizeof(bf16) * off) == hipSuccess &&
              hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        qcs_mlp_destroy(n);
        return set_error(QCS_ENOMEM, "mlp allocation failed");
    }
    std::vector<bf16> wb(off);
    for (size_t i = 0; i < off; ++i) wb[i] = bf16_from(w[i]);
    std::vector<float> zero(off, 0.f);
    ok = hipMemcpy(n->w, w.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->wb, wb.data(), sizeof(bf16) * off, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->m, zero.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->v, zero.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        qcs_mlp_destroy(n);
        return set_error(QCS_EDEVICE, "mlp upload failed");
    }
    *out = n;
    return QCS_OK;
}

uint64_t qcs_mlp_num_params(const qcs_mlp* n) { return n ? n->nparams : 0; }

// FP32 master parameters (the BF16 copy is refreshed on set).
int qcs_mlp_get_params(qcs_mlp* n, float* p) {
    if (!n || !p) return set_error(QCS_EINVAL, "null argument");
    return hipMemcpy(p, n->w, sizeof(float) * n->nparams, hipMemcpyDeviceToHost) == hipSuccess
               ? QCS_OK
               : set_error(QCS_EDEVICE, "download failed");
}

int qcs_mlp_set_params(qcs_mlp* n, const float* p) {
    if (!n || !p) return set_error(QCS_EINVAL, "null argument");
    if (hipMemcpy(n->w, p, sizeof(float) * n->nparams, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(QCS_EDEVICE, "upload failed");
    hipLaunchKernelGGL(k_to_bf16, dim3(unsigned((n->nparams + 255) / 256)), dim3(256), 0, n->stream, int(n->nparams),
                       n->w, n->wb);
    return hipStreamSynchronize(n->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "sync failed");
}

static int upload_inputs(qcs_mlp* n, const float* x, int B) {
    if (hipStreamSynchronize(n->stream) != hipSuccess) return set_error(QCS_EDEVICE, "sync failed");
    std::vector<bf16>& xb = n->hx;
    xb.resize(size_t(B) * n->in);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = bf16_from(x[i]);
    return hipMemcpyAsync(n->x0, xb.data(), sizeof(bf16) * xb.size(), hipMemcpyHostToDevice, n->stream) == hipSuccess
               ? QCS_OK
               : set_error(QCS_EDEVICE, "input upload failed");
}

int qcs_mlp_forward(qcs_mlp* n, const float* x, uint32_t batch, float* y) {
    if (!n || !x || !y || batch == 0) return set_error(QCS_EINVAL, "bad argument");
    for (uint32_t b0 = 0; b0 < batch; b0 += 65536) {
        const int B = int(std::min<uint32_t>(65536, batch - b0));
        int e = mlp_reserve(n, B);
        if (!e) e = upload_inputs(n, x + size_t(b0) * n->in, B);
        if (!e) e = mlp_forward(n, B);
        if (e) return e;
        if (hipMemcpyAsync(y + b0, n->out, sizeof(float) * B, hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
            hipStreamSynchronize(n->stream) != hipSuccess)
            return set_error(QCS_EDEVICE, "forward failed");
    }
    return QCS_OK;
}

// Loss and FP32 gradients (flat, same layout as the parameters) of one batch.
int qcs_mlp_gradients(qcs_mlp* n, const float* x, const float* t, uint32_t batch, float* grad, float* loss) {
    if (!n || !x || !t || batch == 0) return set_error(QCS_EINVAL, "bad argument");
    const int B = int(batch);
    int e = mlp_reserve(n, B);
    if (!e) e = upload_inputs(n, x, B);
    if (!e && hipMemcpyAsync(n->yt, t, sizeof(float) * B, hipMemcpyHostToDevice, n->stream) != hipSuccess)
        e = set_error(QCS_EDEVICE, "target upload failed");
    if (!e) e = mlp_forward(n, B);
    if (!e) e = mlp_backward(n, B);
    if (e) return e;
    float l = 0;
    if ((grad && hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparams, hipMemcpyDeviceToHost, n->stream) != hipSuccess) ||
        hipMemcpyAsync(&l, n->lsum, sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
        hipStreamSynchronize(n->stream) != hipSuccess)
        return set_error(QCS_EDEVICE, "gradient download failed");
    if (loss) *loss = l / float(B);
    return QCS_OK;
}

// Train on a device-resident copy of (X, Y): `epochs` passes of shuffled
// mini-batches, Adam with learning rate lr. loss_per_epoch (optional)
// receives the mean training loss of each epoch.
int qcs_mlp_train(qcs_mlp* n, const float* X, const float* Y, uint64_t rows, uint32_t batch, uint32_t epochs, float lr,
                  uint64_t seed, float* loss_per_epoch) {
    if (!n || !X || !Y || rows == 0 || batch == 0) return set_error(QCS_EINVAL, "bad argument");
    if (n->nrows != rows) {
        if (n->dX) (void)hipFree(n->dX);
        if (n->dY) (void)hipFree(n->dY);
        if (n->didx) (void)hipFree(n->didx);
        n->dX = nullptr;
        n->dY = nullptr;
        n->didx = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&n->dX), sizeof(float) * rows * n->in) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&n->dY), sizeof(float) * rows) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&n->didx), sizeof(uint32_t) * rows) != hipSuccess)
            return set_error(QCS_ENOMEM, "dataset allocation failed");
        n->nrows = rows;
    }
    if (hipMemcpy(n->dX, X, sizeof(float) * rows * n->in, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->dY, Y, sizeof(float) * rows, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(QCS_EDEVICE, "dataset upload failed");
    int e = mlp_reserve(n, int(batch));
    if (e) return e;
    std::vector<uint32_t> perm(rows);
    for (uint64_t i = 0; i < rows; ++i) perm[i] = uint32_t(i);
    std::mt19937_64 rng(seed);
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpy(n->didx, perm.data(), sizeof(uint32_t) * rows, hipMemcpyHostToDevice) != hipSuccess)
            return set_error(QCS_EDEVICE, "index upload failed");
        double lsum = 0;
        uint64_t nb = 0;
        for (uint64_t b0 = 0; b0 + batch <= rows; b0 += batch) {
            const int B = int(batch);
            hipLaunchKernelGGL(k_gather, dim3(unsigned((B * n->in + 255) / 256)), dim3(256), 0, n->stream, B, n->in,
                               n->didx + b0, n->dX, n->dY, n->x0, n->yt);
            if ((e = mlp_forward(n, B)) || (e = mlp_backward(n, B)) || (e = mlp_adam(n, lr))) return e;
            if (loss_per_epoch) {
                float l = 0;
                if (hipMemcpyAsync(&l, n->lsum, sizeof(float), hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
                    hipStreamSynchronize(n->stream) != hipSuccess)
                    return set_error(QCS_EDEVICE, "loss download failed");
                lsum += l / B;
            }
            ++nb;
        }
        if (loss_per_epoch) loss_per_epoch[ep] = nb ? float(lsum / double(nb)) : 0.f;
    }
    return hipStreamSynchronize(n->stream) == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "train sync failed");
}

}  // extern "C"
