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
// Work decomposition for an n-qubit instance with group width G:
//   n <= Tmax: one SMALL unit runs every layer and the energy in LDS;
//   else, with tile bits T = Tmax, per layer
//     stage A: 2^(n-T) units, each a 2^T tile: (layer 0: |+> generated in
//              place) cost diagonal, Rx on qubits 0..T-1 in LDS;
//     stage B_j (j = 0 .. ceil((n-T)/G) - 1): Rx on qubits T+jG .. (the
//              last group may be narrower) with each thread's orbit in
//              registers; the last stage of the last layer also
//              accumulates the energy.
// A stage's kernel pulls units from an atomic ticket until none are left.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace qcs {

struct QInst {
    uint64_t off;        // first amplitude in the arena
    uint32_t n, ne;      // qubits, edges
    uint32_t eoff, hoff; // into edge / site arrays
    uint32_t tbits;      // tile bits T (== n for SMALL instances)
    uint32_t pslot0;     // first energy-partial slot
    uint32_t npslots;    // number of energy partials
    uint32_t pad;
};

enum : uint32_t { U_SMALL = 0, U_A = 1, U_B = 2 };

struct QUnit {
    uint32_t inst;
    uint32_t kind;   // U_* | (energy << 8) | (group width << 16)
    uint32_t idx;    // tile (A) or 256-orbit chunk (B)
    uint32_t layer;
    uint32_t qlo;    // first qubit of the B group
    uint32_t slot;   // energy partial slot
};

constexpr uint32_t kMaxSites = 64, kMaxEdges = 256;

// C(s) and E(s) for one basis state.
__device__ __forceinline__ void ising_terms(uint64_t s, uint32_t n, uint32_t ne, const double* sh, const double* sJ,
                                            const uint8_t* si, const uint8_t* sj, double& cost, double& energy) {
    double hz = 0, jzz = 0;
    for (uint32_t i = 0; i < n; ++i) hz += ((s >> i) & 1) ? -sh[i] : sh[i];
    for (uint32_t k = 0; k < ne; ++k) jzz += (((s >> si[k]) ^ (s >> sj[k])) & 1) ? -sJ[k] : sJ[k];
    cost = jzz + 0.5 * hz;
    energy = -(hz + jzz);
}

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

// In-LDS layer pieces for a tile of 2^tb amplitudes starting at index `base`.
__device__ void lds_cost(float2* t, uint32_t tb, uint64_t base, double gamma, const QInst& I, const double* sh,
                         const double* sJ, const uint8_t* si, const uint8_t* sj) {
    const uint32_t C = 1u << tb;
    for (uint32_t i = threadIdx.x; i < C; i += 256) {
        double c, e;
        ising_terms(base + i, I.n, I.ne, sh, sJ, si, sj, c, e);
        t[i] = cost_phase(t[i], gamma, c);
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

// Rx(2 beta) on qubits qlo .. qlo+g-1, each thread one 2^g-amplitude orbit
// in registers; returns this thread's share of the energy when asked.
template <int g>
__device__ double stage_b(float2* __restrict__ psi, const QUnit& U, const QInst& I, const double* ga, bool want_e,
                          const double* sh, const double* sJ, const uint8_t* si, const uint8_t* sj) {
    constexpr int D = 1 << g;
    const uint64_t o = uint64_t(U.idx) * 256 + threadIdx.x;
    const uint64_t norb = (1ull << I.n) >> g;
    double acc = 0;
    if (o >= norb) return 0;
    const uint64_t base = insert_run(o, U.qlo, g);
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = ldamp<float, true>(psi + (base | (uint64_t(c) << U.qlo)));
    double sb, cb;
    sincos(ga[2 * U.layer + 1], &sb, &cb);
    const float fc = float(cb), fs = float(sb);
#pragma unroll
    for (int q = 0; q < g; ++q)
#pragma unroll
        for (int c = 0; c < D; ++c)
            if (!((c >> q) & 1)) rx_pair(x[c], x[c | (1 << q)], fc, fs);
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const uint64_t s = base | (uint64_t(c) << U.qlo);
        if (want_e) {
            double cc, e;
            ising_terms(s, I.n, I.ne, sh, sJ, si, sj, cc, e);
            const double a = x[c].x, b = x[c].y;
            acc += (a * a + b * b) * e;
        }
        stamp<float, true>(psi + s, x[c]);
    }
    return acc;
}

template <int G>
__global__ __launch_bounds__(256) void k_qaoa_stage(float2* __restrict__ arena, const QInst* __restrict__ insts,
                                                    const QUnit* __restrict__ units, uint32_t nunits,
                                                    uint32_t* __restrict__ ticket, const double* __restrict__ ang,
                                                    uint32_t p, const double* __restrict__ hv,
                                                    const uint8_t* __restrict__ ei, const uint8_t* __restrict__ ej,
                                                    const double* __restrict__ Jv, double* __restrict__ partial,
                                                    double* __restrict__ energy) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    float2* tile = reinterpret_cast<float2*>(smem_raw);
    __shared__ double sh[kMaxSites], sJ[kMaxEdges], red[4];
    __shared__ uint8_t si[kMaxEdges], sj[kMaxEdges];
    __shared__ uint32_t su;
    for (;;) {
        if (threadIdx.x == 0) su = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t u = su;
        __syncthreads();
        if (u >= nunits) break;  // every wave of the block leaves together
        const QUnit U = units[u];
        const QInst I = insts[U.inst];
        for (uint32_t i = threadIdx.x; i < I.n; i += 256) sh[i] = hv[I.hoff + i];
        for (uint32_t k = threadIdx.x; k < I.ne; k += 256) {
            sJ[k] = Jv[I.eoff + k];
            si[k] = ei[I.eoff + k];
            sj[k] = ej[I.eoff + k];
        }
        __syncthreads();
        float2* psi = arena + I.off;
        const uint32_t kind = U.kind & 0xff;
        const bool want_e = (U.kind >> 8) & 1;
        const double* ga = ang + size_t(U.inst) * p * 2;
        if (kind == U_SMALL) {
            const uint32_t C = 1u << I.n;
            const float amp = float(1.0 / sqrt(double(C)));
            for (uint32_t i = threadIdx.x; i < C; i += 256) tile[i] = make_float2(amp, 0.f);
            __syncthreads();
            for (uint32_t l = 0; l < p; ++l) {
                lds_cost(tile, I.n, 0, ga[2 * l], I, sh, sJ, si, sj);
                double sb, cb;
                sincos(ga[2 * l + 1], &sb, &cb);
                lds_rx(tile, I.n, float(cb), float(sb));
            }
            double acc = 0;
            for (uint32_t i = threadIdx.x; i < C; i += 256) {
                double c, e;
                ising_terms(i, I.n, I.ne, sh, sJ, si, sj, c, e);
                const double x = tile[i].x, y = tile[i].y;
                acc += (x * x + y * y) * e;
                psi[i] = tile[i];
            }
            const double tot = block_sum(acc, red);
            if (threadIdx.x == 0) energy[U.inst] = tot;
        } else if (kind == U_A) {
            const uint32_t tb = I.tbits, C = 1u << tb;
            const uint64_t base = uint64_t(U.idx) << tb;
            if (U.layer == 0) {
                const float amp = float(1.0 / sqrt(double(1ull << I.n)));
                for (uint32_t i = threadIdx.x; i < C; i += 256) tile[i] = make_float2(amp, 0.f);
            } else {
                for (uint32_t i = threadIdx.x; i < C; i += 256) tile[i] = ldamp<float, true>(psi + base + i);
            }
            __syncthreads();
            lds_cost(tile, tb, base, ga[2 * U.layer], I, sh, sJ, si, sj);
            double sb, cb;
            sincos(ga[2 * U.layer + 1], &sb, &cb);
            lds_rx(tile, tb, float(cb), float(sb));
            for (uint32_t i = threadIdx.x; i < C; i += 256) stamp<float, true>(psi + base + i, tile[i]);
        } else {
            const uint32_t g = (U.kind >> 16) & 0xff;
            double acc = 0;
            switch (g) {
                case 1: acc = stage_b<1>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
                case 2: acc = stage_b<2>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
                case 3: acc = stage_b<3>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
                case 4: acc = stage_b<G >= 4 ? 4 : G>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
                case 5: acc = stage_b<G >= 5 ? 5 : G>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
                default: acc = stage_b<G>(psi, U, I, ga, want_e, sh, sJ, si, sj); break;
            }
            if (want_e) {
                const double tot = block_sum(acc, red);
                if (threadIdx.x == 0) partial[U.slot] = tot;
            }
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
    std::vector<uint32_t> ei_h, ej_h;
    std::vector<double> J_h, h_h;
    uint64_t total_amps = 0;
    float2* arena = nullptr;
    QInst* d_inst = nullptr;
    uint8_t *d_ei = nullptr, *d_ej = nullptr;
    double *d_J = nullptr, *d_h = nullptr, *d_ang = nullptr, *d_energy = nullptr, *d_partial = nullptr;
    uint32_t* d_large = nullptr;
    uint32_t nlarge = 0;
    // stage plan
    std::vector<QUnit*> d_units;
    std::vector<uint32_t> stage_units;
    uint32_t* d_tickets = nullptr;
    uint32_t tmax = 12, group = 6;
    int use_graph = 1;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;
    std::vector<QInst> inst_h;
};

static void free_plan(qcs_qaoa* q) {
    for (auto* u : q->d_units) (void)hipFree(u);
    q->d_units.clear();
    q->stage_units.clear();
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
    const uint32_t nst = uint32_t(q->stage_units.size());
    hipLaunchKernelGGL(k_qaoa_reset, dim3(1), dim3(256), 0, q->stream, q->d_tickets, nst);
    const size_t shm = sizeof(float2) << q->tmax;
    for (uint32_t s = 0; s < nst; ++s) {
        const uint32_t nu = q->stage_units[s];
        const unsigned grid = std::min<uint32_t>(nu, 2048);
#define QCS_QS(GG)                                                                                            \
    hipLaunchKernelGGL((k_qaoa_stage<GG>), dim3(grid), dim3(256), shm, q->stream, q->arena, q->d_inst,        \
                       q->d_units[s], nu, q->d_tickets + s, q->d_ang, q->p, q->d_h, q->d_ei, q->d_ej, q->d_J, \
                       q->d_partial, q->d_energy)
        switch (q->group) {
            case 4: QCS_QS(4); break;
            case 5: QCS_QS(5); break;
            default: QCS_QS(6); break;
        }
#undef QCS_QS
    }
    if (q->nlarge)
        hipLaunchKernelGGL(k_qaoa_energy, dim3(std::min<uint32_t>(q->nlarge, 1024)), dim3(256), 0, q->stream,
                           q->d_inst, q->d_large, q->nlarge, q->d_partial, q->d_energy);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? QCS_OK : set_error(QCS_EDEVICE, "qaoa launch failed: %s", hipGetErrorString(e));
}

// Build the stage plan for (tmax, group) and capture it.
static int build_plan(qcs_qaoa* q) {
    free_plan(q);
    const uint32_t G = q->group, Tm = q->tmax;
    std::vector<std::vector<QUnit>> stages;
    auto stage = [&](size_t s) -> std::vector<QUnit>& {
        if (stages.size() <= s) stages.resize(s + 1);
        return stages[s];
    };
    std::vector<uint32_t> large;
    uint32_t pslot = 0;
    // stage index of (layer, part): part 0 = A, part 1.. = B groups
    uint32_t maxgroups = 0;
    for (uint32_t i = 0; i < q->ninst; ++i)
        if (q->nq[i] > Tm) maxgroups = std::max(maxgroups, (q->nq[i] - Tm + G - 1) / G);
    const uint32_t per_layer = 1 + maxgroups;
    for (uint32_t i = 0; i < q->ninst; ++i) {
        QInst& I = q->inst_h[i];
        const uint32_t n = q->nq[i];
        if (n <= Tm) {
            I.tbits = n;
            I.pslot0 = 0;
            I.npslots = 0;
            stage(0).push_back(QUnit{i, U_SMALL, 0, 0, 0, 0});
            continue;
        }
        const uint32_t ng = (n - Tm + G - 1) / G;
        const uint32_t tb = Tm;
        I.tbits = tb;
        large.push_back(i);
        const uint32_t ntiles = 1u << (n - tb);
        const uint32_t glast = (n - Tm) - G * (ng - 1);  // width of the last group
        const uint32_t nchunks_last = uint32_t((((1ull << n) >> glast) + 255) / 256);
        I.pslot0 = pslot;
        I.npslots = nchunks_last;
        pslot += nchunks_last;
        for (uint32_t l = 0; l < q->p; ++l) {
            auto& A = stage(size_t(l) * per_layer);
            for (uint32_t t = 0; t < ntiles; ++t) A.push_back(QUnit{i, U_A, t, l, 0, 0});
            for (uint32_t gi = 0; gi < ng; ++gi) {
                const uint32_t gw = gi == ng - 1 ? glast : G;
                const uint32_t nchunks = uint32_t((((1ull << n) >> gw) + 255) / 256);
                const bool last = (l == q->p - 1) && (gi == ng - 1);
                auto& B = stage(size_t(l) * per_layer + 1 + gi);
                for (uint32_t c = 0; c < nchunks; ++c)
                    B.push_back(QUnit{i, U_B | (last ? 0x100u : 0u) | (gw << 16), c, l, tb + gi * G,
                                      last ? I.pslot0 + c : 0});
            }
        }
    }
    if (hipMemcpy(q->d_inst, q->inst_h.data(), sizeof(QInst) * q->ninst, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(QCS_EDEVICE, "instance upload failed");
    for (auto& st : stages) {
        if (st.empty()) continue;
        QUnit* d = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(QUnit) * st.size()) != hipSuccess)
            return set_error(QCS_ENOMEM, "unit table");
        if (hipMemcpy(d, st.data(), sizeof(QUnit) * st.size(), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(QCS_EDEVICE, "unit upload failed");
        q->d_units.push_back(d);
        q->stage_units.push_back(uint32_t(st.size()));
    }
    if (hipMalloc(reinterpret_cast<void**>(&q->d_tickets), sizeof(uint32_t) * (q->stage_units.size() + 1)) !=
            hipSuccess ||
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
    q->ei_h.assign(ei, ei + eo);
    q->ej_h.assign(ej, ej + eo);
    q->J_h.assign(J, J + eo);
    q->h_h.assign(h, h + ho);
    std::vector<uint8_t> e8i(eo + 1), e8j(eo + 1);
    for (uint32_t k = 0; k < eo; ++k) {
        e8i[k] = uint8_t(ei[k]);
        e8j[k] = uint8_t(ej[k]);
    }
    auto al = [](void** p, size_t b) { return hipMalloc(p, b ? b : 8) == hipSuccess; };
    if (!al(reinterpret_cast<void**>(&q->arena), sizeof(float2) * q->total_amps) ||
        !al(reinterpret_cast<void**>(&q->d_inst), sizeof(QInst) * ninst) ||
        !al(reinterpret_cast<void**>(&q->d_ei), eo + 1) || !al(reinterpret_cast<void**>(&q->d_ej), eo + 1) ||
        !al(reinterpret_cast<void**>(&q->d_J), sizeof(double) * (eo + 1)) ||
        !al(reinterpret_cast<void**>(&q->d_h), sizeof(double) * (ho + 1)) ||
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
        amps[2 * i] = tmp[i].x;
        amps[2 * i + 1] = tmp[i].y;
    }
    return QCS_OK;
}

int qcs_qaoa_info(qcs_qaoa* q, uint32_t* stages, uint64_t* amplitudes, uint32_t* tile_bits, uint32_t* group) {
    if (!q) return set_error(QCS_EINVAL, "null batch");
    if (stages) *stages = uint32_t(q->stage_units.size());
    if (amplitudes) *amplitudes = q->total_amps;
    if (tile_bits) *tile_bits = q->tmax;
    if (group) *group = q->group;
    return QCS_OK;
}

}  // extern "C"
