// program.hip - recorded gate sequences replayed as one hipGraph launch.
//
// The reference records gates between ComputeStart and ComputeEnd and
// replays them with Compute, or in reverse with adjoint matrices with
// Uncompute (QubitRegister.h:536-590). A program is such a sequence
// compiled once for one register: optionally fused, its dense matrices
// uploaded into memory the program owns, and its launches captured from
// the register's stream into a hipGraph. Running it is one graph launch on
// the same stream, so a sequence of small-state gates costs one launch
// instead of one per gate. Programs hold only gate launches (no
// reductions, measurements or host work) and no memset nodes; a program
// whose register has since exchanged its state buffer (qcs_swap_state) is
// captured again before it runs.
#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "common.h"

namespace qcs {
int fuse(const qcs_gate* gates, uint32_t n, uint32_t width, std::vector<qcs_gate>& out);
int kdense(qcs_register* r, uint32_t k, const uint32_t* qubits, const double* mat_host);

// The inverse of a gate: conjugate transpose of its matrix; the phase of
// the iSWAP kernels flips (their matrix is implied by the trait).
static qcs_gate adjoint_gate(const qcs_gate& g) {
    qcs_gate a = g;
    const int D = 1 << g.k;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            a.mat[2 * (i * D + j)] = g.mat[2 * (j * D + i)];
            a.mat[2 * (i * D + j) + 1] = -g.mat[2 * (j * D + i) + 1];
        }
    const bool is = g.traits & QCS_TRAIT_ISWAP, isd = g.traits & QCS_TRAIT_ISWAPDAG;
    a.traits &= ~uint32_t(QCS_TRAIT_ISWAP | QCS_TRAIT_ISWAPDAG);
    if (is) a.traits |= QCS_TRAIT_ISWAPDAG;
    if (isd) a.traits |= QCS_TRAIT_ISWAP;
    return a;
}
// Upper bound of the parameter bytes one block's launch stages (each
// staging is rounded up to 256 B): a dense block takes at most one staging
// slot; a generic 2- or 3-qubit gate may run k_nq_split, which stages its
// matrix as D*D real parts then D*D imaginary parts; every other launch
// passes its parameters as kernel arguments.
static size_t param_bound(const qcs_register* r, const qcs_gate& g) {
    if ((g.flags & QCS_GATE_DENSE) || g.k > 3) return (r->mat_bytes + 255) & ~size_t(255);
    if (g.k < 2) return 0;
    const size_t D = size_t(1) << g.k;
    return (2 * D * D * sizeof(double) + 255) & ~size_t(255);
}
}  // namespace qcs

using namespace qcs;

struct qcs_program {
    qcs_register* r = nullptr;
    std::vector<qcs_gate> blocks;  // the (fused) sequence, kept for re-capture
    size_t param_bytes = 0;        // upper bound of what the launches stage
    void* psi = nullptr;  // state buffer the graph was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* params = nullptr;
    uint32_t launches = 0;
};

// Captures p->blocks on p->r's stream into a new graph (dropping any older
// one) and uploads the dense matrices the launches read.
static int capture(qcs_program* p) {
    qcs_register* r = p->r;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    ParamSink sink;
    std::vector<char> host(std::max<size_t>(p->param_bytes, 256));
    sink.host = host.data();
    sink.cap = host.size();
    if (!p->params && hipMalloc(&p->params, sink.cap) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(QCS_ENOMEM, "program parameter memory");
    }
    sink.dev = static_cast<char*>(p->params);
    if (hipStreamSynchronize(r->stream) != hipSuccess ||
        hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return set_error(QCS_EDEVICE, "stream capture failed");
    r->sink = &sink;
    const uint64_t l0 = r->launches;
    int e = QCS_OK;
    for (size_t i = 0; i < p->blocks.size() && !e; ++i) {
        const qcs_gate& g = p->blocks[i];
        e = (g.flags & QCS_GATE_DENSE) || g.k > 3
                ? kdense(r, g.k, g.qubits, g.mat)
                : qcs_apply_gate(r, g.k, g.traits, g.qubits[0], g.qubits[1], g.qubits[2], g.mat);
    }
    r->sink = nullptr;
    p->launches = uint32_t(r->launches - l0);
    r->launches = l0;
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(r->stream, &graph);
    p->graph = graph;
    if (e) return e;
    if (ce != hipSuccess) return set_error(QCS_EDEVICE, "end of stream capture failed");
    if ((sink.used && hipMemcpy(p->params, host.data(), sink.used, hipMemcpyHostToDevice) != hipSuccess) ||
        hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
        return set_error(QCS_EDEVICE, "graph instantiation failed");
    p->psi = r->psi;
    return QCS_OK;
}

extern "C" {

int qcs_program_destroy(qcs_program* p) {
    if (!p) return QCS_OK;
    if (p->r && p->r->stream) (void)hipStreamSynchronize(p->r->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->params) (void)hipFree(p->params);
    delete p;
    return QCS_OK;
}

int qcs_program_create(qcs_register* r, const qcs_gate* gates, uint32_t n, uint32_t fuse_width, int adjoint,
                       qcs_program** out) {
    if (!r || !out || (!gates && n)) return set_error(QCS_EINVAL, "null argument");
    if (fuse_width == 1 || (fuse_width > 5 && fuse_width != QCS_FUSE_AUTO))
        return set_error(QCS_EINVAL, "fusion width must be 0, 2..5 or QCS_FUSE_AUTO");
    *out = nullptr;
    std::vector<qcs_gate> seq(gates, gates + n);
    for (uint32_t i = 0; i < n; ++i) {
        if (seq[i].k < 1 || seq[i].k > 5) return set_error(QCS_EINVAL, "gate %u: arity must be 1..5", i);
        for (uint32_t j = 0; j < seq[i].k; ++j)
            if (seq[i].qubits[j] >= r->n) return set_error(QCS_EINVAL, "gate %u: qubit out of range", i);
    }
    if (adjoint) {
        std::reverse(seq.begin(), seq.end());
        for (auto& g : seq) g = adjoint_gate(g);
    }
    auto* p = new qcs_program;
    p->r = r;
    if (fuse_width >= 2) {
        uint32_t picked = fuse_width;
        const int e = fuse_width == QCS_FUSE_AUTO ? pick_fusion(r, seq.data(), uint32_t(seq.size()), p->blocks, &picked)
                                                  : fuse(seq.data(), uint32_t(seq.size()), fuse_width, p->blocks);
        if (e) {
            delete p;
            return e;
        }
    } else {
        p->blocks.swap(seq);
    }
    for (const auto& g : p->blocks) p->param_bytes += param_bound(r, g);
    if (const int e = capture(p)) {
        qcs_program_destroy(p);
        return e;
    }
    *out = p;
    return QCS_OK;
}

int qcs_program_run(qcs_register* r, qcs_program* p) {
    if (!r || !p) return set_error(QCS_EINVAL, "null argument");
    if (p->r != r) return set_error(QCS_EINVAL, "program was compiled for another register");
    // qcs_swap_state exchanges state buffers between registers: capture
    // again for the buffer the register holds now
    if (p->psi != r->psi)
        if (const int e = capture(p)) return e;
    if (hipGraphLaunch(p->exec, r->stream) != hipSuccess) return set_error(QCS_EDEVICE, "graph launch failed");
    r->launches += p->launches;
    return QCS_OK;
}

int qcs_program_info(const qcs_program* p, uint32_t* launches) {
    if (!p) return set_error(QCS_EINVAL, "null program");
    if (launches) *launches = p->launches;
    return QCS_OK;
}

}  // extern "C"
