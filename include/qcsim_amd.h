/*
 * qcsim_amd.h - C-ABI of the MI355X (gfx950) statevector engine.
 *
 * This is the drop-in boundary under the C++ facade in include/QCSim/. Each
 * entry point replaces one piece of the reference's QC::QubitRegister /
 * QubitRegisterCalculator path (reference files cited per function). Only
 * plain C types cross the boundary: no HIP, no torch, no Eigen.
 *
 * Conventions (same as the reference, QubitRegister.h:7, SimpleGates.h:5):
 *   - qubit 0 is the least-significant bit of a basis-state index;
 *   - amplitudes cross the boundary as interleaved (re, im) doubles;
 *   - gate matrices are row-major 2^k x 2^k complex doubles (interleaved),
 *     local bit 0 = `qubit`, bit 1 = `c1`, bit 2 = `c2`
 *     (QubitRegister.h:444-471, SimpleGates.h:329,430).
 *
 * Every function returns QCS_OK (0) or an error code; qcs_last_error()
 * gives a thread-local message for the last failure on the calling thread.
 * Calls on different registers may run concurrently from different host
 * threads (one HIP stream per register); calls on one register must be
 * serialised by the caller, as with the reference's QubitRegister.
 */
#ifndef QCSIM_AMD_H
#define QCSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define QCS_OK      0
#define QCS_EINVAL  1 /* bad argument (CheckQubits, QubitRegister.h:677-690) */
#define QCS_ENOMEM  2 /* device allocation failed                          */
#define QCS_EDEVICE 3 /* HIP runtime / launch failure                      */

/* ---- storage precision ----------------------------------------------- */
#define QCS_COMPLEX128 0 /* std::complex<double>, the reference's only type */
#define QCS_COMPLEX64  1 /* complex<float> storage, FP64 reductions          */

/* ---- gate traits: the QuantumGate virtuals, SimpleGates.h:27-60 ------- */
#define QCS_TRAIT_CONTROLLED   0x01u /* isControlled()                    */
#define QCS_TRAIT_CTRL_Q1      0x02u /* isControlQubit(1) (3-qubit gates) */
#define QCS_TRAIT_DIAGONAL     0x04u /* isDiagonal()                      */
#define QCS_TRAIT_ANTIDIAGONAL 0x08u /* isAntidiagonal()                  */
#define QCS_TRAIT_SWAP         0x10u /* isSwapGate()                      */
#define QCS_TRAIT_ISWAP        0x20u /* IsISwapGate()                     */
#define QCS_TRAIT_ISWAPDAG     0x40u /* IsISwapDagGate()                  */

typedef struct qcs_register qcs_register;

/* One gate of a circuit batch (qcs_apply_circuit, qcs_fuse). */
#define QCS_GATE_DENSE 0x1u /* flags: dense block on qubits[0..k)           */
typedef struct qcs_gate {
    uint32_t k;        /* 1..3: reference ApplyGate; 4..5: dense block      */
    uint32_t traits;   /* QCS_TRAIT_* (ignored for dense blocks)             */
    uint32_t qubits[5];/* ApplyGate: qubit, c1, c2; dense: bit i -> qubits[i]*/
    uint32_t flags;    /* QCS_GATE_DENSE; k > 3 is always dense              */
    double   mat[2 * 32 * 32]; /* row-major interleaved, first 2*4^k used   */
} qcs_gate;

/* ---- lifecycle -------------------------------------------------------- */

/* QubitRegister(N) ctor, QubitRegister.h:17-37: allocates 2^n amplitudes in
 * HBM (in place, no results buffer) and sets |0...0>. */
int qcs_create(uint32_t nqubits, int precision, qcs_register** out);
int qcs_destroy(qcs_register* r);
uint32_t qcs_get_nqubits(const qcs_register* r);  /* getNrQubits, :59 */
int qcs_get_precision(const qcs_register* r);
/* Device copy of the state (SaveState/RestoreState/Clone, :600-616,662-674).
 * Registers must have equal size and precision. */
int qcs_copy_state(qcs_register* dst, const qcs_register* src);
/* O(1) exchange of the two device buffers (setRegisterStorageFastNoNormalize
 * and RestoreStateDestructive swap semantics, :521-524, :611-616). */
int qcs_swap_state(qcs_register* a, qcs_register* b);

/* ---- state access (QubitRegister.h:62-130, 507-524) ------------------- */

/* Upload / download all 2^n amplitudes (interleaved re,im doubles). */
int qcs_set_state(qcs_register* r, const double* amps, uint64_t n_amps);
int qcs_get_state(qcs_register* r, double* amps, uint64_t n_amps);
/* getBasisStateAmplitude / setRawAmplitude: index out of range is a silent
 * no-op / returns 0, as in the reference (:62-66, :112-117). */
int qcs_get_amplitude(qcs_register* r, uint64_t idx, double* re_im);
int qcs_set_amplitude(qcs_register* r, uint64_t idx, double re, double im);
/* setToBasisState (:74-80), out-of-range index is a silent no-op. */
int qcs_set_basis_state(qcs_register* r, uint64_t idx);
/* setConstant: Clear() is (0,0), setToEqualSuperposition (:106-109) is
 * (1/sqrt(2^n), 0). */
int qcs_fill(qcs_register* r, double re, double im);
/* psi *= (re + i im) for every amplitude. */
int qcs_scale(qcs_register* r, double re, double im);

/* ---- gates ------------------------------------------------------------ */

/* QubitRegister::ApplyGate(gate, qubit, c1, c2), QubitRegister.h:434-486.
 * k = gate.getQubitsNumber() in 1..3, traits = the gate's trait bits, mat =
 * gate.getRawOperatorMatrix(). Dispatch follows the reference exactly
 * (QubitRegisterCalculator.h:39-227, 599-639): the traits choose the kernel
 * and the kernel reads only the matrix entries the reference reads. Qubit
 * checks and error cases are those of CheckQubits (:677-690). */
int qcs_apply_gate(qcs_register* r, uint32_t k, uint32_t traits,
                   uint32_t qubit, uint32_t c1, uint32_t c2,
                   const double* mat);

/* Dense k-qubit block (1 <= k <= 5) on arbitrary distinct qubits; local bit
 * i of the matrix index is qubits[i]. Produced by the fusion pass; for k<=3
 * it is the reference's AppliedGate generic path (SimpleGates.h:439-578). */
int qcs_apply_dense(qcs_register* r, uint32_t k, const uint32_t* qubits,
                    const double* mat);

/* A batch of gates in order; fuse_width 0/1 applies them one by one, 2..5
 * runs the fusion pass first (blocks of at most fuse_width qubits), and
 * QCS_FUSE_AUTO picks the width (0 or 2..5) whose blocks cost least under
 * the fusion cost table (qcs_autotune_fusion, measured on first use for
 * this precision and qubit count). No counterpart in the reference, which
 * applies gates one by one (QubitRegister.h:488-492). */
#define QCS_FUSE_AUTO 0xFFFFFFFFu
int qcs_apply_circuit(qcs_register* r, const qcs_gate* gates, uint32_t n,
                      uint32_t fuse_width);
/* The width the last qcs_apply_circuit on r used (0: unfused). */
uint32_t qcs_last_fuse_width(const qcs_register* r);

/* The fusion pass alone (host only, no device): rewrites `gates` into at
 * most n blocks of <= width qubits whose product equals the sequence, each
 * gate taken with the matrix its traits make the reference apply. Blocks of
 * a single gate keep that gate unchanged; merged blocks are dense
 * (QCS_GATE_DENSE, qubits ascending). No counterpart in the reference. */
int qcs_fuse(const qcs_gate* gates, uint32_t n, uint32_t width,
             qcs_gate* out, uint32_t* n_out);

/* Recorded-gate replay as one graph launch: replaces the gate-by-gate loops
   of QubitRegister::Compute / Uncompute (QubitRegister.h:553-590). A
   program is compiled for one register: the sequence is fused to blocks of
   <= fuse_width qubits (0: no fusion, else 2..5); with adjoint != 0 it is
   the inverse sequence (gates reversed, matrices conjugate-transposed, as
   Uncompute applies them); its matrices are uploaded once into memory the
   program owns and its launches are captured into a hipGraph. qcs_program_run
   replays the graph on the register's stream. */
typedef struct qcs_program qcs_program;
int qcs_program_create(qcs_register* r, const qcs_gate* gates, uint32_t n,
                       uint32_t fuse_width, int adjoint, qcs_program** out);
int qcs_program_run(qcs_register* r, qcs_program* p);
int qcs_program_info(const qcs_program* p, uint32_t* launches);
int qcs_program_destroy(qcs_program* p);

/* Diagonal operator on the whole register: psi[s] *= diag[s]
 * (ApplyOperatorMatrix with a diagonal matrix, QubitRegister.h:499-505).
 * The host diagonal streams through the register's transfer buffer (at most
 * 64 MiB, allocated on first use and kept) in chunks. */
int qcs_apply_diagonal(qcs_register* r, const double* diag, uint64_t n);

/* Dense operator on the whole register: psi <- M psi, M a row-major
 * dim x dim complex matrix (interleaved re, im) in host memory, dim = 2^n
 * (ApplyOperatorMatrix, QubitRegister.h:499-505; the reference's callers
 * are the oracle-based algorithms, GroverAlgorithm.h:97,111,
 * BernsteinVazirani.h:85, DeutschJozsa.h:126, NQubitsQuantumGate.h:22).
 * M streams through the transfer buffer in panels of whole rows; each row
 * is one FP64 dot product on the device. Registers of up to 21 qubits
 * (complex128: 2 * 16 * 2^n bytes must fit the 64 MiB buffer). */
int qcs_apply_matrix(qcs_register* r, const double* mat, uint64_t dim);

/* ---- QFT (QuantumFourierTransform.h:35-87, QubitsSwapper.h:23-33) ----- */

/* QFT (inverse=0) or IQFT (inverse=1) on qubits [sq, eq]; eq is clamped as
 * QuantumSubAlgorithmOnSubregister does (QuantumAlgorithm.h:228-229).
 * Result equals the reference's H / controlled-phase / SWAP sequence. */
int qcs_qft(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap,
            int inverse);

/* ---- measurement (uniform u in [0,1) supplied by the caller; the
 *      reference draws it from its mt19937_64 and uses prob = 1 - u) ---- */

/* MeasureAll, QubitRegister.h:169-195: first i with prob <= cdf(i), default
 * 2^n-1; collapses to |i>. */
int qcs_measure_all(qcs_register* r, double u, uint64_t* outcome);
/* Measure(first, second) / MeasureQubit (:198-224, Calc:948-1059,
 * 1124-1225): returns the measured sub-register value and collapses. */
int qcs_measure(qcs_register* r, uint32_t first, uint32_t second, double u,
                uint64_t* outcome);
/* MeasureNoCollapse() (:619-642, default 0) when all != 0, else the
 * protected MeasureNoCollapse(first, second) (:705-713, Calc:1061-1086,
 * 1227-1254). The state is left untouched. */
int qcs_measure_no_collapse(qcs_register* r, int all, uint32_t first,
                            uint32_t second, double u, uint64_t* outcome);
/* RepeatedMeasure[(first, second)] sampling, QubitRegister.h:227-429: one
 * outcome per supplied uniform, from the CDF truncated once it exceeds
 * 1 - DBL_EPSILON (lower_bound semantics, so an outcome may equal the
 * truncated length). With all != 0 outcomes are full basis-state indices;
 * otherwise (state & mask) >> first. nshots == 1 follows the reference's
 * single-shot shortcut (MeasureNoCollapse). */
int qcs_sample(qcs_register* r, int all, uint32_t first, uint32_t second,
               uint64_t nshots, const double* u, uint64_t* outcomes);
/* qcs_sample without allocation: the same outcomes, with the per-shot
 * device buffers in a fixed 1 MiB array per GPU (held until the process
 * ends) instead of per-call allocations. No counterpart in the reference:
 * for callers that cannot allocate (code under stream capture, serving
 * loops, harnesses that forbid allocation calls). At most 65536 shots per
 * call (QCS_EINVAL beyond). Calls on registers of the same GPU share the
 * array: overlapping calls on one GPU race and must be serialised. */
int qcs_sample_noalloc(qcs_register* r, int all, uint32_t first,
                       uint32_t second, uint64_t nshots, const double* u,
                       uint64_t* outcomes);

/* GetQubitProbability(q), QubitRegister.h:592-598, Calc:1088-1122. FP64. */
int qcs_qubit_probability(qcs_register* r, uint32_t q, double* p);
/* sum |psi|^2 in FP64 (the square of Eigen's norm(), :126). */
int qcs_norm2(qcs_register* r, double* n2);
/* Normalize, :124-130 (no-op when the norm is below 1e-20). */
int qcs_normalize(qcs_register* r);
/* <a|b> = sum conj(a_s) b_s in FP64 (stateFidelity :527-534 and
 * ExpectationValue :646-660). a and b may differ in precision. */
int qcs_inner(qcs_register* a, qcs_register* b, double* re_im);
/* <r|v> = sum conj(r_s) v_s for a host vector v of 2^n interleaved
 * doubles, FP64 (stateFidelity, QubitRegister.h:527-534, is its |.|^2).
 * v streams through the transfer buffer; no second register is made. */
int qcs_inner_host(qcs_register* r, const double* amps, uint64_t n_amps,
                   double* re_im);
/* Exact expectation values of Pauli strings, FP64: out[t] = <psi|P_t|psi>
 * for P_t = i^{|x_t & z_t|} X^{x_t} Z^{z_t} (qubit q: X if bit q of x_t
 * only, Z if bit q of z_t only, Y if both, identity if neither). This is
 * the exact value of what VQE.h:116-142 estimates from 10000 samples after
 * its change of basis (PauliString.h's Z default included: set its z bit),
 * and the Pauli-string special case of ExpectationValue
 * (QubitRegister.h:646-660). Terms sharing an x mask are evaluated in one
 * read-only pass over the state (up to 16 per pass). */
int qcs_pauli_expectation(qcs_register* r, uint32_t nterms,
                          const uint64_t* xmask, const uint64_t* zmask,
                          double* out);
/* max_s |a_s - b_s| and <a|b>, for error reports between precisions. */
int qcs_compare(qcs_register* a, qcs_register* b, double* max_abs,
                double* inner_re_im);
/* sum_s |psi_s|^2 * diag_energy(s) for the Ising energy
 * E(s) = -(sum_i h_i z_i + sum_(i,j) J_ij z_i z_j), z = +1 for bit 0
 * (IsingModel.h:106-119, 203-206): the exact expectation that the
 * reference estimates by sampling (IsingModel.h:569-578). FP64. */
int qcs_ising_energy(qcs_register* r, uint32_t n_edges, const uint32_t* ei,
                     const uint32_t* ej, const double* J, const double* h,
                     double* energy);

/* ---- batched QAOA (caller of the path: IsingModel.h:215-665) ------------- */

/* A batch of weighted Ising instances of uneven size, each evaluated with
 * QAOAIsingSubalgorithm::Exec's circuit (IsingModel.h:557-567: H on all,
 * then per layer the cost operator of :491-531 and Rx(2 beta) on all qubits,
 * :533-554) on a complex64 state, and its exact FP64 energy
 * sum_s |psi_s|^2 E(s), E(s) = -(sum h z + sum J z z) (:106-119, 203-206),
 * where the reference samples 100000 shots (:569-578). Instance i has
 * nqubits[i] qubits and nedges[i] edges (ei, ej, J concatenated; h
 * concatenated per instance). One persistent launch per stage; the stages of
 * an evaluation are captured in one hipGraph. */
typedef struct qcs_qaoa qcs_qaoa;
int qcs_qaoa_create(uint32_t ninst, const uint32_t* nqubits, const uint32_t* nedges,
                    const uint32_t* ei, const uint32_t* ej, const double* J,
                    const double* h, uint32_t p, qcs_qaoa** out);
int qcs_qaoa_destroy(qcs_qaoa* q);
/* tile_bits (6..12): instances up to 2^tile_bits amplitudes stay in LDS;
 * group (4..6): high qubits per register pass; use_graph: replay a hipGraph. */
int qcs_qaoa_configure(qcs_qaoa* q, uint32_t tile_bits, uint32_t group, int use_graph);
/* gamma, beta: ninst x p (row per instance); energy: ninst. */
int qcs_qaoa_evaluate(qcs_qaoa* q, const double* gamma, const double* beta, double* energy);
int qcs_qaoa_last_ms(qcs_qaoa* q, float* ms);
/* The final state of one instance as interleaved doubles (2^n of them). */
int qcs_qaoa_get_state(qcs_qaoa* q, uint32_t inst, double* amps);
int qcs_qaoa_info(qcs_qaoa* q, uint32_t* stages, uint64_t* amplitudes,
                  uint32_t* tile_bits, uint32_t* group);
/* The passes over the state that one evaluation of an n-qubit instance
 * makes under (tile_bits, group) at depth p, in order (host only; no
 * device or batch needed): 5 words per pass, (kind, width, mode, qlo,
 * layer). kind 0: the whole evaluation in LDS (n <= tile_bits); 1: tile
 * pass on qubits 0..tile_bits-1, mode 0 = cost diagonal of `layer` then its
 * Rx, 1 = Rx only; 2: mixer pass on qubits qlo..qlo+width-1, mode 0 = Rx of
 * `layer`, 1 = Rx then the energy, 2 = Rx of `layer`, the cost diagonal of
 * layer + 1, Rx of layer + 1. The first pass starts from |+...+>.
 * Writes at most cap passes; *npasses receives the full count. */
int qcs_qaoa_plan(uint32_t n, uint32_t tile_bits, uint32_t group, uint32_t p, uint32_t* out,
                  uint32_t cap, uint32_t* npasses);

/* ---- batched VQE energies (caller of the path: VQE.h:78-212) ----------- */

/* Many registers of 1..12 qubits of uneven size, each prepared by
 * VQE::PauliStringVQE::Ansatz (VQE.h:78-114: Ry / Rz layers and alternating
 * CNOT ladders from |0...0>, no gates with fewer parameters than qubits)
 * and scored with the exact FP64 energy sum_t c_t <psi|P_t|psi> of its
 * Pauli-string Hamiltonian, the value VariationalQuantumEigensolver::
 * EstimateEnergy (VQE.h:204-212) estimates from sampled parities. Instance
 * i has nqubits[i] qubits and nterms[i] terms (x masks, z masks and
 * coefficients concatenated; masks as in qcs_pauli_expectation). One launch
 * per size class (<= 8, <= 10, <= 12 qubits), complex128 state in LDS. */
typedef struct qcs_vqe qcs_vqe;
int qcs_vqe_create(uint32_t ninst, const uint32_t* nqubits, const uint32_t* nterms,
                   const uint64_t* xmask, const uint64_t* zmask, const double* coef,
                   qcs_vqe** out);
int qcs_vqe_destroy(qcs_vqe* v);
/* nparams[i] parameters per instance, concatenated in params; energy: ninst. */
int qcs_vqe_evaluate(qcs_vqe* v, const uint32_t* nparams, const double* params,
                     double* energy);
int qcs_vqe_last_ms(qcs_vqe* v, float* ms);
/* The final state of one instance after the last evaluation (2^n
 * interleaved doubles). */
int qcs_vqe_get_state(qcs_vqe* v, uint32_t inst, double* amps);

/* ---- MLP surrogate of QAOA energies (no counterpart in the reference) ---- */

/* in -> hidden (x layers, ReLU) -> 1; BF16 weights and activations, FP32
 * accumulation, FP32 master weights and gradients, Adam, MSE loss. The
 * parameters are one flat FP32 vector: per layer W (in x out, row-major)
 * then b (out). */
typedef struct qcs_mlp qcs_mlp;
int qcs_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t layers, uint64_t seed, qcs_mlp** out);
int qcs_mlp_destroy(qcs_mlp* n);
uint64_t qcs_mlp_num_params(const qcs_mlp* n);
int qcs_mlp_get_params(qcs_mlp* n, float* params);
int qcs_mlp_set_params(qcs_mlp* n, const float* params);
/* x: batch x in_dim (FP32, rounded to BF16 on upload); y: batch. */
int qcs_mlp_forward(qcs_mlp* n, const float* x, uint32_t batch, float* y);
/* Loss (mean squared error against t) and the flat FP32 gradient. */
int qcs_mlp_gradients(qcs_mlp* n, const float* x, const float* t, uint32_t batch, float* grad, float* loss);
/* Shuffled mini-batch Adam over a device-resident copy of (X, Y). */
int qcs_mlp_train(qcs_mlp* n, const float* X, const float* Y, uint64_t rows, uint32_t batch,
                  uint32_t epochs, float lr, uint64_t seed, float* loss_per_epoch);

/* ---- autotuning (no counterpart in the reference) ------------------------ */

/* Times every launch variant of the streaming gate kernels on this register
 * (identity gates: the state keeps its values) for each qubit, and keeps the
 * fastest per (precision, qubit count, qubit) for all later launches in the
 * process. gbps, if not NULL, receives nqubits * qcs_autotune_variants()
 * algorithmic GB/s figures. */
int qcs_autotune(qcs_register* r, uint32_t reps, double* gbps);
int qcs_autotune_variants(void);
/* The fusion cost table of QCS_FUSE_AUTO: times a generic 1-qubit gate and
 * dense 2..5-qubit blocks (identity matrices: the state keeps its values)
 * on this register, mean over low, high and spread qubits, and keeps the
 * figures for its precision and qubit count. ms, if not NULL, receives 6
 * values: ms[0] = ms[1] = 1-qubit gate, ms[k] = dense k-qubit block. */
int qcs_autotune_fusion(qcs_register* r, uint32_t reps, double* ms);
/* Preset / read that table (6 values as above; zeros: not measured). */
int qcs_set_fusion_costs(int precision, uint32_t n, const double* ms);
int qcs_get_fusion_costs(int precision, uint32_t n, double* ms);
/* Read / preset one entry of that table (vid = -1: untuned default), so a
 * tuning result saved by one process can be reused by the next. */
int qcs_set_tuned_variant(int precision, uint32_t n, uint32_t q, int vid);
int qcs_get_tuned_variant(int precision, uint32_t n, uint32_t q);

/* ---- stream / timing --------------------------------------------------- */
int qcs_synchronize(qcs_register* r);
/* HIP events on the register's stream around the enclosed calls. */
int qcs_timer_begin(qcs_register* r);
int qcs_timer_end(qcs_register* r, float* ms);
/* Number of kernels launched on the register since creation. */
uint64_t qcs_launch_count(const qcs_register* r);

const char* qcs_last_error(void);
const char* qcs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QCSIM_AMD_H */
