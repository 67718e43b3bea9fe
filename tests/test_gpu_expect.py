"""Pauli-string expectations, dense and diagonal operators, and host vectors
streamed through the device, against brute force and the oracle.

- Pauli strings: <psi|P|psi> with P built by Kronecker products (n <= 9),
  and the exact value of VQE::PauliStringVQE::EstimateEnergy's estimator
  (VQE.h:116-142: basis change Ry(-pi/2) / Rx(pi/2) on X / Y qubits, then
  the parity of all qubits) computed on the oracle's state at 18 qubits.
- ExpectationValue (QubitRegister.h:646-660): Pauli-product gate lists take
  the one-pass path; they must equal the copy path and the reference's own
  answers (Tests.cpp:665-723).
- ApplyOperatorMatrix (QubitRegister.h:499-505): dense and diagonal
  operators against numpy's product.
- stateFidelity / complex64 set and get at 23 qubits: two transfer chunks.
"""
import math

import numpy as np
import pytest

from qcsim_amd import QubitRegister
from qcsim_amd import gates as G
from qcsim_amd import pauli as P
from refmath import dense_pauli, pauli_evs, random_state

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 2e-6}


def random_terms(rng, n, count):
    xs, zs = [], []
    pool_x = [0, 1, (1 << n) - 1] + [int(v) for v in rng.integers(0, 1 << n, size=5)]
    for _ in range(count):
        xs.append(int(pool_x[int(rng.integers(len(pool_x)))]))
        zs.append(int(rng.integers(0, 1 << n)))
    return xs, zs


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 5, 9])
def test_pauli_strings_match_dense_operator(n, precision):
    rng = np.random.default_rng(100 + n)
    psi = random_state(rng, n)
    # 80 terms over 8 x masks: groups of more than 32 terms take several passes
    xs, zs = random_terms(rng, n, 80)
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    got = r.pauli_expectation(xs, zs)
    ref_state = r.getRegisterStorage()  # the stored (possibly rounded) state
    want = np.array([np.vdot(ref_state, dense_pauli(x, z, n) @ ref_state).real for x, z in zip(xs, zs)])
    err = float(np.max(np.abs(got - want)))
    print(f"n={n} {precision}: 80 Pauli terms, max error {err:.2e}")
    assert err < TOL[precision]
    # a single Y: <0|Y|0> = 0 and <+i|Y|+i> = 1 (S H |0> = |+i>)
    r2 = QubitRegister(1)
    r2.ApplyGate(G.HadamardGate(), 0)
    r2.ApplyGate(G.SGate(), 0)
    assert abs(r2.pauli_expectation([1], [1])[0] - 1) < 1e-15
    assert abs(r2.pauli_expectation([0], [1])[0]) < 1e-15
    r.close()
    r2.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", [10, 11, 13])
def test_pauli_strings_across_pass_shapes(n, precision):
    """Sizes whose passes take 1, 2, 4 and 8 loop iterations per work-item
    (with and without the 4- and 2-iteration trips), 24 terms per X mask
    (a pass of 16 and one of 8). Reference: the row form on the stored
    state."""
    rng = np.random.default_rng(40 + n)
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(random_state(rng, n))
    state = r.getRegisterStorage()
    xs, zs = [], []
    for x in (0, 1, 1 << (n - 1), int(rng.integers(1, 1 << n))):
        for _ in range(24):
            xs.append(x)
            zs.append(int(rng.integers(0, 1 << n)))
    err = float(np.max(np.abs(r.pauli_expectation(xs, zs) - pauli_evs(state, xs, zs))))
    print(f"n={n} {precision}: {len(xs)} Pauli terms, max error {err:.2e}")
    assert err < TOL[precision]
    r.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_pauli_strings_at_24_qubits(precision):
    """At 24 qubits a pass loops over 16 (X != 0) or 32 (X = 0) grid strides
    in Gray order, flipping the terms' signs from a per-bit table; 18 terms
    per X mask take a full pass of 16 and one of 2. Reference: the row form
    of each string on the stored state."""
    n = 24
    rng = np.random.default_rng(24)
    psi = random_state(rng, n)
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(psi)
    state = r.getRegisterStorage()
    xs, zs = [], []
    for x in (0, (1 << 23) | 5, 1):
        for _ in range(18):
            xs.append(x)
            zs.append(int(rng.integers(0, 1 << n)))
    zs[0] = 0  # <I> = 1
    got = r.pauli_expectation(xs, zs)
    want = pauli_evs(state, xs, zs)
    err = float(np.max(np.abs(got - want)))
    print(f"24 qubits {precision}: {len(xs)} Pauli terms, max error {err:.2e}")
    assert abs(got[0] - float(np.vdot(state, state).real)) < 1e-12
    assert err < TOL[precision]
    r.close()


def test_pauli_masks_beyond_register_rejected():
    r = QubitRegister(4)
    with pytest.raises(ValueError, match="beyond the register"):
        r.pauli_expectation([1 << 4], [0])
    assert r.pauli_expectation([], []).shape == (0,)
    r.close()


def ansatz_state(n, seed):
    """A VQE-style ansatz (Ry, Rz on every qubit, then a CNOT ladder,
    twice) applied gate by gate; returns the gate list."""
    rng = np.random.default_rng(seed)
    seq = []
    for layer in range(2):
        for q in range(n):
            seq.append((G.RyGate(float(rng.uniform(0, math.pi))), [q]))
            seq.append((G.RzGate(float(rng.uniform(0, 2 * math.pi))), [q]))
        for q in range(n - 1):
            seq.append((G.CNOTGate(), [q + 1, q]))
    return seq


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_vqe_term_energy_is_the_estimators_exact_value(oracle_lib, precision):
    """Per term: the oracle applies the reference's change of basis to the
    ansatz state and sums |phi_s|^2 (-1)^popcount(s); the engine computes the
    same number from the un-rotated state in one FP64 pass."""
    from oracle.oracle import OracleRegister

    n = 18
    seq = ansatz_state(n, 7)
    reg = QubitRegister(n, precision=precision)
    for g, qs in seq:
        reg.ApplyGate(g, *qs)
    rng = np.random.default_rng(8)
    parity = np.array([bin(s).count("1") & 1 for s in range(1 << n)], dtype=np.int8)
    sign = 1.0 - 2.0 * parity
    worst = 0.0
    for t in range(6):
        term = P.PauliString(n)
        for q in range(n):
            term.setOperatorForQubit(q, int(rng.integers(3)))
        term.setCoefficient(float(rng.normal()))
        ref = OracleRegister(n)
        for g, qs in seq:
            ref.apply_gate(g, *qs)
        for q in range(n):  # VQE.h:33-44
            op = term.getOperatorForQubit(q)
            if op == P.OP_X:
                ref.apply_gate(G.RyGate(-0.5 * math.pi), q)
            elif op == P.OP_Y:
                ref.apply_gate(G.RxGate(0.5 * math.pi), q)
        phi = ref.state()
        ref.close()
        want = term.getCoefficient() * float(np.sum(np.abs(phi) ** 2 * sign))
        got = P.vqe_term_energy(reg, term)
        worst = max(worst, abs(got - want))
    print(f"VQE term energies at {n} qubits ({precision}): max error {worst:.2e}")
    assert worst < (1e-12 if precision == "c128" else 2e-6)
    reg.close()


def test_expectation_value_pauli_path_matches_copy_path():
    n = 10
    rng = np.random.default_rng(12)
    psi = random_state(rng, n)
    r = QubitRegister(n)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    mats = {"X": G.PauliXGate(), "Y": G.PauliYGate(), "Z": G.PauliZGate()}
    for trial in range(5):
        gates = []
        for _ in range(7):
            q = int(rng.integers(n))
            m = mats["XYZ"[int(rng.integers(3))]].getRawOperatorMatrix() * complex(rng.normal(), rng.normal())
            gates.append(G.AppliedGate(m, q))
        assert P.gates_as_pauli(gates, n) is not None
        fast = r.ExpectationValue(gates)
        full = np.eye(1 << n, dtype=np.complex128)
        for g in gates:  # dense product of the list, later gates on the left
            k = np.kron(np.eye(1 << (n - 1 - g.q1)), np.kron(g.getRawOperatorMatrix(), np.eye(1 << g.q1)))
            full = k @ full
        want = np.vdot(psi, full @ psi)
        assert abs(fast - want) < 1e-12, (trial, fast, want)
    # a non-Pauli gate in the list takes the copy path
    gates = [G.AppliedGate(G.HadamardGate().getRawOperatorMatrix(), 3)]
    assert P.gates_as_pauli(gates, n) is None
    h = np.kron(np.eye(1 << (n - 4)), np.kron(G.HadamardGate().getRawOperatorMatrix(), np.eye(8)))
    assert abs(r.ExpectationValue(gates) - np.vdot(psi, h @ psi)) < 1e-12
    assert np.max(np.abs(r.getRegisterStorage() - psi)) < 1e-15  # state restored
    r.close()


def test_expectation_value_known_answers():
    # Tests.cpp:665-723
    z0 = [G.AppliedGate(G.PauliZGate().getRawOperatorMatrix(), 0)]
    r = QubitRegister(3)
    assert abs(r.ExpectationValue([]) - 1) < 1e-6
    assert abs(r.ExpectationValue(z0) - 1) < 1e-6
    r.ApplyGate(G.PauliXGate(), 0)
    assert abs(r.ExpectationValue(z0) + 1) < 1e-6
    r.ApplyGate(G.HadamardGate(), 0)
    assert abs(r.ExpectationValue(z0)) < 1e-6
    r.setToBasisState(0)
    r.ApplyGate(G.HadamardGate(), 1)
    assert abs(r.ExpectationValue([G.AppliedGate(G.PauliZGate().getRawOperatorMatrix(), 1)])) < 1e-6
    r.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 4, 9])
def test_dense_operator_matrix(n, precision):
    rng = np.random.default_rng(200 + n)
    psi = random_state(rng, n)
    N = 1 << n
    m = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / math.sqrt(N)  # not unitary
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    start = r.getRegisterStorage()
    r.ApplyOperatorMatrix(m)
    err = float(np.max(np.abs(r.getRegisterStorage() - m @ start)))
    print(f"dense operator n={n} {precision}: max error {err:.2e}")
    assert err < TOL[precision] * (10 if precision == "c128" else 1)
    # a Grover-style oracle matrix (diagonal of +-1) and a permutation
    d = np.where(rng.random(N) < 0.3, -1.0, 1.0).astype(np.complex128)
    before = r.getRegisterStorage()
    r.ApplyOperatorMatrix(np.diag(d))
    assert np.max(np.abs(r.getRegisterStorage() - d * before)) < TOL[precision]
    perm = rng.permutation(N)
    pm = np.zeros((N, N), dtype=np.complex128)
    pm[np.arange(N), perm] = 1
    before = r.getRegisterStorage()
    r.ApplyOperatorMatrix(pm)
    assert np.array_equal(r.getRegisterStorage(), before[perm])
    r.close()


def test_dense_operator_against_oracle_gates(oracle_lib):
    """The dense operator of a small circuit (built by the oracle, column by
    column from basis states) applied with ApplyOperatorMatrix equals the
    oracle's gate-by-gate run."""
    from oracle.oracle import OracleRegister

    n = 8
    seq = ansatz_state(n, 3)
    N = 1 << n
    m = np.zeros((N, N), dtype=np.complex128)
    for c in range(N):
        ref = OracleRegister(n)
        e = np.zeros(N, dtype=np.complex128)
        e[c] = 1
        ref.set_state(e)
        for g, qs in seq:
            ref.apply_gate(g, *qs)
        m[:, c] = ref.state()
        ref.close()
    rng = np.random.default_rng(4)
    psi = random_state(rng, n)
    ref = OracleRegister(n)
    ref.set_state(psi)
    for g, qs in seq:
        ref.apply_gate(g, *qs)
    r = QubitRegister(n)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    r.ApplyOperatorMatrix(m)
    assert np.max(np.abs(r.getRegisterStorage() - ref.state())) < 1e-12
    ref.close()
    r.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_streamed_host_vectors_over_two_chunks(precision):
    """23 qubits = two 64 MiB transfer chunks: the diagonal operator,
    stateFidelity / inner products with a host vector and (complex64) the
    converting upload and download."""
    n = 23
    rng = np.random.default_rng(23)
    psi = random_state(rng, n)
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    stored = r.getRegisterStorage()
    tol = TOL[precision]
    assert np.max(np.abs(stored - psi)) < tol
    phase = np.exp(1j * rng.uniform(0, 2 * math.pi, size=1 << n))
    r.apply_diagonal(phase)
    after = r.getRegisterStorage()
    assert np.max(np.abs(after - phase * stored)) < tol
    other = random_state(rng, n)
    z = r.inner_host(other)
    want = np.vdot(after, other)
    assert abs(z - want) < (1e-12 if precision == "c128" else 1e-6)
    f = r.stateFidelity(phase * stored * np.exp(0.3j))
    assert abs(f - 1) < (1e-12 if precision == "c128" else 1e-6)
    r.close()
