"""Pins the CPU oracle's gate kernels (K1..K15) against independent math.

- every gate family at several placements against the dense Kronecker
  embedding (SimpleGates.h:211-232, 311-333, 411-434) applied by brute force;
- closed forms: Bell/GHZ, CNOT and Toffoli truth tables, SWAP permutation;
- the serial and OpenMP twins agree bit for bit above the 8192-amplitude
  threshold (QubitRegisterCalculator.h:1276), which the reference never tests;
- CheckQubits error messages (QubitRegister.h:677-690).
"""
import numpy as np
import pytest

from gatecatalog import catalog, placements
from oracle.oracle import OracleRegister
from qcsim_amd import gates as G
from refmath import apply_dense, random_state

N = 6


@pytest.mark.parametrize("group", [0, 1, 2])
def test_gates_match_dense_kronecker(group):
    rng = np.random.default_rng(7 + group)
    fams = catalog()[group]
    k = group + 1
    for name, gate in fams:
        for qs in placements(N, k, rng, 6):
            psi = random_state(rng, N)
            r = OracleRegister(N, multithreading=False)
            r.set_state(psi)
            r.apply_gate(gate, *qs)
            want = apply_dense(psi, gate.getRawOperatorMatrix(), N, qs)
            err = np.max(np.abs(r.state() - want))
            assert err < 1e-14, (name, qs, err)


def test_serial_and_omp_twins_agree():
    n = 14  # 16384 amplitudes: above the OpenMP threshold
    rng = np.random.default_rng(3)
    one, two, three = catalog(5)
    for fams, k in ((one, 1), (two, 2), (three, 3)):
        for name, gate in fams:
            qs = placements(n, k, rng, 3)[-1]
            psi = random_state(rng, n)
            a = OracleRegister(n, multithreading=False)
            b = OracleRegister(n, multithreading=True)
            a.set_state(psi)
            b.set_state(psi)
            a.apply_gate(gate, *qs)
            b.apply_gate(gate, *qs)
            assert np.array_equal(a.state(), b.state()), name


def test_bell_ghz_and_truth_tables():
    r = OracleRegister(3)
    r.apply_gate(G.HadamardGate(), 0)
    r.apply_gate(G.CNOTGate(), 1, 0)
    r.apply_gate(G.CNOTGate(), 2, 1)
    s = 1 / np.sqrt(2)
    assert np.allclose(r.state(), [s, 0, 0, 0, 0, 0, 0, s], atol=1e-15)
    # CNOT (target q=0, control c1=1) truth table on basis states
    for b in range(4):
        r = OracleRegister(2)
        psi = np.zeros(4, complex)
        psi[b] = 1
        r.set_state(psi)
        r.apply_gate(G.CNOTGate(), 0, 1)
        want = b ^ 1 if b & 2 else b
        assert abs(r.state()[want] - 1) < 1e-15
    # Toffoli: controls c1, c2 flip qubit
    for b in range(8):
        r = OracleRegister(3)
        psi = np.zeros(8, complex)
        psi[b] = 1
        r.set_state(psi)
        r.apply_gate(G.ToffoliGate(), 0, 1, 2)
        want = b ^ 1 if (b & 6) == 6 else b
        assert abs(r.state()[want] - 1) < 1e-15
    # SWAP permutes bits
    for b in range(8):
        r = OracleRegister(3)
        psi = np.zeros(8, complex)
        psi[b] = 1
        r.set_state(psi)
        r.apply_gate(G.SwapGate(), 0, 2)
        want = (b & 2) | ((b & 1) << 2) | ((b >> 2) & 1)
        assert abs(r.state()[want] - 1) < 1e-15


def test_x_y_h_closed_forms():
    r = OracleRegister(1)
    r.apply_gate(G.PauliXGate(), 0)
    assert np.allclose(r.state(), [0, 1])
    r.apply_gate(G.PauliYGate(), 0)  # Y|1> = -i|0>
    assert np.allclose(r.state(), [-1j, 0])
    r = OracleRegister(1)
    r.apply_gate(G.HadamardGate(), 0)
    assert np.allclose(r.state(), [2**-0.5, 2**-0.5], atol=1e-16)


@pytest.mark.parametrize(
    "k,q,c1,c2,msg",
    [
        (1, 3, 0, 0, "Qubit number is too high"),
        (2, 0, 3, 0, "Controlling qubit number is too high"),
        (2, 1, 1, 0, "Qubit and controlling qubit are the same"),
        (3, 0, 1, 5, "Controlling qubit number is too high"),
        (3, 0, 1, 1, "Qubits must be different"),
    ],
)
def test_check_qubits_messages(k, q, c1, c2, msg):
    r = OracleRegister(3)
    gate = {1: G.HadamardGate(), 2: G.CNOTGate(), 3: G.ToffoliGate()}[k]
    with pytest.raises(ValueError, match=msg):
        r.apply_gate(gate, q, c1, c2)
