"""Host logic of Pauli strings (qcsim_amd.pauli), no GPU: masks from the
reference's PauliString (Z by default, PauliString.h:9-27) and from labels,
and the recognition of ExpectationValue gate lists that are Pauli products,
checked against dense Kronecker products."""
import numpy as np

from qcsim_amd import gates as G
from qcsim_amd import pauli as P
from refmath import dense_pauli, pauli_ev, pauli_evs, random_state


def test_pauli_string_masks_default_to_z():
    t = P.PauliString(4)
    assert t.masks() == (0, 0b1111)
    t.setOperatorForQubit(1, P.OP_X)
    t.setOperatorForQubit(3, P.OP_Y)
    t.setOperatorForQubit(7, P.OP_X)  # beyond the string: ignored
    assert t.getOperatorForQubit(9) == P.OP_Z
    assert t.masks() == (0b1010, 0b1101)
    assert P.masks_from_label("XIZY") == (0b1001, 0b1100)


def embed(m, q, n):
    return np.kron(np.eye(1 << (n - 1 - q)), np.kron(m, np.eye(1 << q)))


def test_gate_lists_recognised_as_pauli_products():
    n = 4
    rng = np.random.default_rng(3)
    mats = [np.eye(2), G.PauliXGate().getRawOperatorMatrix(), G.PauliYGate().getRawOperatorMatrix(),
            G.PauliZGate().getRawOperatorMatrix()]
    for trial in range(50):
        gates = []
        full = np.eye(1 << n, dtype=np.complex128)
        for _ in range(int(rng.integers(1, 8))):
            q = int(rng.integers(n))
            m = mats[int(rng.integers(4))] * complex(rng.normal(), rng.normal())
            gates.append(G.AppliedGate(m, q))
            full = embed(m, q, n) @ full
        coef, x, z = P.gates_as_pauli(gates, n)
        assert np.allclose(full, coef * dense_pauli(x, z, n), atol=1e-12), trial


def test_non_pauli_lists_are_not_recognised():
    h = G.AppliedGate(G.HadamardGate().getRawOperatorMatrix(), 0)
    assert P.gates_as_pauli([h], 2) is None
    cz = G.AppliedGate(G.ControlledZGate().getRawOperatorMatrix(), 0, 1)
    assert P.gates_as_pauli([cz], 2) is None
    s = G.AppliedGate(G.SGate().getRawOperatorMatrix(), 0)
    assert P.gates_as_pauli([s], 2) is None
    assert P.gates_as_pauli([s, s], 2) is not None  # S S = Z


def test_dense_pauli_y_convention():
    # Y = i X Z on one qubit, and qubit 0 is the least significant bit
    x = dense_pauli(1, 0, 1)
    z = dense_pauli(0, 1, 1)
    assert np.allclose(dense_pauli(1, 1, 1), 1j * x @ z)
    assert np.allclose(dense_pauli(1, 0, 2), np.kron(np.eye(2), x))


def test_row_form_matches_the_dense_string():
    """refmath.pauli_ev / pauli_evs (the row form, used as the reference at
    sizes where the dense matrix does not fit) against the Kronecker
    product."""
    rng = np.random.default_rng(5)
    n = 6
    psi = random_state(rng, n)
    xs, zs, want = [], [], []
    for _ in range(100):
        x, z = int(rng.integers(0, 1 << n)), int(rng.integers(0, 1 << n))
        xs.append(x)
        zs.append(z)
        want.append(np.vdot(psi, dense_pauli(x, z, n) @ psi).real)
        assert abs(pauli_ev(psi, x, z) - want[-1]) < 1e-14, (x, z)
    assert np.max(np.abs(pauli_evs(psi, xs, zs) - np.array(want))) < 1e-14
