"""Host side of batched VQE (no GPU): the ansatz gate list follows
VQE::PauliStringVQE::Ansatz (VQE.h:78-114) — gate counts and ladder order —
and Hamiltonian masks follow PauliString's Z default."""
from qcsim_amd import gates as G
from qcsim_amd.pauli import OP_X, PauliString
from qcsim_amd.vqe import ansatz_gates, hamiltonian


def kinds(seq):
    return ["cx" if isinstance(g, G.CNOTGate) else type(g).__name__[:2] for g, _ in seq]


def test_no_ansatz_below_n_parameters():
    assert ansatz_gates(4, [0.1, 0.2, 0.3]) == []


def test_first_layer_and_ladders():
    n = 3
    seq = ansatz_gates(n, [0.1] * 16)
    # layer of 3 (Ry, Rz) pairs, ladder 0->1, 1->2, 2->0, layer of 3, reversed ladder, layer of 2
    cx = [(qs[0], qs[1]) for g, qs in seq if isinstance(g, G.CNOTGate)]
    assert cx == [(1, 0), (2, 1), (0, 2), (1, 2), (0, 1), (2, 0)]
    assert kinds(seq).count("Ry") == 8 and kinds(seq).count("Rz") == 8
    assert [qs[0] for g, qs in seq if isinstance(g, G.RyGate)] == [0, 1, 2, 0, 1, 2, 0, 1]
    # two qubits: no wrap-around CNOT
    seq2 = ansatz_gates(2, [0.2] * 8)
    assert [(qs[0], qs[1]) for g, qs in seq2 if isinstance(g, G.CNOTGate)] == [(1, 0)]


def test_odd_parameter_count_leaves_the_last_unused():
    seq = ansatz_gates(2, [0.1] * 5)
    assert kinds(seq) == ["Ry", "Rz", "Ry", "Rz"]


def test_hamiltonian_masks():
    t = PauliString(3)
    t.setOperatorForQubit(1, OP_X)
    t.setCoefficient(-0.5)
    assert hamiltonian([t]) == ([0b010], [0b101], [-0.5])
