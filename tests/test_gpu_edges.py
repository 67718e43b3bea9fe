"""Edge cases on the GPU against the oracle: the smallest registers (1-3
qubits, where no kernel has a full block of work), empty circuits and shot
lists, single-target QFT ranges, fusion on states smaller than one MFMA
tile, and QAOA instances without edges or with a single site."""
import numpy as np
import pytest

from gatecatalog import catalog
from oracle.oracle import OracleRegister
from qcsim_amd import QubitRegister
from qcsim_amd import gates as G
from qcsim_amd.qaoa import IsingInstance, QaoaBatch
from refmath import qft_closed, random_state

pytestmark = pytest.mark.gpu


def pair(n, psi):
    g = QubitRegister(n)
    g.setRegisterStorageFastNoNormalize(psi.copy())
    o = OracleRegister(n)
    o.set_state(psi)
    return g, o


@pytest.mark.parametrize("n", [1, 2, 3])
def test_smallest_registers_every_family(n):
    rng = np.random.default_rng(n)
    one, two, three = catalog(n)
    fams = one + (two if n >= 2 else []) + (three if n >= 3 else [])
    psi = random_state(rng, n)
    g, o = pair(n, psi)
    for _, gate in fams:
        k = gate.getQubitsNumber()
        qs = [int(x) for x in rng.permutation(n)[:k]]
        g.ApplyGate(gate, *qs)
        o.apply_gate(gate, *qs)
    assert np.max(np.abs(g.getRegisterStorage() - o.state())) < 1e-12
    # QFT over the whole (tiny) register and over single-target ranges
    g.qft()
    o.qft(0, n - 1)
    assert np.max(np.abs(g.getRegisterStorage() - o.state())) < 1e-12
    for q in range(n):
        g.qft(q, q)
        o.qft(q, q)
    assert np.max(np.abs(g.getRegisterStorage() - o.state())) < 1e-12
    # measurement on a 1..3-qubit state with supplied uniforms
    for u in (0.123, 0.37, 0.81):
        assert g.MeasureNoCollapse(u=u) == o.measure_no_collapse(u)
    g.close()


def test_empty_circuit_and_zero_shots():
    n = 5
    psi = random_state(np.random.default_rng(2), n)
    g, _ = pair(n, psi)
    g.apply_circuit([], fuse_width=0)
    g.apply_circuit([], fuse_width=5)
    assert np.array_equal(g.getRegisterStorage(), psi)
    out = g.sample(np.zeros(0))
    assert len(out) == 0
    prog = g.compile_program([])
    prog.run()
    assert np.array_equal(g.getRegisterStorage(), psi)
    prog.close()
    g.close()


@pytest.mark.parametrize("n", [5, 8, 10])
def test_fused_blocks_on_states_below_one_mfma_tile(n):
    """Fused 4- and 5-qubit blocks when 2^n is smaller than an MFMA tile
    (the non-MFMA dense kernel) or just one tile."""
    rng = np.random.default_rng(40 + n)
    one, two, three = catalog(n)
    fams = one + two + three
    seq = []
    for _ in range(80):
        _, gate = fams[int(rng.integers(len(fams)))]
        seq.append((gate, [int(x) for x in rng.choice(n, size=gate.getQubitsNumber(), replace=False)]))
    psi = random_state(rng, n)
    for width in (4, 5):
        g, o = pair(n, psi)
        g.apply_circuit(seq, fuse_width=width)
        for gate, qs in seq:
            o.apply_gate(gate, *qs)
        assert np.max(np.abs(g.getRegisterStorage() - o.state())) < 1e-12, width
        g.close()


def test_qaoa_instances_without_edges_or_with_one_site():
    """n = 1 (no edges), n = 2 with one edge, n = 5 with no edges (a pure
    field model: every layer is a product of single-qubit rotations)."""
    insts = [
        IsingInstance(1, [], np.zeros(0), np.array([0.7])),
        IsingInstance(2, [(0, 1)], np.array([-1.3]), np.array([0.2, -0.4])),
        IsingInstance(5, [], np.zeros(0), np.array([0.5, -0.1, 0.3, 0.0, 1.1])),
    ]
    p = 2
    rng = np.random.default_rng(9)
    Gm = rng.uniform(0, np.pi, (len(insts), p))
    Bm = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p)
    E = batch.evaluate(Gm, Bm)
    for i, inst in enumerate(insts):
        o = OracleRegister(inst.n)
        o.qaoa_exec(inst.edges, inst.J, inst.h, Gm[i], Bm[i])
        e_ref = o.ising_energy(inst.edges, inst.J, inst.h)
        assert abs(E[i] - e_ref) < 2e-5 * (1 + np.sum(np.abs(inst.h)) + np.sum(np.abs(inst.J))), i
        assert np.max(np.abs(batch.state(i) - o.state())) < 3e-6, i
    batch.close()
