"""Every gate family of the reference (and one instance per kernel K1..K15),
shared by the oracle tests and the GPU parity tests."""
import numpy as np

from qcsim_amd import gates as G

from refmath import random_unitary


class CCGeneric(G.ThreeQubitsControlledGate):
    """Two-control gate with a generic 2x2 target block: kernel K13."""

    traits = G.T_CONTROLLED | G.T_CTRL_Q1

    def __init__(self, u):
        super().__init__()
        self._m[6:8, 6:8] = u


class CCDiag(G.ThreeQubitsControlledGate):
    """Two-control diagonal gate with m66 != 1: kernel K11 (both entries)."""

    traits = G.T_CONTROLLED | G.T_CTRL_Q1 | G.T_DIAGONAL

    def __init__(self, a, b):
        super().__init__()
        self._m[6, 6] = np.exp(1j * a)
        self._m[7, 7] = np.exp(1j * b)


def catalog(seed=1234):
    """(name, gate) pairs covering K1..K15."""
    rng = np.random.default_rng(seed)
    one = [
        ("H", G.HadamardGate()),
        ("Hy", G.HyGate()),
        ("S", G.SGate()),
        ("Sdg", G.SDGGate()),
        ("T", G.TGate()),
        ("Tdg", G.TDGGate()),
        ("P", G.PhaseShiftGate(0.7)),
        ("X", G.PauliXGate()),
        ("Y", G.PauliYGate()),
        ("Z", G.PauliZGate()),
        ("SX", G.SquareRootNOTGate()),
        ("SXdg", G.SquareRootNOTDagGate()),
        ("Split", G.SplitterGate()),
        ("Rx", G.RxGate(0.3)),
        ("Ry", G.RyGate(1.1)),
        ("Rz", G.RzGate(-0.4)),
        ("U", G.UGate(0.3, 0.5, 0.7, 0.9)),
        ("Haar", G.SingleQubitGate(G.haar_u2(rng))),
    ]
    cu = G.TwoQubitsControlledGate()
    cu.SetOperation(random_unitary(rng, 2))
    two = [
        ("SWAP", G.SwapGate()),
        ("iSWAP", G.iSwapGate()),
        ("iSWAPdg", G.iSwapDagGate()),
        ("Dec", G.DecrementGate()),
        ("CNOT", G.CNOTGate()),
        ("CY", G.ControlledYGate()),
        ("CZ", G.ControlledZGate()),
        ("CH", G.ControlledHadamardGate()),
        ("CSX", G.ControlledSquareRootNOTGate()),
        ("CSXdg", G.ControlledSquareRootNOTDagGate()),
        ("CPhase", G.ControlledPhaseGate()),
        ("CP", G.ControlledPhaseShiftGate(0.9)),
        ("CU", G.ControlledUGate(0.4, 1.2, -0.3, 0.2)),
        ("CRx", G.ControlledRxGate(0.8)),
        ("CRy", G.ControlledRyGate(-1.3)),
        ("CRz", G.ControlledRzGate(0.6)),
        ("CSetOp", cu),
        ("U4", G.TwoQubitsGate(random_unitary(rng, 4))),
    ]
    c14 = G.ThreeQubitsControlledGate()
    c14.SetOperation(random_unitary(rng, 4))
    three = [
        ("Toffoli", G.ToffoliGate()),
        ("Fredkin", G.FredkinGate()),
        ("CCZ", G.CCZGate()),
        ("CCdiag", CCDiag(0.4, -1.1)),
        ("CCgeneric", CCGeneric(random_unitary(rng, 2))),
        ("C4x4", c14),
        ("U8", G.ThreeQubitsGate(random_unitary(rng, 8))),
    ]
    return one, two, three


def placements(n, k, rng, count):
    """Distinct qubit tuples of length k: extremes first, then random."""
    out = []
    if k == 1:
        out = [(0,), (n - 1,)]
    elif k == 2:
        out = [(0, 1), (1, 0), (0, n - 1), (n - 1, 0)]
    else:
        out = [(0, 1, 2), (2, 1, 0), (0, n - 1, 1), (n - 1, 0, n // 2)]
    while len(out) < count:
        t = tuple(int(x) for x in rng.choice(n, size=k, replace=False))
        if t not in out:
            out.append(t)
    return out[:count]
