"""Every launch variant of the streaming gate kernels (orbits per thread,
nontemporal access, block regions, orbit spread) forced in turn for every
qubit and checked against the oracle: the autotuner may pick any of them,
and it times them on identity gates only."""
import numpy as np
import pytest

from gatecatalog import catalog
from oracle.oracle import OracleRegister
from qcsim_amd import QubitRegister
from qcsim_amd import gates as G
from qcsim_amd._lib import QCS_COMPLEX64, QCS_COMPLEX128, check, lib
from refmath import random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_every_variant_matches_the_oracle(precision, tol):
    n = 22  # 2^21 orbits: room for every spread (up to 2^20 apart, two per thread)
    prec = QCS_COMPLEX128 if precision == "c128" else QCS_COMPLEX64
    rng = np.random.default_rng(22)
    one, two, three = catalog(22)
    fams = one + two + three
    psi = random_state(rng, n)
    o = OracleRegister(n)
    o.set_state(psi)
    g = QubitRegister(n, precision=precision)
    g.setRegisterStorageFastNoNormalize(psi.copy())
    nv = lib().qcs_autotune_variants()
    try:
        for v in range(nv):
            for q in range(n):
                check(lib().qcs_set_tuned_variant(prec, n, q, v))
            # generic 1-qubit gates on a low, a middle and a high target
            # (k_1q: every layout), then gates of random families
            for q in (int(rng.integers(0, 4)), int(rng.integers(4, 16)), int(rng.integers(16, n))):
                u = G.SingleQubitGate(G.haar_u2(rng))
                g.ApplyGate(u, q)
                o.apply_gate(u, q)
            for _ in range(4):
                _, gate = fams[int(rng.integers(len(fams)))]
                qs = [int(x) for x in rng.choice(n, size=gate.getQubitsNumber(), replace=False)]
                g.ApplyGate(gate, *qs)
                o.apply_gate(gate, *qs)
            err = np.max(np.abs(g.getRegisterStorage() - o.state()))
            assert err < tol, (v, err)
    finally:
        for q in range(n):
            lib().qcs_set_tuned_variant(prec, n, q, -1)
    g.close()
    o.close()


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_every_variant_every_target_and_control(precision, tol):
    """Each variant forced in turn, on every target of a 12-qubit state, for
    the three k_1q modes (generic H / U, antidiagonal X / CNOT / Toffoli,
    diagonal Rz / CRz) with 0, 1 and 2 controls at random positions below
    and above the target: the one-amplitude-per-lane layouts take their
    partner from another lane of the block, in a block of 256 consecutive
    amplitudes or of two 128-amplitude halves, depending on the target's
    position among the non-control bits."""
    n = 12
    prec = QCS_COMPLEX128 if precision == "c128" else QCS_COMPLEX64
    rng = np.random.default_rng(12)
    psi = random_state(rng, n)
    o = OracleRegister(n)
    o.set_state(psi)
    g = QubitRegister(n, precision=precision)
    g.setRegisterStorageFastNoNormalize(psi.copy())
    nv = lib().qcs_autotune_variants()
    try:
        for v in range(nv):
            for q in range(n):
                check(lib().qcs_set_tuned_variant(prec, n, q, v))
            for t in range(n):
                others = [q for q in range(n) if q != t]
                c1, c2 = (int(x) for x in rng.choice(others, size=2, replace=False))
                th = float(rng.uniform(0, 2 * np.pi))
                seq = [
                    (G.HadamardGate(), [t]),
                    (G.SingleQubitGate(G.haar_u2(rng)), [t]),
                    (G.PauliXGate(), [t]),
                    (G.RzGate(th), [t]),
                    (G.CNOTGate(), [t, c1]),
                    (G.ControlledUGate(th, 0.3, 1.1, 0.2), [t, c2]),
                    (G.ControlledRzGate(th), [t, c1]),
                    (G.ToffoliGate(), [t, c1, c2]),
                ]
                for gate, qs in seq:
                    g.ApplyGate(gate, *qs)
                    o.apply_gate(gate, *qs)
            err = np.max(np.abs(g.getRegisterStorage() - o.state()))
            assert err < tol, (v, err)
    finally:
        for q in range(n):
            lib().qcs_set_tuned_variant(prec, n, q, -1)
    g.close()
    o.close()
