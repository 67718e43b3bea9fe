"""GPU parity: the HIP engine through the C-ABI against the CPU oracle.

Bars: FP64 amplitudes within 1e-12 max-abs of the oracle; complex64 within
2e-6 (single-precision storage, FP32 arithmetic); measurement outcomes equal
to the oracle's for the same supplied uniforms unless the uniform lies within
the stated rounding margin of a CDF boundary.
"""
import json
import math
import os

import numpy as np
import pytest

from gatecatalog import catalog, placements
from oracle.oracle import OracleRegister
from qcsim_amd import QuantumFourierTransform, QubitRegister
from qcsim_amd import gates as G
from refmath import cdf_first, qft_closed, random_state

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TOL64 = 1e-12
TOL32 = 2e-6


def pair(n, psi, precision="c128"):
    g = QubitRegister(n, precision=precision)
    g.setRegisterStorageFastNoNormalize(psi)
    o = OracleRegister(n)
    o.set_state(psi)
    return g, o


def cdf_margin(psi, probs):
    """Distance from each prob to the nearest CDF value (sequential cumsum)."""
    p = psi.real * psi.real + psi.imag * psi.imag
    c = np.cumsum(p)
    i = np.clip(np.searchsorted(c, probs), 0, c.shape[0] - 1)
    lo = np.abs(c[i] - probs)
    lo2 = np.abs(c[np.clip(i - 1, 0, None)] - probs)
    return np.minimum(lo, lo2)


@pytest.mark.parametrize("precision,tol", [("c128", TOL64), ("c64", TOL32)])
@pytest.mark.parametrize("n", [7, 13])
def test_every_gate_family(precision, tol, n):
    rng = np.random.default_rng(100 + n)
    for group, fams in enumerate(catalog(n)):
        k = group + 1
        for name, gate in fams:
            for qs in placements(n, k, rng, 4):
                psi = random_state(rng, n)
                g, o = pair(n, psi, precision)
                g.ApplyGate(gate, *qs)
                o.apply_gate(gate, *qs)
                err = np.max(np.abs(g.getRegisterStorage() - o.state()))
                assert err < tol, (precision, name, qs, err)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_blocks(k):
    from refmath import random_unitary

    n = 11
    rng = np.random.default_rng(k)
    for _ in range(4):
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
        m = random_unitary(rng, 1 << k)
        psi = random_state(rng, n)
        g, o = pair(n, psi)
        g.apply_dense(qs, m)
        o.apply_dense(qs, m)
        assert np.max(np.abs(g.getRegisterStorage() - o.state())) < TOL64, qs


@pytest.mark.parametrize("n", [1, 2, 5, 9, 12, 14])
def test_qft_matches_oracle_and_closed_form(n):
    rng = np.random.default_rng(n)
    psi = random_state(rng, n)
    for inverse in (False, True):
        g, o = pair(n, psi)
        g.qft(inverse=inverse)
        o.qft(inverse=inverse)
        got = g.getRegisterStorage()
        assert np.max(np.abs(got - o.state())) < TOL64
        assert np.max(np.abs(got - qft_closed(psi, n, inverse=inverse))) < TOL64


def test_qft_subranges_and_facade():
    n = 10
    rng = np.random.default_rng(77)
    psi = random_state(rng, n)
    for sq, eq, swap in [(0, 4, True), (3, 8, True), (2, 9, False), (6, 6, True), (1, 1000, True)]:
        g, o = pair(n, psi)
        qft = QuantumFourierTransform(n, sq, eq)
        qft.QFT(g, swap)
        o.qft(sq, eq, swap)
        assert np.max(np.abs(g.getRegisterStorage() - o.state())) < TOL64, (sq, eq)
        qft.IQFT(g, swap)
        assert np.max(np.abs(g.getRegisterStorage() - psi)) < TOL64


def test_measurement_known_answers():
    with open(os.path.join(GOLDEN, "measurement_known_answers.json")) as f:
        cases = json.load(f)
    psi = np.array([0.5, -0.5j, 0, 1 / math.sqrt(2)])
    for c in cases:
        g = QubitRegister(2)
        g.setRawAmplitude(0, 0.5)
        g.setRawAmplitude(1, -0.5j)
        g.setRawAmplitude(3, 1 / math.sqrt(2))
        g.Normalize()
        assert np.max(np.abs(g.getRegisterStorage() - psi)) < 1e-15
        assert g.MeasureQubit(c["qubit"], u=c["u"]) == c["outcome"]
        want = np.array([complex(*z) for z in c["amplitudes"]])
        assert np.max(np.abs(g.getRegisterStorage() - want)) < 1e-12, c


@pytest.mark.parametrize("n", [3, 11, 16])
def test_cdf_searches(n):
    rng = np.random.default_rng(n + 1)
    psi = random_state(rng, n)
    psi[rng.choice(1 << n, size=(1 << n) // 4, replace=False)] = 0
    psi /= np.linalg.norm(psi)
    g, o = pair(n, psi)
    us = np.concatenate([rng.random(300), [0.0, 0.5, 1 - 2**-53]])
    probs = 1.0 - us
    margin = cdf_margin(psi, probs)
    mism = 0
    for u, m in zip(us, margin):
        a = g.MeasureNoCollapse(u=u)
        b = o.measure_no_collapse(u)
        if a != b:
            mism += 1
            assert m < 1e-13 * (1 << n), (u, a, b, m)
    assert mism <= 1
    # range, single qubit and sampling
    for first, second in [(0, 0), (1, n - 1), (n - 1, n - 1)]:
        for u in us[:40]:
            assert g.MeasureNoCollapse(first, second, u=u) == o.measure_no_collapse(u, first, second)
    s_g = g.sample(us)
    s_o = o.sample(us)
    bad = np.nonzero(s_g != s_o)[0]
    assert all(margin[i] < 1e-13 * (1 << n) for i in bad)
    s_g = g.sample(us, 1, min(3, n - 1))
    s_o = o.sample(us, 1, min(3, n - 1))
    assert np.count_nonzero(s_g != s_o) <= len(bad)


def test_measure_defaults_on_unnormalised_state():
    n = 6
    rng = np.random.default_rng(1)
    psi = 0.5 * random_state(rng, n)
    g, o = pair(n, psi)
    assert g.MeasureNoCollapse(u=0.0) == o.measure_no_collapse(0.0) == 0
    assert g.MeasureAll(u=0.0) == o.measure_all(0.0) == (1 << n) - 1
    assert np.max(np.abs(g.getRegisterStorage() - o.state())) == 0


@pytest.mark.parametrize("n", [4, 12])
def test_measure_collapse(n):
    rng = np.random.default_rng(3 * n)
    psi = random_state(rng, n)
    for first, second in [(0, 0), (n - 1, n - 1), (1, 2), (0, n - 1)]:
        for u in (0.13, 0.5, 0.91):
            g, o = pair(n, psi)
            a = g.Measure(first, second, u=u)
            b = o.measure(first, second, u)
            assert a == b
            assert np.max(np.abs(g.getRegisterStorage() - o.state())) < TOL64
    for u in (0.2, 0.7):
        g, o = pair(n, psi)
        assert g.MeasureAll(u=u) == o.measure_all(u)
        assert np.max(np.abs(g.getRegisterStorage() - o.state())) == 0


def test_reductions_and_state_utilities():
    n = 12
    rng = np.random.default_rng(8)
    psi = random_state(rng, n)
    g, o = pair(n, psi)
    for q in (0, 5, n - 1):
        assert abs(g.GetQubitProbability(q) - o.qubit_probability(q)) < 1e-14
    assert abs(g.norm2() - 1) < 1e-13
    g.setRegisterStorageFastNoNormalize(psi * 3)
    g.Normalize()
    assert np.max(np.abs(g.getRegisterStorage() - psi)) < 1e-15
    h = QubitRegister(n)
    h.setRegisterStorageFastNoNormalize(psi)
    assert abs(h.inner(g) - 1) < 1e-13
    assert abs(g.stateFidelity(psi) - 1) < 1e-13
    g.setToBasisState(5)
    assert g.getBasisStateAmplitude(5) == 1 and g.getBasisStateAmplitude(4) == 0
    assert g.getBasisStateAmplitude(1 << n) == 0  # out of range: 0, silently
    g.setToBasisState(1 << n)  # out of range: no-op
    assert g.getBasisStateAmplitude(5) == 1
    g.setToCatState()
    s = 1 / math.sqrt(2)
    assert abs(g.getBasisStateAmplitude(0) - s) < 1e-16 and abs(g.getBasisStateAmplitude((1 << n) - 1) - s) < 1e-16
    g.setToEqualSuperposition()
    assert abs(g.norm2() - 1) < 1e-12
    g.setRegisterStorageFastNoNormalize(psi * (0.3 - 0.4j))
    g.AdjustPhaseAndNormalize()
    v = g.getRegisterStorage()
    assert abs(v[0].imag) < 1e-15 and v[0].real > 0 and abs(np.linalg.norm(v) - 1) < 1e-13


def test_expectation_values_and_record_replay():
    # Tests.cpp:665-723
    z0 = G.AppliedGate(G.PauliZGate().getRawOperatorMatrix(), 0)
    r = QubitRegister(3)
    assert r.ExpectationValue([]) == 1.0
    assert abs(r.ExpectationValue([z0]) - 1) < 1e-12
    r.ApplyGate(G.PauliXGate(), 0)
    assert abs(r.ExpectationValue([z0]) + 1) < 1e-12
    r.ApplyGate(G.HadamardGate(), 0)
    assert abs(r.ExpectationValue([z0])) < 1e-12
    r.setToBasisState(0)
    r.ApplyGate(G.HadamardGate(), 1)
    assert abs(r.ExpectationValue([G.AppliedGate(G.PauliZGate().getRawOperatorMatrix(), 1)])) < 1e-12
    # Compute / Uncompute restores the state
    n = 8
    rng = np.random.default_rng(4)
    psi = random_state(rng, n)
    r = QubitRegister(n)
    r.setRegisterStorageFastNoNormalize(psi)
    r.ComputeStart()
    r.ApplyGate(G.ToffoliGate(), 0, 3, 5)
    r.ApplyGate(G.SingleQubitGate(G.haar_u2(rng)), 2)
    r.ApplyGate(G.ControlledRyGate(0.4), 6, 1)
    r.ComputeEnd()
    r.Uncompute()
    assert np.max(np.abs(r.getRegisterStorage() - psi)) < 1e-13
    r.SaveState()
    r.ApplyGate(G.HadamardGate(), 0)
    r.RestoreState()
    assert np.max(np.abs(r.getRegisterStorage() - psi)) < 1e-13


@pytest.mark.parametrize(
    "gate,args,msg",
    [
        (G.HadamardGate(), (9,), "Qubit number is too high"),
        (G.CNOTGate(), (0, 9), "Controlling qubit number is too high"),
        (G.CNOTGate(), (2, 2), "Qubit and controlling qubit are the same"),
        (G.ToffoliGate(), (0, 1, 1), "Qubits must be different"),
    ],
)
def test_invalid_arguments(gate, args, msg):
    r = QubitRegister(4)
    with pytest.raises(ValueError, match=msg):
        r.ApplyGate(gate, *args)


def test_ising_energy_matches_oracle():
    n = 10
    rng = np.random.default_rng(12)
    edges = [(i, (i + 1) % n) for i in range(n)] + [(i, (i + 3) % n) for i in range(0, n, 2)]
    edges = [(min(a, b), max(a, b)) for a, b in edges]
    J = rng.standard_normal(len(edges))
    h = 0.5 * rng.standard_normal(n)
    psi = random_state(rng, n)
    g, o = pair(n, psi)
    assert abs(g.ising_energy(edges, J, h) - o.ising_energy(edges, J, h)) < 1e-12


def test_random_circuit_20_qubits_with_qft_and_samples():
    """20 qubits complex128: 40 layers of Haar U(2) on every qubit + CNOT/CZ
    on a random pairing, then a 20-qubit QFT; 1024 supplied uniforms."""
    n, layers = 20, 40
    rng = np.random.default_rng(20240)
    g = QubitRegister(n)
    o = OracleRegister(n)
    for _ in range(layers):
        for q in range(n):
            u = G.SingleQubitGate(G.haar_u2(rng))
            g.ApplyGate(u, q)
            o.apply_gate(u, q)
        perm = rng.permutation(n)
        for a, b in zip(perm[0::2], perm[1::2]):
            gate = G.CNOTGate() if rng.random() < 0.5 else G.ControlledZGate()
            g.ApplyGate(gate, int(a), int(b))
            o.apply_gate(gate, int(a), int(b))
    g.qft()
    o.qft()
    ref = o.state()
    got = g.getRegisterStorage()
    err = float(np.max(np.abs(got - ref)))
    infid = 1 - abs(np.vdot(ref, got)) ** 2
    print(f"\n20-qubit circuit: max-abs {err:.3e} infidelity {infid:.3e}")
    assert err < TOL64
    us = rng.random(1024)
    margin = cdf_margin(ref, 1 - us)
    sg = g.sample(us)
    so = o.sample(us)
    bad = np.nonzero(sg != so)[0]
    print(f"20-qubit circuit: {len(bad)} sample mismatches, min CDF margin {margin.min():.3e}")
    assert all(margin[i] < 1e-10 for i in bad)


@pytest.mark.parametrize("precision,tol", [("c128", TOL64), ("c64", TOL32)])
@pytest.mark.parametrize("sq,eq", [(0, 15), (1, 15), (2, 13), (4, 15), (3, 12), (7, 15), (0, 9), (5, 5)])
def test_qft_fused_paths(sq, eq, precision, tol):
    """Every pass type of the fused QFT (register groups, LDS chunk, tiled /
    pair / chunk bit reversal) against the oracle's gate sequence, in both
    storage precisions."""
    n = 16
    rng = np.random.default_rng(sq * 31 + eq)
    psi = random_state(rng, n)
    for inverse in (False, True):
        for swap in (True, False):
            g, o = pair(n, psi, precision)
            g.qft(sq, eq, swap, inverse)
            o.qft(sq, eq, swap, inverse)
            err = np.max(np.abs(g.getRegisterStorage() - o.state()))
            assert err < tol, (sq, eq, precision, inverse, swap, err)


@pytest.mark.parametrize("precision,tol", [("c128", TOL64), ("c64", TOL32)])
@pytest.mark.parametrize("sq,eq", [(0, 19), (0, 16), (1, 18), (5, 19), (12, 19)])
def test_qft_register_groups_20_qubits(sq, eq, precision, tol):
    """The register-group passes of widths 5 and 6 (orbits split over lane
    pairs) beside the 12-target LDS pass, on a random 20-qubit state against
    the oracle's gate sequence: (0, 19) takes groups of 6 and 2, (0, 16) one
    of 5, (1, 18) one of 6 plus 1, (5, 19) the LDS pass on 5..11 and a group
    of 6 plus 2, (12, 19) groups of 6 and 2 with no LDS pass."""
    n = 20
    rng = np.random.default_rng(sq * 37 + eq)
    psi = random_state(rng, n)
    for inverse in (False, True):
        for swap in (True, False):
            g, o = pair(n, psi, precision)
            g.qft(sq, eq, swap, inverse)
            o.qft(sq, eq, swap, inverse)
            err = np.max(np.abs(g.getRegisterStorage() - o.state()))
            assert err < tol, (sq, eq, precision, inverse, swap, err)


def _product_state(rng, n):
    """Haar U(2) on every qubit of |0...0>: psi[s] = prod_q u_q[bit_q(s)]."""
    us = [G.haar_u2(rng) for _ in range(n)]
    cols = [u[:, 0] for u in us]

    def amp(s):
        v = 1.0 + 0j
        for q in range(n):
            v *= cols[q][(s >> q) & 1]
        return v

    return us, amp


@pytest.mark.parametrize("n", [24, 30])
def test_qft_full_size_properties(n):
    """Size-independent checks at full size: QFT|j> has amplitudes
    exp(2 pi i j k / N) / sqrt(N) (the closed form sqrt(N) ifft), and
    IQFT(QFT(psi)) = psi on a product state whose amplitudes are known."""
    rng = np.random.default_rng(n)
    N = 1 << n
    r = QubitRegister(n)
    j = int(rng.integers(N))
    r.setToBasisState(j)
    r.qft()
    ks = rng.integers(N, size=32)
    for k in ks:
        want = np.exp(2j * np.pi * ((j * int(k)) % N) / N) / math.sqrt(N)
        assert abs(r.getBasisStateAmplitude(int(k)) - want) < 1e-12 / math.sqrt(N) * 64
    us, amp = _product_state(rng, n)
    r.setToBasisState(0)
    for q, u in enumerate(us):
        r.ApplyGate(G.SingleQubitGate(u), q)
    r.qft()
    r.qft(inverse=True)
    for s in rng.integers(N, size=64):
        assert abs(r.getBasisStateAmplitude(int(s)) - amp(int(s))) < 1e-13
    assert abs(r.norm2() - 1) < 1e-12
    r.close()


@pytest.mark.parametrize("precision,tol", [("c128", TOL64), ("c64", 5e-6)])
@pytest.mark.parametrize("n,k", [(11, 4), (11, 5), (12, 5), (16, 5), (16, 4), (12, 3)])
def test_dense_blocks_mfma(precision, tol, n, k):
    """Dense k-qubit blocks through the MFMA kernel (k = 4, 5 when n >= k + 6
    for complex128 with 64-orbit tiles, n >= k + 7 for complex64 with 128-orbit
    tiles; below that the register kernel) at low, high, mixed and reversed
    qubit orders: (11, 5) is the smallest complex128 MFMA case and a complex64
    fallback, (12, 5) the smallest complex64 MFMA case."""
    from refmath import random_unitary

    rng = np.random.default_rng(100 * n + k)
    choices = [list(range(k)), list(range(n - k, n)), list(range(k))[::-1],
               [int(x) for x in rng.choice(n, size=k, replace=False)]]
    for qs in choices:
        m = random_unitary(rng, 1 << k)
        psi = random_state(rng, n)
        g, o = pair(n, psi, precision)
        g.apply_dense(qs, m)
        o.apply_dense(qs, m)
        err = np.max(np.abs(g.getRegisterStorage() - o.state()))
        assert err < tol, (qs, err)


@pytest.mark.parametrize("precision,tol", [("c128", TOL64), ("c64", 2e-5)])
@pytest.mark.parametrize("width", [2, 3, 4, 5])
def test_fused_circuit(precision, tol, width):
    """qcs_apply_circuit with fusion against the oracle's gate-by-gate run."""
    n = 14
    rng = np.random.default_rng(width)
    one, two, three = catalog(width)
    fams = one + two + three
    seq = []
    for _ in range(150):
        name, gate = fams[int(rng.integers(len(fams)))]
        seq.append((gate, [int(x) for x in rng.choice(n, size=gate.getQubitsNumber(), replace=False)]))
    psi = random_state(rng, n)
    g, o = pair(n, psi, precision)
    g.apply_circuit(seq, fuse_width=width)
    for gate, qs in seq:
        o.apply_gate(gate, *qs)
    err = np.max(np.abs(g.getRegisterStorage() - o.state()))
    assert err < tol, err


def test_precision_toffoli_circuit_c64_against_fp64():
    """The layered circuit with Toffoli-class gates at 20 qubits: the complex64 register
    running fused FP32 MFMA blocks stays within 1e-5 max-abs and 1e-8
    infidelity of the FP64 register running the same gates unfused."""
    import bench

    out = bench.gpu_precision_fused_c64(20, 20)
    print(out)
    assert out["max_abs_error"] < 1e-5
    assert out["infidelity"] < 1e-8
    assert abs(out["c64_norm2_fp64"] - 1) < 1e-5


def test_clone_survives_overwrite_of_source():
    """Clone (qcs_copy_state) on a busy 27-qubit register, source reset
    right after: the copy holds the pre-reset state."""
    n = 27
    reg = QubitRegister(n)
    for q in range(n):
        reg.ApplyGate(G.HadamardGate(), q)
    c = reg.Clone()
    reg.setToBasisState(3)
    amp = 2.0 ** (-n / 2)
    assert abs(c.getBasisStateAmplitude(0) - amp) < 1e-12
    assert abs(c.getBasisStateAmplitude((1 << n) - 1) - amp) < 1e-12
    assert abs(c.norm2() - 1) < 1e-10
    assert abs(reg.getBasisStateAmplitude(3) - 1) < 1e-15
    c.close()
    reg.close()


def test_fused_matches_unfused_long_circuit():
    """The layered random circuit at 24 qubits (a long queue of launches on the
    register's stream): fused widths 4 and 5 agree with the unfused run."""
    import bench

    out = bench.gpu_circuit(24, 6)
    print(out)
    for k in ("fused4", "fused5"):
        assert out[k]["max_abs_vs_unfused"] < 1e-12, out[k]
        assert out[k]["infidelity_vs_unfused"] < 1e-12, out[k]


def test_layered_circuit_against_the_oracle_24_qubits():
    """The layered random circuit (60 layers: Haar U(2) on every qubit, then CNOT /
    CZ / controlled-Haar-U(2) on a random pairing) at 24 qubits, complex128,
    unfused and fused to widths 4 and 5 (MFMA blocks), against the oracle's
    gate-by-gate run from |0...0>: max-abs <= 1e-12, infidelity printed."""
    import bench

    n = 24
    seq = bench.random_layered_circuit(n, 60)
    o = OracleRegister(n)
    for gate, qs in seq:
        o.apply_gate(gate, *qs)
    ref = o.state()
    o.close()
    for width in (0, 4, 5):
        g = QubitRegister(n)
        g.apply_circuit(seq, fuse_width=width)
        psi = g.getRegisterStorage()
        err = float(np.max(np.abs(psi - ref)))
        infid = max(0.0, 1 - abs(np.vdot(ref, psi)) ** 2)
        print(f"width {width}: max-abs {err:.3e}, infidelity {infid:.3e}")
        assert err < 1e-12, (width, err)
        g.close()


def test_toffoli_circuit_c64_fused_against_oracle_24_qubits():
    """The precision workload's circuit (per layer: Haar U(2) on every qubit, CZ /
    CNOT on a random pairing, one Toffoli / CCZ / Fredkin) at 24 qubits and
    40 layers: the complex64 register fused to width-5 blocks (FP32 MFMA)
    against the oracle's FP64 gate-by-gate run from |0...0>. Max-abs and
    infidelity are printed; bars 1e-6 and 1e-9."""
    import bench

    n = 24
    seq = bench.random_toffoli_circuit(n, 40)
    o = OracleRegister(n)
    for gate, qs in seq:
        o.apply_gate(gate, *qs)
    ref = o.state()
    o.close()
    g = QubitRegister(n, precision="c64")
    g.apply_circuit(seq, fuse_width=5)
    psi = g.getRegisterStorage()
    err = float(np.max(np.abs(psi - ref)))
    infid = max(0.0, 1 - abs(np.vdot(ref, psi)) ** 2 / (np.vdot(psi, psi).real * np.vdot(ref, ref).real))
    print(f"24 qubits, {len(seq)} gates, complex64 fused width 5 vs oracle: max-abs {err:.3e}, infidelity {infid:.3e}, "
          f"max |amplitude| {np.max(np.abs(ref)):.3e}")
    assert err < 1e-6
    assert infid < 1e-9
    g.close()


def test_fusion_width_autotuner():
    """QCS_FUSE_AUTO: the measured cost table is positive and grows with the
    block width; the auto-fused layered circuit equals the oracle at 1e-12;
    and the width follows the table: with dense blocks priced as cheap as a
    gate wide blocks (4 or 5 qubits) win, with dense blocks priced 100 times a gate no
    fusion does."""
    import ctypes as C

    import bench
    from qcsim_amd._lib import QCS_COMPLEX128, check, lib

    n = 20
    seq = bench.random_layered_circuit(n, 12)
    reg = QubitRegister(n)
    t = reg.autotune_fusion()
    print("fusion cost table (ms):", t)
    assert np.all(t > 0) and t[5] >= t[2] * 0.5
    reg.apply_circuit(seq, fuse_width="auto")
    w = reg.last_fuse_width()
    assert w in (0, 2, 3, 4, 5)
    o = OracleRegister(n)
    for gate, qs in seq:
        o.apply_gate(gate, *qs)
    err = float(np.max(np.abs(reg.getRegisterStorage() - o.state())))
    print(f"auto width {w}: max-abs vs oracle {err:.2e}")
    assert err < 1e-12
    cheap = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    check(lib().qcs_set_fusion_costs(QCS_COMPLEX128, n, cheap.ctypes.data_as(C.POINTER(C.c_double))))
    reg.apply_circuit(seq, fuse_width="auto")
    assert reg.last_fuse_width() >= 4
    dear = np.array([1.0, 1.0, 100.0, 100.0, 100.0, 100.0])
    check(lib().qcs_set_fusion_costs(QCS_COMPLEX128, n, dear.ctypes.data_as(C.POINTER(C.c_double))))
    reg.apply_circuit(seq, fuse_width="auto")
    assert reg.last_fuse_width() == 0
    check(lib().qcs_set_fusion_costs(QCS_COMPLEX128, n, t.ctypes.data_as(C.POINTER(C.c_double))))
    reg.close()
    o.close()
