"""Compiled gate programs (qcs_program_*): the recorded-gate replay of
QubitRegister::Compute / Uncompute (QubitRegister.h:553-590) captured as a
hipGraph. A replay must equal the eager launches bit for bit, replays must
stay equal with eager work between them, and the adjoint program must undo
the forward one."""
import numpy as np
import pytest

from gatecatalog import catalog
from qcsim_amd import QubitRegister
from qcsim_amd import gates as G
from refmath import random_state

pytestmark = pytest.mark.gpu


def random_sequence(rng, n, count):
    one, two, three = catalog(7)
    fams = one + two + three
    seq = []
    for _ in range(count):
        _, gate = fams[int(rng.integers(len(fams)))]
        seq.append((gate, [int(x) for x in rng.choice(n, size=gate.getQubitsNumber(), replace=False)]))
    return seq


def fresh(n, psi, precision="c128"):
    r = QubitRegister(n, precision=precision)
    r.setRegisterStorageFastNoNormalize(psi.copy())
    return r


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("width", [0, 3, 5])
def test_program_replays_equal_eager_launches(precision, width):
    n = 14
    rng = np.random.default_rng(10 + width)
    seq = random_sequence(rng, n, 160)
    psi = random_state(rng, n)
    g = fresh(n, psi, precision)
    e = fresh(n, psi, precision)
    prog = g.compile_program(seq, fuse_width=width)
    assert prog.launches() >= 1
    h = G.HadamardGate()
    for rep in range(3):
        prog.run()
        e.apply_circuit(seq, fuse_width=width)
        assert np.array_equal(g.getRegisterStorage(), e.getRegisterStorage()), rep
        # eager work between replays
        g.ApplyGate(h, rep)
        e.ApplyGate(h, rep)
    prog.close()
    g.close()
    e.close()


@pytest.mark.parametrize("width", [0, 4, 5])
def test_adjoint_program_undoes_the_forward_one(width):
    n = 13
    rng = np.random.default_rng(20 + width)
    seq = random_sequence(rng, n, 120)
    psi = random_state(rng, n)
    r = fresh(n, psi)
    fwd = r.compile_program(seq, fuse_width=width)
    inv = r.compile_program(seq, fuse_width=width, adjoint=True)
    fwd.run()
    assert np.max(np.abs(r.getRegisterStorage() - psi)) > 1e-3
    inv.run()
    assert np.max(np.abs(r.getRegisterStorage() - psi)) < 1e-12
    # the adjoint program is the eager inverse sequence
    e = fresh(n, psi)
    fwd.run()
    for gate, qs in reversed(seq):
        m = gate.getRawOperatorMatrix().conj().T
        e.ApplyGate(G.AppliedGate(m, *(qs + [0] * (3 - len(qs)))))
    e_state = e.getRegisterStorage()
    r2 = fresh(n, psi)
    inv2 = r2.compile_program(seq, fuse_width=0, adjoint=True)
    inv2.run()
    assert np.max(np.abs(r2.getRegisterStorage() - e_state)) < 1e-12
    for x in (fwd, inv, inv2):
        x.close()
    for x in (r, e, r2):
        x.close()


def test_program_is_bound_to_its_register():
    from qcsim_amd._lib import check, lib

    a, b = QubitRegister(6), QubitRegister(6)
    prog = a.compile_program([(G.HadamardGate(), [0])])
    with pytest.raises(ValueError, match="another register"):
        check(lib().qcs_program_run(b._h, prog._h))
    with pytest.raises(ValueError, match="out of range"):
        a.compile_program([(G.HadamardGate(), [6])])
    prog.close()
    a.close()
    b.close()


def test_program_replay_timing_small_state():
    """16 qubits, 600 gates: one graph launch against 600 launches (printed,
    and the replay must not be slower)."""
    import time

    n = 16
    rng = np.random.default_rng(5)
    seq = [(G.SingleQubitGate(G.haar_u2(rng)), [int(rng.integers(n))]) for _ in range(600)]
    r = QubitRegister(n)
    prog = r.compile_program(seq)
    for _ in range(2):
        prog.run()
        r.apply_circuit(seq)
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        r.apply_circuit(seq)
    r.synchronize()
    t_eager = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    for _ in range(5):
        prog.run()
    r.synchronize()
    t_graph = (time.perf_counter() - t0) / 5
    print(f"600 gates at 16 qubits: eager {t_eager * 1e3:.2f} ms, graph {t_graph * 1e3:.2f} ms")
    assert t_graph < 1.2 * t_eager
    prog.close()
    r.close()


def test_program_follows_state_buffer_swaps():
    """RestoreStateDestructive exchanges the register's state buffer
    (qcs_swap_state) and releases the other one: the program must run on the
    buffer the register holds now."""
    n = 12
    rng = np.random.default_rng(31)
    seq = random_sequence(rng, n, 60)
    psi = random_state(rng, n)
    r = fresh(n, psi)
    e = fresh(n, psi)
    prog = r.compile_program(seq, fuse_width=4)
    r.SaveState()
    prog.run()
    r.RestoreStateDestructive()
    prog.run()
    e.apply_circuit(seq, fuse_width=4)
    assert np.array_equal(r.getRegisterStorage(), e.getRegisterStorage())
    prog.close()
    r.close()
    e.close()


def test_circuit_file_runs_on_the_engine(tmp_path):
    """A circuit written to a binary file and read back runs bit for bit as
    the original list, eagerly and as a program."""
    from qcsim_amd import circuit as cf

    n = 12
    rng = np.random.default_rng(41)
    seq = random_sequence(rng, n, 80)
    path = str(tmp_path / "c.qcsc")
    cf.save(path, cf.Circuit(n, seq))
    loaded = cf.load(path)
    psi = random_state(rng, n)
    a, b, c = fresh(n, psi), fresh(n, psi), fresh(n, psi)
    a.apply_circuit(seq)
    b.apply_circuit(loaded.gates)
    prog = c.compile_program(loaded.gates, fuse_width=4)
    prog.run()
    assert np.array_equal(a.getRegisterStorage(), b.getRegisterStorage())
    assert np.max(np.abs(a.getRegisterStorage() - c.getRegisterStorage())) < 1e-12
    prog.close()
    for r in (a, b, c):
        r.close()


def oracle_run(n, psi, seq, adjoint=False):
    """The oracle's gate-by-gate run of seq, or with adjoint=True of its
    inverse: the gates reversed with conjugate-transposed matrices and no
    traits, which is what the reference's Uncompute applies
    (QubitRegister.h:571-590)."""
    from oracle.oracle import OracleRegister

    ref = OracleRegister(n)
    ref.set_state(psi)
    if not adjoint:
        for gate, qs in seq:
            ref.apply_gate(gate, *qs)
    else:
        for gate, qs in reversed(seq):
            qq = list(qs) + [0] * (3 - len(qs))
            ref.apply_raw(gate.getQubitsNumber(), 0, qq[0], qq[1], qq[2], gate.getRawOperatorMatrix().conj().T)
    out = ref.state()
    ref.close()
    return out


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("width", [0, 4, 5])
def test_program_matches_oracle(width, adjoint):
    """Forward (Compute) and adjoint (Uncompute) programs, unfused and fused,
    against the oracle's gate-by-gate run at 1e-12."""
    n = 14
    rng = np.random.default_rng(50 + width + 7 * adjoint)
    seq = random_sequence(rng, n, 160)
    psi = random_state(rng, n)
    r = fresh(n, psi)
    prog = r.compile_program(seq, fuse_width=width, adjoint=adjoint)
    prog.run()
    err = float(np.max(np.abs(r.getRegisterStorage() - oracle_run(n, psi, seq, adjoint))))
    print(f"program width {width} adjoint {adjoint}: max-abs vs oracle {err:.2e}")
    assert err < 1e-12
    prog.close()
    r.close()


def qft_applied_gates(n):
    """The reference's QFT sequence on qubits 0..n-1 with swaps
    (QuantumFourierTransform.h:35-60, QubitsSwapper.h:23-33), as traitless
    recorded gates, which is how Compute/Uncompute replays it."""
    import math

    seq = []
    h = G.HadamardGate().getRawOperatorMatrix()
    seq.append((G.AppliedGate(h), [n - 1]))
    for cur in range(n - 1, 0, -1):
        phase = math.pi / 2
        for ctrl in range(cur - 1, -1, -1):
            seq.append((G.AppliedGate(G.ControlledPhaseShiftGate(phase).getRawOperatorMatrix()), [cur, ctrl]))
            phase *= 0.5
        seq.append((G.AppliedGate(h), [cur - 1]))
    sw = G.SwapGate().getRawOperatorMatrix()
    a, b = 0, n - 1
    while a < b:
        seq.append((G.AppliedGate(sw), [a, b]))
        a += 1
        b -= 1
    return seq


@pytest.mark.parametrize("adjoint", [False, True])
def test_program_with_many_generic_gates(adjoint):
    """A 17-qubit QFT recorded as traitless gates (136 controlled phases and 8
    swaps run as generic 4x4 blocks, each staging its matrix) compiles and
    replays like the oracle; so does a run of 220 generic 2- and 3-qubit
    gates."""
    n = 17
    rng = np.random.default_rng(61)
    psi = random_state(rng, n)
    seq = qft_applied_gates(n)
    assert sum(g.getQubitsNumber() == 2 for g, _ in seq) == 144
    r = fresh(n, psi)
    prog = r.compile_program(seq, adjoint=adjoint)
    prog.run()
    err = float(np.max(np.abs(r.getRegisterStorage() - oracle_run(n, psi, seq, adjoint))))
    print(f"17-qubit recorded QFT adjoint {adjoint}: max-abs vs oracle {err:.2e}")
    assert err < 1e-12
    prog.close()
    from refmath import random_unitary

    seq = []
    for i in range(220):
        k = 2 if i % 4 else 3
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
        seq.append((G.AppliedGate(random_unitary(rng, 1 << k)), qs))
    r.setRegisterStorageFastNoNormalize(psi.copy())
    prog = r.compile_program(seq, adjoint=adjoint)
    prog.run()
    err = float(np.max(np.abs(r.getRegisterStorage() - oracle_run(n, psi, seq, adjoint))))
    print(f"220 generic gates adjoint {adjoint}: max-abs vs oracle {err:.2e}")
    assert err < 1e-12
    prog.close()
    r.close()
