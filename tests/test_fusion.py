"""The fusion pass (qcs_fuse, host-only) against the oracle: applying the
fused blocks equals applying the original sequence gate by gate; blocks
respect the width limit; single-gate blocks keep the original gate."""
import ctypes as C

import numpy as np
import pytest

from gatecatalog import catalog
from oracle.oracle import OracleRegister
from qcsim_amd import _lib
from qcsim_amd import gates as G
from refmath import random_state


def to_struct(seq):
    arr = (_lib.QcsGate * len(seq))()
    for i, (gate, qs) in enumerate(seq):
        m = np.ascontiguousarray(gate.getRawOperatorMatrix(), dtype=np.complex128).view(np.float64).reshape(-1)
        arr[i].k = gate.getQubitsNumber()
        arr[i].traits = gate.traits
        for j, q in enumerate(qs):
            arr[i].qubits[j] = q
        C.memmove(arr[i].mat, m.ctypes.data, m.nbytes)
    return arr


def fuse(arr, width):
    lib = _lib.load()
    out = (_lib.QcsGate * len(arr))()
    nout = C.c_uint32()
    _lib.check(lib.qcs_fuse(arr, len(arr), width, out, C.byref(nout)))
    return out, nout.value


def random_circuit(rng, n, ngates):
    one, two, three = catalog(int(rng.integers(1000)))
    fams = one + two + three
    seq = []
    for _ in range(ngates):
        name, gate = fams[int(rng.integers(len(fams)))]
        k = gate.getQubitsNumber()
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
        seq.append((gate, qs))
    return seq


def apply_struct(reg, g):
    k = g.k
    m = np.frombuffer(bytes(g.mat), dtype=np.float64)[: 2 * 4**k].view(np.complex128).reshape(1 << k, 1 << k)
    if (g.flags & _lib.QCS_GATE_DENSE) or k > 3:
        reg.apply_dense([g.qubits[j] for j in range(k)], m)
    else:
        reg.apply_raw(k, g.traits, g.qubits[0], g.qubits[1], g.qubits[2], m)


@pytest.mark.parametrize("width", [1, 2, 3, 4, 5])
def test_fused_blocks_equal_sequence(width):
    n = 7
    rng = np.random.default_rng(width)
    for trial in range(4):
        seq = random_circuit(rng, n, 60)
        psi = random_state(rng, n)
        ref = OracleRegister(n)
        ref.set_state(psi)
        for gate, qs in seq:
            ref.apply_gate(gate, *qs)
        arr = to_struct(seq)
        out, nout = fuse(arr, width)
        assert nout <= len(seq)
        fr = OracleRegister(n)
        fr.set_state(psi)
        for i in range(nout):
            g = out[i]
            if g.flags & _lib.QCS_GATE_DENSE:
                assert g.k <= width
            apply_struct(fr, g)
        assert np.max(np.abs(fr.state() - ref.state())) < 1e-12


def test_layered_circuit_fuses_and_keeps_single_gates():
    n = 10
    rng = np.random.default_rng(3)
    seq = []
    for _ in range(3):
        for q in range(n):
            seq.append((G.SingleQubitGate(G.haar_u2(rng)), [q]))
        perm = rng.permutation(n)
        for a, b in zip(perm[0::2], perm[1::2]):
            seq.append((G.CNOTGate(), [int(a), int(b)]))
    out, nout = fuse(to_struct(seq), 5)
    assert nout < len(seq) // 2
    # a lone CZ stays a CZ (keeps its quarter-traffic kernel)
    out1, n1 = fuse(to_struct([(G.ControlledZGate(), [2, 7])]), 5)
    assert n1 == 1 and out1[0].flags == 0 and out1[0].traits == G.ControlledZGate.traits


def test_layered_30_qubit_circuit_block_count():
    """60 layers on 30 qubits (a Haar U(2) on every qubit, then CNOT / CZ on
    a random pairing): the greedy block growth takes 211 blocks of <= 5
    qubits with this seed (gate-order growth took about 380); the bound leaves a
    little room and catches a pass that stops taking gates past the first
    blocked one."""
    n, layers = 30, 60
    rng = np.random.default_rng(11)
    seq = []
    for _ in range(layers):
        for q in range(n):
            seq.append((G.SingleQubitGate(G.haar_u2(rng)), [q]))
        perm = rng.permutation(n)
        for a, b in zip(perm[0::2], perm[1::2]):
            gate = G.CNOTGate() if rng.integers(2) else G.ControlledZGate()
            seq.append((gate, [int(a), int(b)]))
    out, nout = fuse(to_struct(seq), 5)
    assert nout <= 225
    assert all(out[i].k <= 5 for i in range(nout))
    # the blocks of a slice (the gates on qubits 0..7) reproduce it exactly
    small = [(g, qs) for g, qs in seq[:200] if max(qs) < 8]
    psi = random_state(rng, 8)
    ref, fr = OracleRegister(8), OracleRegister(8)
    ref.set_state(psi)
    fr.set_state(psi)
    for gate, qs in small:
        ref.apply_gate(gate, *qs)
    o2, n2 = fuse(to_struct(small), 5)
    for i in range(n2):
        apply_struct(fr, o2[i])
    assert np.max(np.abs(fr.state() - ref.state())) < 1e-12
