"""Circuit files (qcsim_amd/circuit.py, include/QCSim/CircuitFile.h): exact
round trips in both encodings, the byte layout shared with the C++ reader,
the oracle running a loaded circuit exactly as the original, and the errors
on malformed files."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from gatecatalog import catalog
from oracle.oracle import OracleRegister
from qcsim_amd import circuit as cf
from qcsim_amd import gates as G
from qcsim_amd._lib import QcsGate, check, lib
from refmath import random_state, random_unitary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sample_circuit(seed=3, n=8, count=60, dense=True):
    rng = np.random.default_rng(seed)
    one, two, three = catalog(seed)
    fams = one + two + three
    gates, flags = [], []
    for _ in range(count):
        _, gate = fams[int(rng.integers(len(fams)))]
        gates.append((gate, [int(x) for x in rng.choice(n, size=gate.getQubitsNumber(), replace=False)]))
        flags.append(0)
    if dense:
        for k in (4, 5):
            qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
            gates.append((cf.MatrixGate(random_unitary(rng, 1 << k), 0, "dense", cf.GATE_DENSE), qs))
            flags.append(cf.GATE_DENSE)
    return cf.Circuit(n, gates, flags, results={"note": "sample"}, meta={"seed": seed})


def same_gates(a, b):
    assert a.qubits == b.qubits and len(a.gates) == len(b.gates)
    for (ga, qa), (gb, qb), fa, fb in zip(a.gates, b.gates, a.flags, b.flags):
        assert qa == qb and fa == fb
        assert getattr(ga, "traits", 0) == gb.traits
        assert np.array_equal(np.asarray(ga.getRawOperatorMatrix()), gb.getRawOperatorMatrix())


@pytest.mark.parametrize("ext", [".json", ".qcsc"])
def test_round_trip_is_exact(tmp_path, ext):
    c = sample_circuit()
    path = str(tmp_path / ("c" + ext))
    cf.save(path, c)
    d = cf.load(path)
    same_gates(c, d)
    if ext == ".json":
        assert d.results == c.results and d.meta == c.meta


def test_binary_layout_by_hand(tmp_path):
    """Header and one record written by hand, per the documented layout."""
    import struct

    path = str(tmp_path / "h.qcsc")
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    with open(path, "wb") as f:
        f.write(b"QCSC" + struct.pack("<III", 1, 3, 1))
        f.write(struct.pack("<8I", 1, 0, 0, 2, 0, 0, 0, 0))
        f.write(h.astype(np.complex128).ravel().view(np.float64).astype("<f8").tobytes())
    c = cf.load(path)
    assert c.qubits == 3 and c.gates[0][1] == [2]
    assert np.array_equal(c.gates[0][0].getRawOperatorMatrix(), h.astype(np.complex128))


def test_oracle_runs_a_loaded_circuit_exactly(tmp_path):
    c = sample_circuit(dense=False)
    path = str(tmp_path / "c.qcsc")
    cf.save(path, c)
    d = cf.load(path)
    psi = random_state(np.random.default_rng(1), c.qubits)
    a, b = OracleRegister(c.qubits), OracleRegister(c.qubits)
    a.set_state(psi)
    b.set_state(psi)
    for gate, qs in c.gates:
        a.apply_gate(gate, *qs)
    for gate, qs in d.gates:
        b.apply_gate(gate, *qs)
    assert np.array_equal(a.state(), b.state())


def test_fusion_output_round_trips(tmp_path):
    """Blocks from qcs_fuse (host only) written and read back keep their
    dense flag and matrices."""
    c = sample_circuit(dense=False)
    arr = (QcsGate * len(c.gates))()
    for i, (gate, qs) in enumerate(c.gates):
        m = np.ascontiguousarray(gate.getRawOperatorMatrix(), dtype=np.complex128).ravel().view(np.float64)
        arr[i].k, arr[i].traits = len(qs), gate.traits
        for j, q in enumerate(qs):
            arr[i].qubits[j] = q
        C.memmove(arr[i].mat, m.ctypes.data, m.nbytes)
    out = (QcsGate * len(c.gates))()
    nout = C.c_uint32()
    check(lib().qcs_fuse(arr, len(c.gates), 5, out, C.byref(nout)))
    gates, flags = [], []
    for i in range(nout.value):
        g = out[i]
        k = g.k
        m = np.array(g.mat[: 2 * 4**k]).view(np.complex128).reshape(1 << k, 1 << k)
        gates.append((cf.MatrixGate(m, g.traits, None, g.flags), list(g.qubits[:k])))
        flags.append(g.flags)
    fused = cf.Circuit(c.qubits, gates, flags)
    path = str(tmp_path / "f.qcsc")
    cf.save(path, fused)
    same_gates(fused, cf.load(path))
    assert any(f & cf.GATE_DENSE for f in flags)


def test_cpp_reader_writes_back_identical_bytes(tmp_path):
    """include/QCSim/CircuitFile.h reads a file written here and writes it
    back byte for byte."""
    src = tmp_path / "rt.cpp"
    src.write_text(
        '#include "CircuitFile.h"\n'
        "int main(int argc, char** argv) {\n"
        "    QC::CircuitFile c = QC::ReadCircuitFile(argv[1]);\n"
        "    QC::WriteCircuitFile(argv[2], c);\n"
        "    return c.gates.size() > 0 ? 0 : 1;\n"
        "}\n"
    )
    exe = tmp_path / "rt"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include/QCSim", f"-I{ROOT}/include", str(src), "-o",
                    str(exe)], check=True)
    a, b = str(tmp_path / "a.qcsc"), str(tmp_path / "b.qcsc")
    cf.save(a, sample_circuit())
    subprocess.run([str(exe), a, b], check=True)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_malformed_files_are_rejected(tmp_path):
    c = sample_circuit(count=4, dense=False)
    good = str(tmp_path / "g.qcsc")
    cf.save(good, c)
    data = open(good, "rb").read()
    for bad, msg in ((b"XXXX" + data[4:], "not a qcsim"), (data[:-8], "truncated"), (data + b"\0", "trailing")):
        p = str(tmp_path / "bad.qcsc")
        open(p, "wb").write(bad)
        with pytest.raises(ValueError, match=msg):
            cf.load(p)
    with pytest.raises(ValueError, match="out of range"):
        cf.save(str(tmp_path / "x.json"), cf.Circuit(2, [(G.HadamardGate(), [3])]))
