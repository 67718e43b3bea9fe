"""Circuit files: gate sequences (and optionally their results) shared by the
oracle, the engine, the C++ façade and test fixtures.

A gate is stored as the engine's `qcs_gate` describes it (include/qcsim_amd.h):
arity k, the seven reference traits packed as QCS_TRAIT_* bits, the qubit
list in ApplyGate order (qubit, controllingQubit1, controllingQubit2) or,
for dense blocks (flags & QCS_GATE_DENSE), block bit i -> qubits[i], and the
full 2^k x 2^k complex matrix. A reference `AppliedGate` list
(SimpleGates.h:439-578: matrix + q1, q2, q3, no traits) is the traits = 0
case.

Two encodings:

* JSON (``.json``): ``{"format": "qcsim-circuit", "version": 1, "qubits": n,
  "gates": [{"k", "traits", "flags", "qubits", "matrix": [[re, im], ...],
  "name"?}], "results"?: {...}, "meta"?: {...}}``; the matrix is row-major,
  values are written as Python's shortest round-trip repr (exact).
* Binary (``.qcsc``), little-endian: the 4 bytes ``QCSC``, uint32 version
  (1), uint32 qubits, uint32 gate count; then per gate uint32 k, traits,
  flags, qubits[5], followed by 2 * 4^k float64 (re, im row-major). The C++
  reader is include/QCSim/CircuitFile.h.
"""
from __future__ import annotations

import json
import struct
from dataclasses import dataclass, field

import numpy as np

from . import gates as G

MAGIC = b"QCSC"
VERSION = 1
GATE_DENSE = 0x1


class MatrixGate(G.QuantumGateWithOp):
    """A gate read back from a file: its matrix with the stored traits and
    flags (QCS_GATE_DENSE for fused blocks)."""

    def __init__(self, m, traits=0, name=None, flags=0):
        super().__init__(np.asarray(m, dtype=np.complex128))
        self.traits = int(traits)
        self.name = name
        self.flags = int(flags)

    def getQubitsNumber(self):
        return int(round(np.log2(self._m.shape[0])))


@dataclass
class Circuit:
    qubits: int
    gates: list  # (gate, [qubits]) pairs, as QubitRegister.apply_circuit takes them
    flags: list = field(default_factory=list)  # per gate, QCS_GATE_DENSE or 0
    results: dict = field(default_factory=dict)
    meta: dict = field(default_factory=dict)

    def applied_gates(self):
        """The reference's AppliedGate list (no traits), k <= 3 only."""
        out = []
        for gate, qs in self.gates:
            q = list(qs) + [0] * (3 - len(qs))
            out.append(G.AppliedGate(gate.getRawOperatorMatrix(), q[0], q[1], q[2]))
        return out


def _record(gate, qubits, flags):
    m = np.asarray(gate.getRawOperatorMatrix(), dtype=np.complex128)
    k = len(qubits)
    if m.shape != (1 << k, 1 << k):
        raise ValueError(f"matrix of shape {m.shape} for {k} qubits")
    if not 1 <= k <= 5 or len(set(qubits)) != k:
        raise ValueError("a gate acts on 1..5 distinct qubits")
    return k, int(getattr(gate, "traits", 0)), int(flags), [int(q) for q in qubits], m


def _check_qubits(n, k, qubits):
    if any(q >= n for q in qubits):
        raise ValueError(f"qubit out of range for a {n}-qubit circuit")


def save_json(path, circuit: Circuit):
    gates = []
    flags = circuit.flags or [getattr(g, "flags", 0) for g, _ in circuit.gates]
    for (gate, qs), fl in zip(circuit.gates, flags):
        k, traits, fl, qs, m = _record(gate, qs, fl)
        _check_qubits(circuit.qubits, k, qs)
        entry = {"k": k, "traits": traits, "flags": fl, "qubits": qs,
                 "matrix": [[float(z.real), float(z.imag)] for z in m.ravel()]}
        name = getattr(gate, "name", None) or type(gate).__name__
        entry["name"] = name
        gates.append(entry)
    doc = {"format": "qcsim-circuit", "version": VERSION, "qubits": circuit.qubits, "gates": gates}
    if circuit.results:
        doc["results"] = circuit.results
    if circuit.meta:
        doc["meta"] = circuit.meta
    with open(path, "w") as f:
        json.dump(doc, f)


def load_json(path) -> Circuit:
    with open(path) as f:
        doc = json.load(f)
    if doc.get("format") != "qcsim-circuit" or doc.get("version") != VERSION:
        raise ValueError("not a version-1 qcsim circuit file")
    n = int(doc["qubits"])
    gates, flags = [], []
    for e in doc["gates"]:
        k = int(e["k"])
        a = np.asarray(e["matrix"], dtype=np.float64)
        if a.shape != (4**k, 2) or len(e["qubits"]) != k:
            raise ValueError("malformed gate entry")
        _check_qubits(n, k, e["qubits"])
        m = (a[:, 0] + 1j * a[:, 1]).reshape(1 << k, 1 << k)
        gates.append((MatrixGate(m, e.get("traits", 0), e.get("name"), e.get("flags", 0)),
                      [int(q) for q in e["qubits"]]))
        flags.append(int(e.get("flags", 0)))
    return Circuit(n, gates, flags, doc.get("results", {}), doc.get("meta", {}))


def save_binary(path, circuit: Circuit):
    flags = circuit.flags or [getattr(g, "flags", 0) for g, _ in circuit.gates]
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<III", VERSION, circuit.qubits, len(circuit.gates)))
        for (gate, qs), fl in zip(circuit.gates, flags):
            k, traits, fl, qs, m = _record(gate, qs, fl)
            _check_qubits(circuit.qubits, k, qs)
            f.write(struct.pack("<8I", k, traits, fl, *(qs + [0] * (5 - k))))
            f.write(np.ascontiguousarray(m.ravel()).view(np.float64).astype("<f8").tobytes())


def load_binary(path) -> Circuit:
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != MAGIC:
        raise ValueError("not a qcsim binary circuit file")
    version, n, count = struct.unpack_from("<III", data, 4)
    if version != VERSION:
        raise ValueError(f"unsupported version {version}")
    off = 16
    gates, flags = [], []
    for _ in range(count):
        if off + 32 > len(data):
            raise ValueError("truncated circuit file")
        k, traits, fl, *qs = struct.unpack_from("<8I", data, off)
        off += 32
        if not 1 <= k <= 5:
            raise ValueError("malformed gate record")
        if off + 16 * 4**k > len(data):
            raise ValueError("truncated circuit file")
        vals = np.frombuffer(data, dtype="<f8", count=2 * 4**k, offset=off)
        off += 16 * 4**k
        m = vals.view(np.complex128).reshape(1 << k, 1 << k).copy()
        _check_qubits(n, k, qs[:k])
        gates.append((MatrixGate(m, traits, None, fl), list(qs[:k])))
        flags.append(fl)
    if off != len(data):
        raise ValueError("trailing bytes in circuit file")
    return Circuit(n, gates, flags)


def save(path, circuit: Circuit):
    (save_binary if str(path).endswith(".qcsc") else save_json)(path, circuit)


def load(path) -> Circuit:
    return (load_binary if str(path).endswith(".qcsc") else load_json)(path)
