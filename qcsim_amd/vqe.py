"""Batched VQE energies on the HIP engine (qcs_vqe_*).

Each instance is a register of n <= 12 qubits prepared by the reference's
hardware-efficient ansatz (VQE::PauliStringVQE::Ansatz, VQE.h:78-114) and
scored with the exact FP64 energy sum_t c_t <psi|P_t|psi> of a Pauli-string
Hamiltonian: the value that VariationalQuantumEigensolver::EstimateEnergy
(VQE.h:204-212) estimates from 10000 samples per term. Instances of uneven
size and parameter count run together, one launch per size class.

``ansatz_gates`` gives the same ansatz as a gate list, for a single register
(QubitRegister.ApplyGate) of any size.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gates as G
from ._lib import check, lib
from .pauli import PauliString

_PD = C.POINTER(C.c_double)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)

MAX_QUBITS = 12


def ansatz_gates(n: int, params):
    """VQE.h:78-114 as a list of (gate, [qubit, control]): Ry then Rz per
    qubit while two parameters remain, then alternating CNOT ladders each
    followed by another Ry / Rz layer; no gates when len(params) < n."""
    params = list(params)
    P = len(params)
    seq = []
    if P < n:
        return seq
    pos = 0

    def layer():
        nonlocal pos
        q = 0
        while q < n and pos + 1 < P:
            seq.append((G.RyGate(params[pos]), [q]))
            seq.append((G.RzGate(params[pos + 1]), [q]))
            pos += 2
            q += 1

    layer()
    flip = False
    while pos + 1 < P:
        if flip:
            for q in range(n - 1, 0, -1):
                seq.append((G.CNOTGate(), [q - 1, q]))
            if n > 2:
                seq.append((G.CNOTGate(), [n - 1, 0]))
        else:
            for q in range(n - 1):
                seq.append((G.CNOTGate(), [q + 1, q]))
            if n > 2:
                seq.append((G.CNOTGate(), [0, n - 1]))
        flip = not flip
        layer()
    return seq


def hamiltonian(terms):
    """(x masks, z masks, coefficients) of a list of PauliString terms
    (PauliString.h; the coefficient is the term's)."""
    xs, zs, cs = [], [], []
    for t in terms:
        x, z = t.masks()
        xs.append(x)
        zs.append(z)
        cs.append(float(t.getCoefficient()))
    return xs, zs, cs


class VqeBatch:
    """Instances (n, (xs, zs, cs)) evaluated together: qcs_vqe_*."""

    def __init__(self, instances):
        self.instances = list(instances)
        nq = np.array([n for n, _ in self.instances], dtype=np.uint32)
        nt = np.array([len(h[0]) for _, h in self.instances], dtype=np.uint32)
        xs = np.array([x for _, h in self.instances for x in h[0]] or [0], dtype=np.uint64)
        zs = np.array([z for _, h in self.instances for z in h[1]] or [0], dtype=np.uint64)
        cs = np.array([c for _, h in self.instances for c in h[2]] or [0.0], dtype=np.float64)
        self._keep = (nq, nt, xs, zs, cs)
        h = C.c_void_p()
        check(lib().qcs_vqe_create(len(self.instances), nq.ctypes.data_as(_PU32), nt.ctypes.data_as(_PU32),
                                   xs.ctypes.data_as(_PU64), zs.ctypes.data_as(_PU64), cs.ctypes.data_as(_PD),
                                   C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().qcs_vqe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, params) -> np.ndarray:
        """params: one sequence of ansatz parameters per instance (lengths may
        differ). Returns the energies."""
        counts = np.array([len(p) for p in params], dtype=np.uint32)
        assert counts.shape[0] == len(self.instances)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64) for p in params])
                                    if counts.sum() else np.zeros(1))
        out = np.zeros(len(self.instances))
        check(lib().qcs_vqe_evaluate(self._h, counts.ctypes.data_as(_PU32), flat.ctypes.data_as(_PD),
                                     out.ctypes.data_as(_PD)))
        return out

    def state(self, i: int) -> np.ndarray:
        out = np.zeros(2 << self.instances[i][0])
        check(lib().qcs_vqe_get_state(self._h, i, out.ctypes.data_as(_PD)))
        return out.view(np.complex128)

    def last_ms(self) -> float:
        ms = C.c_float()
        check(lib().qcs_vqe_last_ms(self._h, C.byref(ms)))
        return float(ms.value)


__all__ = ["MAX_QUBITS", "PauliString", "VqeBatch", "ansatz_gates", "hamiltonian"]
