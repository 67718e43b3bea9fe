"""Pauli strings and their exact expectation values on the engine.

A string is stored as two bit masks over the register's qubits, P =
i^{|x & z|} X^x Z^z: qubit q carries X when only bit q of ``x`` is set, Z
when only bit q of ``z`` is, Y = iXZ when both are and the identity when
neither is. ``QubitRegister.pauli_expectation`` returns <psi|P|psi> for
many strings in FP64 (qcs_pauli_expectation).

``PauliString`` mirrors PauliString::PauliString (PauliString.h:7-60): one
operator per qubit, Z by default (there is no identity there), plus a real
coefficient. ``vqe_term_energy`` is the exact value of what
VQE::PauliStringVQE::EstimateEnergy samples (VQE.h:116-142): that routine
rotates X terms with Ry(-pi/2) and Y terms with Rx(pi/2) (VQE.h:33-44),
samples the parity of all qubits and scales by the coefficient, which
measures coefficient * <psi| (x) op_q |psi>.
"""
from __future__ import annotations

import numpy as np

OP_Z, OP_X, OP_Y = 0, 1, 2  # PauliString::PauliOp values (PauliString.h:9-14)


class PauliString:
    """PauliString::PauliString: per-qubit Z/X/Y operators and a coefficient."""

    def __init__(self, nrQubits: int = 2):
        self.ops = [OP_Z] * nrQubits
        self.coeff = 1.0

    def setOperatorForQubit(self, qubit: int, op: int):
        if qubit < len(self.ops):
            self.ops[qubit] = op

    def getOperatorForQubit(self, qubit: int) -> int:
        return self.ops[qubit] if qubit < len(self.ops) else OP_Z

    def getCoefficient(self) -> float:
        return self.coeff

    def setCoefficient(self, c: float):
        self.coeff = c

    def masks(self) -> tuple[int, int]:
        """(x, z) masks of the operator on all qubits."""
        x = z = 0
        for q, op in enumerate(self.ops):
            if op in (OP_X, OP_Y):
                x |= 1 << q
            if op in (OP_Z, OP_Y):
                z |= 1 << q
        return x, z


def masks_from_label(label: str) -> tuple[int, int]:
    """Masks of a label such as "XIZY", whose character k is qubit k."""
    x = z = 0
    for q, c in enumerate(label.upper()):
        if c in "XY":
            x |= 1 << q
        if c in "ZY":
            z |= 1 << q
        if c not in "IXYZ":
            raise ValueError(f"bad Pauli label {label!r}")
    return x, z


def _classify(m: np.ndarray):
    """(c, x, z) with m == c * P(x, z) exactly, or None."""
    a, b, c, d = m[0, 0], m[0, 1], m[1, 0], m[1, 1]
    if b == 0 and c == 0:
        if a == d:
            return a, 0, 0
        if a == -d:
            return a, 0, 1
        return None
    if a == 0 and d == 0:
        if b == c:
            return b, 1, 0
        if c == -b:
            return -1j * c, 1, 1  # c * Y has m10 = i c
    return None


def gates_as_pauli(gates, nqubits: int):
    """For a list of recorded gates (ExpectationValue's argument,
    QubitRegister.h:646-660) that are all single-qubit scalar multiples of
    I, X, Y or Z: (coefficient, x mask, z mask) with the product of the
    gates equal to coefficient * P(x, z). Otherwise None."""
    per = {}
    for g in gates:
        if g.getQubitsNumber() != 1:
            return None
        q = int(g.q1)
        if not 0 <= q < nqubits:
            return None
        m = np.asarray(g.getRawOperatorMatrix(), dtype=np.complex128)
        per[q] = m @ per[q] if q in per else m.copy()  # later gates act after
    coef = 1.0 + 0j
    x = z = 0
    for q, m in per.items():
        r = _classify(m)
        if r is None:
            return None
        c, xb, zb = r
        coef *= c
        x |= xb << q
        z |= zb << q
    return coef, x, z



def vqe_term_energy(reg, term: PauliString) -> float:
    """Exact value of VQE::PauliStringVQE::EstimateEnergy's estimate for a
    state already prepared in ``reg`` (before the change of basis): the
    coefficient times <psi| (x)_q op_q |psi>, from one FP64 reduction
    instead of sampling."""
    x, z = term.masks()
    return term.coeff * float(reg.pauli_expectation([x], [z])[0])
