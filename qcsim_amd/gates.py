"""Gate library mirroring the reference's QC::Gates classes.

Each gate carries its raw operator matrix (``getRawOperatorMatrix``) and the
trait bits of the reference's QuantumGate virtuals (SimpleGates.h:27-60):
isControlled, isControlQubit(1), isDiagonal, isAntidiagonal, isSwapGate,
IsISwapGate, IsISwapDagGate. The register dispatches on those traits exactly
like QubitRegister::ApplyGate (QubitRegister.h:434-486).

Matrices are the textbook ones the reference defines in SimpleGates.h:582-919
and QuantumGate.h:10-469. Local bit 0 of the matrix index is ``qubit``,
bit 1 ``controllingQubit1``, bit 2 ``controllingQubit2``.
"""
from __future__ import annotations

import cmath
import math

import numpy as np

T_CONTROLLED = 0x01
T_CTRL_Q1 = 0x02
T_DIAGONAL = 0x04
T_ANTIDIAGONAL = 0x08
T_SWAP = 0x10
T_ISWAP = 0x20
T_ISWAPDAG = 0x40

_S2 = 1.0 / math.sqrt(2.0)


class QuantumGateWithOp:
    """Raw matrix + trait bits (SimpleGates.h:63-132)."""

    traits = 0

    def __init__(self, matrix):
        m = np.array(matrix, dtype=np.complex128)
        assert m.ndim == 2 and m.shape[0] == m.shape[1]
        self._m = m

    def getRawOperatorMatrix(self) -> np.ndarray:
        return self._m

    def getQubitsNumber(self) -> int:
        return int(round(math.log2(self._m.shape[0])))

    def setOperator(self, m):
        self._m = np.array(m, dtype=np.complex128)

    # trait virtuals
    def isControlled(self):
        return bool(self.traits & T_CONTROLLED)

    def isControlQubit(self, q):
        if not self.isControlled():
            return False
        if q == 0:
            return True
        return q == 1 and bool(self.traits & T_CTRL_Q1)

    def isDiagonal(self):
        return bool(self.traits & T_DIAGONAL)

    def isAntidiagonal(self):
        return bool(self.traits & T_ANTIDIAGONAL)

    def isSwapGate(self):
        return bool(self.traits & T_SWAP)

    def IsISwapGate(self):
        return bool(self.traits & T_ISWAP)

    def IsISwapDagGate(self):
        return bool(self.traits & T_ISWAPDAG)


class SingleQubitGate(QuantumGateWithOp):
    def __init__(self, m=None):
        super().__init__(np.zeros((2, 2)) if m is None else m)


class TwoQubitsGate(QuantumGateWithOp):
    def __init__(self, m=None):
        super().__init__(np.eye(4) if m is None else m)


class ThreeQubitsGate(QuantumGateWithOp):
    def __init__(self, m=None):
        super().__init__(np.eye(8) if m is None else m)


class AppliedGate(QuantumGateWithOp):
    """Recorded gate: matrix + qubits, no traits (SimpleGates.h:439-578)."""

    def __init__(self, m, q1=0, q2=0, q3=0):
        super().__init__(m)
        self.q1, self.q2, self.q3 = q1, q2, q3

    def getQubit1(self):
        return self.q1

    def getQubit2(self):
        return self.q2

    def getQubit3(self):
        return self.q3


# ---- one-qubit gates (SimpleGates.h:582-919) --------------------------------
class HadamardGate(SingleQubitGate):
    def __init__(self):
        super().__init__([[_S2, _S2], [_S2, -_S2]])


class HyGate(SingleQubitGate):
    def __init__(self):
        super().__init__([[_S2, -1j * _S2], [1j * _S2, -_S2]])


class _Diag1(SingleQubitGate):
    traits = T_DIAGONAL


class SGate(_Diag1):
    def __init__(self):
        super().__init__([[1, 0], [0, 1j]])


class SDGGate(_Diag1):
    def __init__(self):
        super().__init__([[1, 0], [0, -1j]])


class TGate(_Diag1):
    def __init__(self):
        super().__init__([[1, 0], [0, cmath.rect(1, math.pi / 4)]])


class TDGGate(_Diag1):
    def __init__(self):
        super().__init__([[1, 0], [0, cmath.rect(1, -math.pi / 4)]])


class PhaseShiftGate(_Diag1):
    def __init__(self, theta=0.0):
        super().__init__([[1, 0], [0, 1]])
        self.SetPhaseShift(theta)

    def SetPhaseShift(self, theta):
        self._m[1, 1] = cmath.rect(1.0, theta)


class PauliXGate(SingleQubitGate):
    traits = T_ANTIDIAGONAL

    def __init__(self):
        super().__init__([[0, 1], [1, 0]])


class PauliYGate(SingleQubitGate):
    traits = T_ANTIDIAGONAL

    def __init__(self):
        super().__init__([[0, -1j], [1j, 0]])


class PauliZGate(_Diag1):
    def __init__(self):
        super().__init__([[1, 0], [0, -1]])


class SquareRootNOTGate(SingleQubitGate):
    def __init__(self):
        a, b = 0.5 + 0.5j, 0.5 - 0.5j
        super().__init__([[a, b], [b, a]])


class SquareRootNOTDagGate(SingleQubitGate):
    def __init__(self):
        a, b = 0.5 - 0.5j, 0.5 + 0.5j
        super().__init__([[a, b], [b, a]])


class SplitterGate(SingleQubitGate):
    def __init__(self):
        super().__init__([[_S2, 1j * _S2], [1j * _S2, _S2]])


class RxGate(SingleQubitGate):
    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        c, s = math.cos(theta * 0.5), math.sin(theta * 0.5)
        self._m = np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


class RyGate(SingleQubitGate):
    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        c, s = math.cos(theta * 0.5), math.sin(theta * 0.5)
        self._m = np.array([[c, -s], [s, c]], dtype=np.complex128)


class RzGate(SingleQubitGate):
    traits = T_DIAGONAL

    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        t2 = theta * 0.5
        self._m = np.array([[cmath.rect(1, -t2), 0], [0, cmath.rect(1, t2)]], dtype=np.complex128)


class UGate(SingleQubitGate):
    def __init__(self, theta=0.0, phi=0.0, lam=0.0, gamma=0.0):
        super().__init__()
        self.SetParams(theta, phi, lam, gamma)

    def SetParams(self, theta=0.0, phi=0.0, lam=0.0, gamma=0.0):
        t2 = theta * 0.5
        c, s = math.cos(t2), math.sin(t2)
        self._m = np.array(
            [
                [cmath.rect(1, gamma) * c, -cmath.rect(1, gamma + lam) * s],
                [cmath.rect(1, gamma + phi) * s, cmath.rect(1, gamma + phi + lam) * c],
            ],
            dtype=np.complex128,
        )


# ---- two-qubit gates (QuantumGate.h:10-372) ---------------------------------
def _swap_like(v):
    m = np.eye(4, dtype=np.complex128)
    m[1, 1] = m[2, 2] = 0
    m[1, 2] = m[2, 1] = v
    return m


class SwapGate(TwoQubitsGate):
    traits = T_SWAP

    def __init__(self):
        super().__init__(_swap_like(1))


class iSwapGate(TwoQubitsGate):
    traits = T_ISWAP

    def __init__(self):
        super().__init__(_swap_like(1j))


class iSwapDagGate(TwoQubitsGate):
    traits = T_ISWAPDAG

    def __init__(self):
        super().__init__(_swap_like(-1j))


class DecrementGate(TwoQubitsGate):
    def __init__(self):
        m = np.zeros((4, 4), dtype=np.complex128)
        m[3, 0] = m[0, 1] = m[1, 2] = m[2, 3] = 1
        super().__init__(m)


class TwoQubitsControlledGate(TwoQubitsGate):
    traits = T_CONTROLLED

    def SetOperation(self, u):
        self._m[2:4, 2:4] = np.asarray(u, dtype=np.complex128)


def _ctrl2(block, traits_extra=0):
    class _G(TwoQubitsControlledGate):
        traits = T_CONTROLLED | traits_extra

        def __init__(self):
            super().__init__()
            self._m[2:4, 2:4] = np.array(block, dtype=np.complex128)

    return _G


class CNOTGate(_ctrl2([[0, 1], [1, 0]], T_ANTIDIAGONAL)):
    pass


class ControlledYGate(_ctrl2([[0, -1j], [1j, 0]], T_ANTIDIAGONAL)):
    pass


class ControlledZGate(_ctrl2([[1, 0], [0, -1]], T_DIAGONAL)):
    pass


class ControlledHadamardGate(_ctrl2([[_S2, _S2], [_S2, -_S2]])):
    pass


class ControlledSquareRootNOTGate(_ctrl2([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])):
    pass


class ControlledSquareRootNOTDagGate(_ctrl2([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])):
    pass


class ControlledPhaseGate(_ctrl2([[1, 0], [0, 1j]], T_DIAGONAL)):
    pass


class ControlledPhaseShiftGate(TwoQubitsControlledGate):
    traits = T_CONTROLLED | T_DIAGONAL

    def __init__(self, theta=0.0):
        super().__init__()
        self.SetPhaseShift(theta)

    def SetPhaseShift(self, theta):
        self._m[3, 3] = cmath.rect(1.0, theta)


class ControlledUGate(TwoQubitsControlledGate):
    def __init__(self, theta=0.0, phi=0.0, lam=0.0, gamma=0.0):
        super().__init__()
        self.SetParams(theta, phi, lam, gamma)

    def SetParams(self, theta=0.0, phi=0.0, lam=0.0, gamma=0.0):
        self._m[2:4, 2:4] = UGate(theta, phi, lam, gamma).getRawOperatorMatrix()


class ControlledRxGate(TwoQubitsControlledGate):
    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        self._m[2:4, 2:4] = RxGate(theta).getRawOperatorMatrix()


class ControlledRyGate(TwoQubitsControlledGate):
    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        self._m[2:4, 2:4] = RyGate(theta).getRawOperatorMatrix()


class ControlledRzGate(TwoQubitsControlledGate):
    traits = T_CONTROLLED | T_DIAGONAL

    def __init__(self, theta=0.0):
        super().__init__()
        self.SetTheta(theta)

    def SetTheta(self, theta):
        self._m[2:4, 2:4] = RzGate(theta).getRawOperatorMatrix()


# ---- three-qubit gates (QuantumGate.h:376-469) ------------------------------
class ThreeQubitsControlledGate(ThreeQubitsGate):
    traits = T_CONTROLLED

    def SetOperation(self, u):
        self._m[4:8, 4:8] = np.asarray(u, dtype=np.complex128)


class ToffoliGate(ThreeQubitsControlledGate):
    traits = T_CONTROLLED | T_CTRL_Q1 | T_ANTIDIAGONAL

    def __init__(self):
        super().__init__()
        self._m[6, 6] = self._m[7, 7] = 0
        self._m[6, 7] = self._m[7, 6] = 1


class FredkinGate(ThreeQubitsControlledGate):
    traits = T_CONTROLLED | T_SWAP

    def __init__(self):
        super().__init__()
        self._m[5, 5] = self._m[6, 6] = 0
        self._m[5, 6] = self._m[6, 5] = 1


class CCZGate(ThreeQubitsControlledGate):
    traits = T_CONTROLLED | T_CTRL_Q1 | T_DIAGONAL

    def __init__(self):
        super().__init__()
        self._m[7, 7] = -1


def haar_unitary(rng: np.random.Generator, dim: int) -> np.ndarray:
    """Haar-random U(dim): QR of a complex Ginibre matrix with the phase fix of
    Mezzadri, "How to generate random matrices from the classical compact
    groups", Notices AMS 54 (2007) 592-604."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def haar_u2(rng: np.random.Generator) -> np.ndarray:
    """Haar-random U(2) (haar_unitary with dim 2, same draws)."""
    return haar_unitary(rng, 2)
