"""Python mirror of QC::QubitRegister over the HIP engine (QubitRegister.h:11-728).

Method names, defaults and semantics follow the reference. Measurements take
an optional supplied uniform ``u`` in [0, 1); without one the register draws
from its own numpy PCG64 stream (the reference draws from a time-seeded
mt19937_64, QubitRegister.h:26-34, so its draws are not reproducible anyway).
The state lives in HBM; ``getRegisterStorage()`` downloads a copy.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import gates as G
from ._lib import QCS_COMPLEX64, QCS_COMPLEX128, QCS_FUSE_AUTO, QcsGate, check, lib

_PD = C.POINTER(C.c_double)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_PD)


# struct qcs_gate as a numpy record: the eight 32-bit header words (k, traits,
# qubits[5], flags) and the interleaved matrix
_GATE_DT = np.dtype([("head", np.uint32, 8), ("mat", np.float64, 2 * 32 * 32)])
assert _GATE_DT.itemsize == C.sizeof(QcsGate)


def _mat_buf(m: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.complex128)
    return m.view(np.float64).reshape(-1)


class QubitRegister:
    """QC::QubitRegister<> (QubitRegister.h:11)."""

    def __init__(self, N: int = 3, addseed: int = 0, precision: str = "c128"):
        assert N > 0
        prec = {"c128": QCS_COMPLEX128, "c64": QCS_COMPLEX64}[precision]
        h = C.c_void_p()
        check(lib().qcs_create(N, prec, C.byref(h)))
        self._h = h
        self.NrQubits = N
        self.NrBasisStates = 1 << N
        self.precision = precision
        self._rng = np.random.default_rng(addseed if addseed else None)
        self._saved = None
        self._record = False
        self._gates: list[G.AppliedGate] = []

    # ---- lifecycle -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().qcs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
            if self._saved is not None:
                self._saved.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def getNrQubits(self):
        return self.NrQubits

    def getNrBasisStates(self):
        return self.NrBasisStates

    def _u(self, u):
        return float(self._rng.random()) if u is None else float(u)

    # ---- state access (QubitRegister.h:62-167) ---------------------------------
    def getBasisStateAmplitude(self, state: int) -> complex:
        out = np.zeros(2)
        check(lib().qcs_get_amplitude(self._h, state, _dptr(out)))
        return complex(out[0], out[1])

    def getBasisStateProbability(self, state: int) -> float:
        return abs(self.getBasisStateAmplitude(state)) ** 2

    def setToBasisState(self, state: int):
        check(lib().qcs_set_basis_state(self._h, state))

    def setToQubitState(self, q: int):
        if q >= self.NrQubits:
            return
        self.setToBasisState(1 << q)

    def setToCatState(self):
        self.Clear()
        s = 1.0 / math.sqrt(2.0)
        self.setRawAmplitude(0, s)
        self.setRawAmplitude(self.NrBasisStates - 1, s)

    def Reset(self):
        self.setToBasisState(0)

    def setToEqualSuperposition(self):
        check(lib().qcs_fill(self._h, 1.0 / math.sqrt(self.NrBasisStates), 0.0))

    def setRawAmplitude(self, state: int, val: complex):
        val = complex(val)
        check(lib().qcs_set_amplitude(self._h, state, val.real, val.imag))

    def Clear(self):
        check(lib().qcs_fill(self._h, 0.0, 0.0))

    def Normalize(self):
        check(lib().qcs_normalize(self._h))

    def AdjustPhaseAndNormalize(self):
        # QubitRegister.h:133-167: divide by the first of psi[0], psi[A/2],
        # psi[A-1] whose modulus is >= 1e-5, then Normalize.
        for idx in (0, self.NrBasisStates >> 1, self.NrBasisStates - 1):
            v0 = self.getBasisStateAmplitude(idx)
            if abs(v0) >= 1e-5:
                inv = 1.0 / v0
                check(lib().qcs_scale(self._h, inv.real, inv.imag))
                break
        self.Normalize()

    def getRegisterStorage(self) -> np.ndarray:
        out = np.empty(self.NrBasisStates, dtype=np.complex128)
        check(lib().qcs_get_state(self._h, _dptr(out.view(np.float64)), self.NrBasisStates))
        return out

    GetRegisterStorage = getRegisterStorage  # capitalised spelling (the reference's is lower-case)

    def setRegisterStorage(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        if v.shape[0] != self.NrBasisStates:
            return
        check(lib().qcs_set_state(self._h, _dptr(v.view(np.float64)), self.NrBasisStates))
        self.Normalize()

    def setRegisterStorageFastNoNormalize(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        check(lib().qcs_set_state(self._h, _dptr(v.view(np.float64)), self.NrBasisStates))

    def stateFidelity(self, state) -> float:
        """|<reg|state>|^2 (QubitRegister.h:527-534); the host vector streams
        through the device, no second register is made."""
        v = np.ascontiguousarray(state, dtype=np.complex128)
        if v.shape[0] != self.NrBasisStates:
            return 0.0
        return abs(self.inner_host(v)) ** 2

    def inner_host(self, vals) -> complex:
        """<reg|vals> for a host vector, FP64."""
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        out = np.zeros(2)
        check(lib().qcs_inner_host(self._h, _dptr(v.view(np.float64)), v.shape[0], _dptr(out)))
        return complex(out[0], out[1])

    def pauli_expectation(self, xmask, zmask) -> np.ndarray:
        """Exact <psi|P_t|psi> for Pauli strings given by masks (qubit q is X
        for bit q of x only, Z for z only, Y for both): qcs_pauli_expectation."""
        x = np.ascontiguousarray(xmask, dtype=np.uint64)
        z = np.ascontiguousarray(zmask, dtype=np.uint64)
        assert x.shape == z.shape and x.ndim == 1
        out = np.zeros(x.shape[0])
        check(
            lib().qcs_pauli_expectation(
                self._h, x.shape[0], x.ctypes.data_as(_PU64), z.ctypes.data_as(_PU64), _dptr(out)
            )
        )
        return out

    def inner(self, other: "QubitRegister") -> complex:
        out = np.zeros(2)
        check(lib().qcs_inner(self._h, other._h, _dptr(out)))
        return complex(out[0], out[1])

    def compare(self, other: "QubitRegister"):
        mx = np.zeros(1)
        ip = np.zeros(2)
        check(lib().qcs_compare(self._h, other._h, _dptr(mx), _dptr(ip)))
        return float(mx[0]), complex(ip[0], ip[1])

    # ---- gates (QubitRegister.h:434-505) ---------------------------------------
    def ApplyGate(self, gate, qubit=None, controllingQubit1=0, controllingQubit2=0):
        if isinstance(gate, G.AppliedGate) and qubit is None:
            qubit, controllingQubit1, controllingQubit2 = gate.q1, gate.q2, gate.q3
        m = gate.getRawOperatorMatrix()
        k = gate.getQubitsNumber()
        check(
            lib().qcs_apply_gate(
                self._h, k, gate.traits, int(qubit), int(controllingQubit1), int(controllingQubit2), _dptr(_mat_buf(m))
            )
        )
        if self._record:
            self._gates.append(G.AppliedGate(m.copy(), qubit, controllingQubit1, controllingQubit2))

    def ApplyGates(self, gates):
        for g in gates:
            self.ApplyGate(g)

    def apply_dense(self, qubits, m):
        q = (C.c_uint32 * len(qubits))(*qubits)
        check(lib().qcs_apply_dense(self._h, len(qubits), q, _dptr(_mat_buf(m))))

    @staticmethod
    def _gate_array(gates):
        """The gates as a struct qcs_gate array. The array is filled through a
        numpy view of the same memory, one assignment per matrix size, and its
        zero pages are left untouched past each matrix's 2*4^k doubles."""
        n = len(gates)
        a = np.zeros(n, dtype=_GATE_DT)
        heads = []
        mats = {}
        for i, (gate, qubits) in enumerate(gates):
            q = list(qubits) + [0] * (5 - len(qubits))
            heads.append([gate.getQubitsNumber(), gate.traits, *q, getattr(gate, "flags", 0)])
            m = gate.getRawOperatorMatrix()
            rows, ms = mats.setdefault(m.size, ([], []))
            rows.append(i)
            ms.append(m)
        if n:
            a["head"] = np.asarray(heads, dtype=np.uint32)
        for size, (rows, ms) in mats.items():
            block = np.asarray(ms, dtype=np.complex128).reshape(len(rows), size)
            a["mat"][rows, : 2 * size] = block.view(np.float64)
        return (QcsGate * n).from_buffer(a)

    def apply_circuit(self, gates, fuse_width=0):
        """The gates in order; fuse_width 0 (none), 2..5 or "auto" (the width
        with the lowest cost under the measured fusion cost table)."""
        arr = self._gate_array(gates)
        w = QCS_FUSE_AUTO if fuse_width == "auto" else int(fuse_width)
        check(lib().qcs_apply_circuit(self._h, arr, len(gates), w))

    def last_fuse_width(self) -> int:
        return int(lib().qcs_last_fuse_width(self._h))

    def autotune_fusion(self, reps: int = 2) -> np.ndarray:
        """Measure the fusion cost table for this precision and size: ms per
        1-qubit gate (entries 0, 1) and per dense k-qubit block (entry k)."""
        out = np.zeros(6)
        check(lib().qcs_autotune_fusion(self._h, reps, _dptr(out)))
        return out

    def compile_program(self, gates, fuse_width=0, adjoint: bool = False) -> "Program":
        """The gate list captured as a graph for this register (qcs_program_*):
        Compute-style replay, or with adjoint=True the Uncompute order."""
        return Program(self, gates, fuse_width, adjoint)

    def ApplyOperatorMatrix(self, m):
        """psi <- m psi (QubitRegister.h:499-505) on the device: one pass for a
        diagonal m, else a dense FP64 product streamed in row panels."""
        m = np.asarray(m, dtype=np.complex128)
        assert m.shape == (self.NrBasisStates, self.NrBasisStates)
        if np.count_nonzero(m - np.diag(np.diagonal(m))) == 0:
            d = np.ascontiguousarray(np.diagonal(m))
            check(lib().qcs_apply_diagonal(self._h, _dptr(d.view(np.float64)), self.NrBasisStates))
        else:
            mm = np.ascontiguousarray(m)
            check(lib().qcs_apply_matrix(self._h, _dptr(mm.view(np.float64)), self.NrBasisStates))
        if self._record:
            self._gates.append(G.AppliedGate(m.copy()))

    def apply_diagonal(self, diag):
        """psi[s] *= diag[s]: ApplyOperatorMatrix's diagonal case without
        building the 2^n x 2^n matrix."""
        d = np.ascontiguousarray(diag, dtype=np.complex128)
        check(lib().qcs_apply_diagonal(self._h, _dptr(d.view(np.float64)), d.shape[0]))

    def qft(self, sq=0, eq=None, do_swap=True, inverse=False):
        eq = self.NrQubits - 1 if eq is None else eq
        check(lib().qcs_qft(self._h, sq, min(eq, 0xFFFFFFFF), int(do_swap), int(inverse)))

    # ---- measurement (QubitRegister.h:169-429, 592-642) ------------------------
    def MeasureAll(self, u=None) -> int:
        out = C.c_uint64()
        check(lib().qcs_measure_all(self._h, self._u(u), C.byref(out)))
        return out.value

    def MeasureQubit(self, qubit, u=None) -> int:
        return self.Measure(qubit, qubit, u)

    def Measure(self, firstQubit, secondQubit, u=None) -> int:
        out = C.c_uint64()
        check(lib().qcs_measure(self._h, firstQubit, secondQubit, self._u(u), C.byref(out)))
        return out.value

    def MeasureNoCollapse(self, firstQubit=None, secondQubit=None, u=None) -> int:
        out = C.c_uint64()
        full = firstQubit is None
        f = 0 if full else firstQubit
        s = 0 if full else (firstQubit if secondQubit is None else secondQubit)
        check(lib().qcs_measure_no_collapse(self._h, int(full), f, s, self._u(u), C.byref(out)))
        return out.value

    def sample(self, u, firstQubit=None, secondQubit=None) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(u.shape[0], dtype=np.uint64)
        full = firstQubit is None
        f = 0 if full else firstQubit
        s = 0 if full else secondQubit
        check(lib().qcs_sample(self._h, int(full), f, s, u.shape[0], _dptr(u), out.ctypes.data_as(_PU64)))
        return out

    def RepeatedMeasure(self, nrTimes=1000, firstQubit=None, secondQubit=None, u=None) -> dict:
        if nrTimes == 0:
            return {}
        if u is None:
            u = self._rng.random(nrTimes)
        vals, counts = np.unique(self.sample(u, firstQubit, secondQubit), return_counts=True)
        return {int(v): int(c) for v, c in zip(vals, counts)}

    RepeatedMeasureUnordered = RepeatedMeasure

    def GetQubitProbability(self, qubit) -> float:
        out = np.zeros(1)
        check(lib().qcs_qubit_probability(self._h, qubit, _dptr(out)))
        return float(out[0])

    def norm2(self) -> float:
        out = np.zeros(1)
        check(lib().qcs_norm2(self._h, _dptr(out)))
        return float(out[0])

    def ising_energy(self, edges, J, h) -> float:
        ei = np.ascontiguousarray([e[0] for e in edges], dtype=np.uint32)
        ej = np.ascontiguousarray([e[1] for e in edges], dtype=np.uint32)
        Jv = np.ascontiguousarray(J, dtype=np.float64)
        hv = np.ascontiguousarray(h, dtype=np.float64)
        out = np.zeros(1)
        check(
            lib().qcs_ising_energy(
                self._h, len(ei), ei.ctypes.data_as(_PU32), ej.ctypes.data_as(_PU32), _dptr(Jv), _dptr(hv), _dptr(out)
            )
        )
        return float(out[0])

    # ---- save / restore / record (QubitRegister.h:536-616, 646-674) ------------
    def SaveState(self):
        if self._saved is None:
            self._saved = QubitRegister(self.NrQubits, precision=self.precision)
        check(lib().qcs_copy_state(self._saved._h, self._h))

    def RestoreState(self):
        if self._saved is None:
            return
        check(lib().qcs_copy_state(self._h, self._saved._h))

    def RestoreStateDestructive(self):
        if self._saved is None:
            return
        check(lib().qcs_swap_state(self._h, self._saved._h))
        self._saved.close()
        self._saved = None

    def ComputeStart(self):
        self._record = True
        self._gates = []

    def ComputeEnd(self):
        self._record = False

    def ComputeClear(self):
        self._gates = []

    def Compute(self):
        rec, self._record = self._record, False
        for g in self._gates:
            if g.getQubitsNumber() > 3:
                self.ApplyOperatorMatrix(g.getRawOperatorMatrix())
            else:
                self.ApplyGate(g)
        self._record = rec

    def Uncompute(self):
        rec, self._record = self._record, False
        for g in reversed(self._gates):
            m = g.getRawOperatorMatrix().conj().T
            if g.getQubitsNumber() > 3:
                self.ApplyOperatorMatrix(m)
            else:
                self.ApplyGate(G.AppliedGate(m, g.q1, g.q2, g.q3))
        self._record = rec

    def ExpectationValue(self, gates) -> complex:
        """<psi|G|psi> for the product G of the recorded gates
        (QubitRegister.h:646-660). Single-qubit gates that are multiples of
        Pauli matrices (the reference's TODO at :649) take one read-only
        FP64 pass; other lists run on a copy of the state."""
        if not gates:
            return 1.0
        from .pauli import gates_as_pauli

        pc = gates_as_pauli(gates, self.NrQubits)
        if pc is not None:
            coef, x, z = pc
            return complex(coef * self.pauli_expectation([x], [z])[0])
        saved = QubitRegister(self.NrQubits, precision=self.precision)
        check(lib().qcs_copy_state(saved._h, self._h))
        self.ApplyGates(gates)
        res = saved.inner(self)
        check(lib().qcs_swap_state(self._h, saved._h))
        saved.close()
        return res

    def Clone(self) -> "QubitRegister":
        qr = QubitRegister(self.NrQubits, precision=self.precision)
        check(lib().qcs_copy_state(qr._h, self._h))
        if self._saved is not None:
            qr.SaveState()
            check(lib().qcs_copy_state(qr._saved._h, self._saved._h))
        qr._gates = list(self._gates)
        qr._record = self._record
        return qr

    def autotune(self, reps: int = 3) -> np.ndarray:
        """Pick the fastest launch variant per qubit (state values unchanged);
        returns GB/s per (qubit, variant)."""
        nv = lib().qcs_autotune_variants()
        out = np.zeros(self.NrQubits * nv)
        check(lib().qcs_autotune(self._h, reps, _dptr(out)))
        return out.reshape(self.NrQubits, nv)

    # ---- timing ------------------------------------------------------------------
    def synchronize(self):
        check(lib().qcs_synchronize(self._h))

    def timer_begin(self):
        check(lib().qcs_timer_begin(self._h))

    def timer_end(self) -> float:
        ms = C.c_float()
        check(lib().qcs_timer_end(self._h, C.byref(ms)))
        return float(ms.value)

    def launch_count(self) -> int:
        return int(lib().qcs_launch_count(self._h))


class QuantumFourierTransform:
    """QC::SubAlgo::QuantumFourierTransform (QuantumFourierTransform.h:12-92)."""

    def __init__(self, N: int, startQubit: int = 0, endQubit: int = 2**31 - 1):
        self.sQubit = startQubit
        # QuantumSubAlgorithmOnSubregister clamp (QuantumAlgorithm.h:228-229)
        self.eQubit = max(startQubit, min(N - 1, endQubit))
        self.hadamard = G.HadamardGate()
        self.cPhaseShift = G.ControlledPhaseShiftGate()

    def getStartQubit(self):
        return self.sQubit

    def getEndQubit(self):
        return self.eQubit

    def Execute(self, reg: QubitRegister, u=None) -> int:
        self.QFT(reg)
        return reg.MeasureAll(u)

    def QFT(self, reg: QubitRegister, doSwap: bool = True):
        if reg._record:
            return self._sequence(reg, doSwap, False)
        reg.qft(self.sQubit, self.eQubit, doSwap, False)

    def IQFT(self, reg: QubitRegister, doSwap: bool = True):
        if reg._record:
            return self._sequence(reg, doSwap, True)
        reg.qft(self.sQubit, self.eQubit, doSwap, True)

    def Swap(self, reg: QubitRegister):
        sw = G.SwapGate()
        a, b = self.sQubit, self.eQubit
        while a < b:
            reg.ApplyGate(sw, a, b)
            a += 1
            b -= 1

    def _sequence(self, reg, doSwap, inverse):
        # gate-by-gate form, used while the register records gates
        sq, eq = self.sQubit, self.eQubit
        if not inverse:
            reg.ApplyGate(self.hadamard, eq)
            for cur in range(eq, sq, -1):
                phase = math.pi / 2
                for ctrl in range(cur - 1, sq - 1, -1):
                    self.cPhaseShift.SetPhaseShift(phase)
                    reg.ApplyGate(self.cPhaseShift, cur, ctrl)
                    phase *= 0.5
                reg.ApplyGate(self.hadamard, cur - 1)
            if doSwap:
                self.Swap(reg)
        else:
            if doSwap:
                self.Swap(reg)
            for cur in range(sq + 1, eq + 1):
                reg.ApplyGate(self.hadamard, cur - 1)
                phase = -math.pi / 2
                for ctrl in range(cur - 1, sq - 1, -1):
                    self.cPhaseShift.SetPhaseShift(phase)
                    reg.ApplyGate(self.cPhaseShift, cur, ctrl)
                    phase *= 0.5
            reg.ApplyGate(self.hadamard, eq)


class Program:
    """A gate sequence compiled for one register and replayed as one graph
    launch (the recorded-gate replay of QubitRegister.h:553-590)."""

    def __init__(self, reg: QubitRegister, gates, fuse_width: int = 0, adjoint: bool = False):
        arr = QubitRegister._gate_array(gates)
        h = C.c_void_p()
        w = QCS_FUSE_AUTO if fuse_width == "auto" else int(fuse_width)
        check(lib().qcs_program_create(reg._h, arr, len(gates), w, int(adjoint), C.byref(h)))
        self._h = h
        self._reg = reg

    def run(self):
        check(lib().qcs_program_run(self._reg._h, self._h))

    def launches(self) -> int:
        n = C.c_uint32()
        check(lib().qcs_program_info(self._h, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().qcs_program_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
