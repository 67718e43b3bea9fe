"""ctypes wrapper of the CPU oracle, oracle/libqcs_oracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libqcs_oracle.so")
# the same source built with -msse2 -mpopcnt (the reference's CMake flags on
# x86-64 Linux) instead of -mavx2; timed beside the AVX2 build
LIB_SSE2 = os.path.join(HERE, "libqcs_oracle_sse2.so")

_PD = C.POINTER(C.c_double)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(isa: str = "avx2"):
    """The oracle library: isa "avx2" (the default, used by the tests) or
    "sse2"."""
    if isa not in _libs:
        path = {"avx2": LIB, "sse2": LIB_SSE2}[isa]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "orc_create": (C.c_void_p, [C.c_uint32]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_set_multithreading": (None, [C.c_void_p, C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_set_state": (None, [C.c_void_p, _PD]),
            "orc_get_state": (None, [C.c_void_p, _PD]),
            "orc_apply_gate": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _PD]),
            "orc_apply_dense": (C.c_int, [C.c_void_p, C.c_uint32, _PU32, _PD]),
            "orc_measure_all": (C.c_uint64, [C.c_void_p, C.c_double]),
            "orc_measure": (C.c_uint64, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double]),
            "orc_measure_no_collapse": (C.c_uint64, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_double]),
            "orc_sample": (None, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, _PD, _PU64]),
            "orc_qubit_probability": (C.c_double, [C.c_void_p, C.c_uint32]),
            "orc_norm2": (C.c_double, [C.c_void_p]),
            "orc_normalize": (None, [C.c_void_p]),
            "orc_inner": (None, [C.c_void_p, C.c_void_p, _PD]),
            "orc_qft": (None, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int]),
            "orc_ising_energy": (C.c_double, [C.c_void_p, C.c_uint32, _PU32, _PU32, _PD, _PD]),
            "orc_ising_state_energy": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint32, _PU32, _PU32, _PD, _PD]),
            "orc_qaoa_exec": (None, [C.c_void_p, C.c_uint32, _PU32, _PU32, _PD, _PD, C.c_uint32, _PD, _PD, C.c_int]),
            "orc_last_error": (C.c_char_p, []),
            "orc_stream_triad": (C.c_double, [C.c_uint64, C.c_int, C.c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _libs[isa] = L
    return _libs[isa]


def _d(a):
    return a.ctypes.data_as(_PD)


def _mat(m):
    return np.ascontiguousarray(m, dtype=np.complex128).view(np.float64).reshape(-1)


class OracleRegister:
    """CPU restatement of QC::QubitRegister (complex128, results buffer + swap)."""

    def __init__(self, n: int, multithreading: bool = True, isa: str = "avx2"):
        self.n = n
        self.A = 1 << n
        self._L = lib(isa)
        self._h = self._L.orc_create(n)
        if not self._h:
            raise MemoryError("oracle register allocation failed")
        self._L.orc_set_multithreading(self._h, int(multithreading))

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, psi):
        v = np.ascontiguousarray(psi, dtype=np.complex128)
        assert v.shape[0] == self.A
        self._L.orc_set_state(self._h, _d(v.view(np.float64)))

    def state(self) -> np.ndarray:
        out = np.empty(self.A, dtype=np.complex128)
        self._L.orc_get_state(self._h, _d(out.view(np.float64)))
        return out

    def apply_gate(self, gate, q, c1=0, c2=0):
        rc = self._L.orc_apply_gate(self._h, gate.getQubitsNumber(), gate.traits, q, c1, c2,
                                  _d(_mat(gate.getRawOperatorMatrix())))
        if rc:
            raise ValueError(self._L.orc_last_error().decode())

    def apply_raw(self, k, traits, q, c1, c2, m):
        rc = self._L.orc_apply_gate(self._h, k, traits, q, c1, c2, _d(_mat(m)))
        if rc:
            raise ValueError(self._L.orc_last_error().decode())

    def apply_dense(self, qubits, m):
        q = (C.c_uint32 * len(qubits))(*qubits)
        if self._L.orc_apply_dense(self._h, len(qubits), q, _d(_mat(m))):
            raise ValueError("bad dense block")

    def measure_all(self, u):
        return int(self._L.orc_measure_all(self._h, u))

    def measure(self, first, second, u):
        return int(self._L.orc_measure(self._h, first, second, u))

    def measure_no_collapse(self, u, first=None, second=None):
        full = first is None
        return int(self._L.orc_measure_no_collapse(self._h, int(full), first or 0,
                                                 (second if second is not None else (first or 0)), u))

    def sample(self, u, first=None, second=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(u.shape[0], dtype=np.uint64)
        full = first is None
        self._L.orc_sample(self._h, int(full), first or 0, second or 0, u.shape[0], _d(u),
                         out.ctypes.data_as(_PU64))
        return out

    def qubit_probability(self, q):
        return float(self._L.orc_qubit_probability(self._h, q))

    def norm2(self):
        return float(self._L.orc_norm2(self._h))

    def normalize(self):
        self._L.orc_normalize(self._h)

    def inner(self, other):
        out = np.zeros(2)
        self._L.orc_inner(self._h, other._h, _d(out))
        return complex(out[0], out[1])

    def qft(self, sq=0, eq=None, do_swap=True, inverse=False):
        eq = self.n - 1 if eq is None else eq
        self._L.orc_qft(self._h, sq, eq, int(do_swap), int(inverse))

    def ising_energy(self, edges, J, h):
        ei = np.ascontiguousarray([e[0] for e in edges], dtype=np.uint32)
        ej = np.ascontiguousarray([e[1] for e in edges], dtype=np.uint32)
        Jv = np.ascontiguousarray(J, dtype=np.float64)
        hv = np.ascontiguousarray(h, dtype=np.float64)
        return float(self._L.orc_ising_energy(self._h, len(ei), ei.ctypes.data_as(_PU32), ej.ctypes.data_as(_PU32),
                                            _d(Jv), _d(hv)))

    def qaoa_exec(self, edges, J, h, gamma, beta, better_mixing=False):
        ei = np.ascontiguousarray([e[0] for e in edges], dtype=np.uint32)
        ej = np.ascontiguousarray([e[1] for e in edges], dtype=np.uint32)
        Jv = np.ascontiguousarray(J, dtype=np.float64)
        hv = np.ascontiguousarray(h, dtype=np.float64)
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        b = np.ascontiguousarray(beta, dtype=np.float64)
        self._L.orc_qaoa_exec(self._h, len(ei), ei.ctypes.data_as(_PU32), ej.ctypes.data_as(_PU32), _d(Jv), _d(hv),
                            len(g), _d(g), _d(b), int(better_mixing))


def ising_state_energy(s, n, edges, J, h):
    ei = np.ascontiguousarray([e[0] for e in edges], dtype=np.uint32)
    ej = np.ascontiguousarray([e[1] for e in edges], dtype=np.uint32)
    Jv = np.ascontiguousarray(J, dtype=np.float64)
    hv = np.ascontiguousarray(h, dtype=np.float64)
    return float(lib().orc_ising_state_energy(s, n, len(ei), ei.ctypes.data_as(_PU32), ej.ctypes.data_as(_PU32),
                                              _d(Jv), _d(hv)))
