"""ctypes binding of libqcsim_amd.so (the C-ABI in include/qcsim_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C qcsim_amd/csrc``). There is no fallback: if the HIP library is
missing or cannot be loaded, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqcsim_amd.so")

QCS_OK, QCS_EINVAL, QCS_ENOMEM, QCS_EDEVICE = 0, 1, 2, 3
QCS_GATE_DENSE = 0x1
QCS_COMPLEX128, QCS_COMPLEX64 = 0, 1
QCS_FUSE_AUTO = 0xFFFFFFFF

# Every symbol include/qcsim_amd.h declares, with its ctypes signature.
_P = C.c_void_p
_U32, _U64, _I, _D = C.c_uint32, C.c_uint64, C.c_int, C.c_double
_PD = C.POINTER(C.c_double)
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)


class QcsGate(C.Structure):
    """struct qcs_gate (include/qcsim_amd.h)."""

    _fields_ = [
        ("k", C.c_uint32),
        ("traits", C.c_uint32),
        ("qubits", C.c_uint32 * 5),
        ("flags", C.c_uint32),
        ("mat", C.c_double * (2 * 32 * 32)),
    ]


SIGNATURES = {
    "qcs_create": (_I, [_U32, _I, C.POINTER(_P)]),
    "qcs_destroy": (_I, [_P]),
    "qcs_get_nqubits": (_U32, [_P]),
    "qcs_get_precision": (_I, [_P]),
    "qcs_copy_state": (_I, [_P, _P]),
    "qcs_swap_state": (_I, [_P, _P]),
    "qcs_set_state": (_I, [_P, _PD, _U64]),
    "qcs_get_state": (_I, [_P, _PD, _U64]),
    "qcs_get_amplitude": (_I, [_P, _U64, _PD]),
    "qcs_set_amplitude": (_I, [_P, _U64, _D, _D]),
    "qcs_set_basis_state": (_I, [_P, _U64]),
    "qcs_fill": (_I, [_P, _D, _D]),
    "qcs_scale": (_I, [_P, _D, _D]),
    "qcs_apply_gate": (_I, [_P, _U32, _U32, _U32, _U32, _U32, _PD]),
    "qcs_apply_dense": (_I, [_P, _U32, _PU32, _PD]),
    "qcs_apply_circuit": (_I, [_P, C.POINTER(QcsGate), _U32, _U32]),
    "qcs_apply_diagonal": (_I, [_P, _PD, _U64]),
    "qcs_apply_matrix": (_I, [_P, _PD, _U64]),
    "qcs_fuse": (_I, [C.POINTER(QcsGate), _U32, _U32, C.POINTER(QcsGate), C.POINTER(_U32)]),
    "qcs_program_create": (_I, [_P, C.POINTER(QcsGate), _U32, _U32, _I, C.POINTER(_P)]),
    "qcs_program_run": (_I, [_P, _P]),
    "qcs_program_info": (_I, [_P, _PU32]),
    "qcs_program_destroy": (_I, [_P]),
    "qcs_qft": (_I, [_P, _U32, _U32, _I, _I]),
    "qcs_measure_all": (_I, [_P, _D, _PU64]),
    "qcs_measure": (_I, [_P, _U32, _U32, _D, _PU64]),
    "qcs_measure_no_collapse": (_I, [_P, _I, _U32, _U32, _D, _PU64]),
    "qcs_sample": (_I, [_P, _I, _U32, _U32, _U64, _PD, _PU64]),
    "qcs_sample_noalloc": (_I, [_P, _I, _U32, _U32, _U64, _PD, _PU64]),
    "qcs_qubit_probability": (_I, [_P, _U32, _PD]),
    "qcs_norm2": (_I, [_P, _PD]),
    "qcs_normalize": (_I, [_P]),
    "qcs_inner": (_I, [_P, _P, _PD]),
    "qcs_compare": (_I, [_P, _P, _PD, _PD]),
    "qcs_inner_host": (_I, [_P, _PD, _U64, _PD]),
    "qcs_pauli_expectation": (_I, [_P, _U32, _PU64, _PU64, _PD]),
    "qcs_ising_energy": (_I, [_P, _U32, _PU32, _PU32, _PD, _PD, _PD]),
    "qcs_autotune": (_I, [_P, _U32, _PD]),
    "qcs_autotune_variants": (_I, []),
    "qcs_autotune_fusion": (_I, [_P, _U32, _PD]),
    "qcs_set_fusion_costs": (_I, [_I, _U32, _PD]),
    "qcs_get_fusion_costs": (_I, [_I, _U32, _PD]),
    "qcs_last_fuse_width": (_U32, [_P]),
    "qcs_set_tuned_variant": (_I, [_I, _U32, _U32, _I]),
    "qcs_get_tuned_variant": (_I, [_I, _U32, _U32]),
    "qcs_qaoa_create": (_I, [_U32, _PU32, _PU32, _PU32, _PU32, _PD, _PD, _U32, C.POINTER(_P)]),
    "qcs_qaoa_destroy": (_I, [_P]),
    "qcs_qaoa_configure": (_I, [_P, _U32, _U32, _I]),
    "qcs_qaoa_evaluate": (_I, [_P, _PD, _PD, _PD]),
    "qcs_qaoa_last_ms": (_I, [_P, C.POINTER(C.c_float)]),
    "qcs_qaoa_get_state": (_I, [_P, _U32, _PD]),
    "qcs_qaoa_info": (_I, [_P, _PU32, _PU64, _PU32, _PU32]),
    "qcs_qaoa_plan": (_I, [_U32, _U32, _U32, _U32, _PU32, _U32, _PU32]),
    "qcs_vqe_create": (_I, [_U32, _PU32, _PU32, _PU64, _PU64, _PD, C.POINTER(_P)]),
    "qcs_vqe_destroy": (_I, [_P]),
    "qcs_vqe_evaluate": (_I, [_P, _PU32, _PD, _PD]),
    "qcs_vqe_last_ms": (_I, [_P, C.POINTER(C.c_float)]),
    "qcs_vqe_get_state": (_I, [_P, _U32, _PD]),
    "qcs_mlp_create": (_I, [_U32, _U32, _U32, _U64, C.POINTER(_P)]),
    "qcs_mlp_destroy": (_I, [_P]),
    "qcs_mlp_num_params": (_U64, [_P]),
    "qcs_mlp_get_params": (_I, [_P, C.POINTER(C.c_float)]),
    "qcs_mlp_set_params": (_I, [_P, C.POINTER(C.c_float)]),
    "qcs_mlp_forward": (_I, [_P, C.POINTER(C.c_float), _U32, C.POINTER(C.c_float)]),
    "qcs_mlp_gradients": (_I, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), _U32, C.POINTER(C.c_float),
                               C.POINTER(C.c_float)]),
    "qcs_mlp_train": (_I, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), _U64, _U32, _U32, C.c_float, _U64,
                           C.POINTER(C.c_float)]),
    "qcs_synchronize": (_I, [_P]),
    "qcs_timer_begin": (_I, [_P]),
    "qcs_timer_end": (_I, [_P, C.POINTER(C.c_float)]),
    "qcs_launch_count": (_U64, [_P]),
    "qcs_last_error": (C.c_char_p, []),
    "qcs_version": (C.c_char_p, []),
}


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: the HIP engine is not built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'`)"
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class QcsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(code: int) -> None:
    if code != QCS_OK:
        msg = lib().qcs_last_error().decode()
        if code == QCS_EINVAL:
            raise ValueError(msg)
        if code == QCS_ENOMEM:
            raise MemoryError(msg)
        raise QcsError(code, msg)
