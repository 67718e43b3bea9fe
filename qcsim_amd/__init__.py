"""qcsim_amd - MI355X (gfx950) statevector engine behind the QC::QubitRegister API.

The compute path is libqcsim_amd.so (hand-written HIP kernels, C-ABI in
include/qcsim_amd.h). This package is its Python host side: a ctypes binding
(`_lib`), the gate library (`gates`) and a mirror of QC::QubitRegister /
QC::SubAlgo::QuantumFourierTransform (`register`). There is no CPU fallback:
calls fail loudly when the HIP library is missing.
"""
from . import gates
from ._lib import LIB_PATH, QcsError, lib
from .register import QuantumFourierTransform, QubitRegister

__all__ = ["gates", "lib", "LIB_PATH", "QcsError", "QubitRegister", "QuantumFourierTransform"]
