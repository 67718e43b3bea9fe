"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/qcsim_amd.h declares, and the Python binding covers all of them.
No compute calls (there is no GPU here)."""
import ctypes
import os
import re

from qcsim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "qcsim_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qcs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_version_string():
    lib = _lib.load()
    assert b"gfx950" in lib.qcs_version()


def test_gate_struct_layout():
    # struct qcs_gate: 8 uint32 then 2048 doubles
    assert ctypes.sizeof(_lib.QcsGate) == 8 * 4 + 2048 * 8
