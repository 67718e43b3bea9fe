"""qcs_sample_noalloc (shot buffers in a fixed per-GPU array) against
qcs_sample (per-call buffers), through the C-ABI: identical outcomes on two
calls in a row, the second on a smaller register with fewer shots, in both
precisions, for whole-register and range sampling and the single-shot
shortcut; more shots than the array holds are rejected."""
import ctypes as C

import numpy as np
import pytest

from qcsim_amd import QubitRegister
from qcsim_amd._lib import check, lib
from refmath import random_state

pytestmark = pytest.mark.gpu

_PD = C.POINTER(C.c_double)
_PU64 = C.POINTER(C.c_uint64)


def both(reg, u, all_=1, first=0, second=0):
    u = np.ascontiguousarray(u, dtype=np.float64)
    a = np.zeros(u.shape[0], dtype=np.uint64)
    b = np.zeros(u.shape[0], dtype=np.uint64)
    check(lib().qcs_sample(reg._h, all_, first, second, u.shape[0], u.ctypes.data_as(_PD), a.ctypes.data_as(_PU64)))
    check(lib().qcs_sample_noalloc(reg._h, all_, first, second, u.shape[0], u.ctypes.data_as(_PD),
                                   b.ctypes.data_as(_PU64)))
    return a, b


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_noalloc_sampling_equals_sampling(precision):
    rng = np.random.default_rng(71)
    cases = [(16, 65536), (11, 1000)]  # the second call: smaller register, fewer shots
    for n, shots in cases:
        reg = QubitRegister(n, precision=precision)
        reg.setRegisterStorageFastNoNormalize(random_state(rng, n))
        u = rng.random(shots)
        a, b = both(reg, u)
        assert np.array_equal(a, b), (n, shots)
        a, b = both(reg, u[:257], 0, 2, n - 3)  # range sampling
        assert np.array_equal(a, b), (n, "range")
        a, b = both(reg, u[:1])  # single-shot shortcut
        assert np.array_equal(a, b), (n, "single")
        assert np.all(a < (1 << n))
        reg.close()


def test_noalloc_rejects_too_many_shots():
    reg = QubitRegister(6)
    u = np.zeros(65537)
    out = np.zeros(65537, dtype=np.uint64)
    with pytest.raises(ValueError, match="exceed the maximum"):
        check(lib().qcs_sample_noalloc(reg._h, 1, 0, 0, 65537, u.ctypes.data_as(_PD), out.ctypes.data_as(_PU64)))
    reg.close()
