"""Registers past 2^32 amplitudes: 33 qubits complex64 (64 GiB) and 32
qubits complex128 (64 GiB). Index arithmetic above 32 bits in the gate,
dense (MFMA), QFT, reduction and sampling kernels, checked through closed
forms that hold at any size (product states, QFT of a basis state)."""
import math

import numpy as np
import pytest

from qcsim_amd import QubitRegister
from qcsim_amd import gates as G
from refmath import random_unitary

pytestmark = pytest.mark.gpu


def product_state(rng, n):
    """Haar U(2) on every qubit of |0...0>: psi[s] = prod_q u_q[bit_q(s), 0]."""
    us = [G.haar_u2(rng) for _ in range(n)]
    cols = [u[:, 0] for u in us]

    def amp(s):
        v = 1.0 + 0j
        for q in range(n):
            v *= cols[q][(s >> q) & 1]
        return v

    return us, cols, amp


def test_33_qubits_complex64_gates_dense_block_and_sampling():
    n = 33
    N = 1 << n
    rng = np.random.default_rng(33)
    r = QubitRegister(n, precision="c64")
    us, cols, amp = product_state(rng, n)
    for q, u in enumerate(us):
        r.ApplyGate(G.SingleQubitGate(u), q)
    probes = [int(x) for x in rng.integers(N, size=48)] + [N - 1, 1 << 32, (1 << 32) + 5, (1 << 32) - 1]
    for s in probes:
        want = amp(s)
        assert abs(r.getBasisStateAmplitude(s) - want) < 2e-6 * abs(want) + 1e-12, s
    for q in (0, 31, 32):
        assert abs(r.GetQubitProbability(q) - abs(cols[q][1]) ** 2) < 2e-6, q
    assert abs(r.norm2() - 1) < 1e-5
    # a dense 5-qubit block across bit 32 (MFMA, FP32): new amplitudes are
    # sum_c M[local(s), c] psi[s with local bits c]
    qs = [3, 17, 30, 31, 32]
    m = random_unitary(rng, 32)
    r.apply_dense(qs, m)

    def dense_amp(s):
        loc = sum(((s >> q) & 1) << i for i, q in enumerate(qs))
        rest = s & ~sum(1 << q for q in qs)
        tot = 0j
        for c in range(32):
            sc = rest | sum(((c >> i) & 1) << q for i, q in enumerate(qs))
            tot += m[loc, c] * amp(sc)
        return tot

    for s in probes[:24]:
        want = dense_amp(s)
        assert abs(r.getBasisStateAmplitude(s) - want) < 5e-6 * (abs(want) + 2 ** (-n / 2)), s
    # sampling: the top qubit's marginal is unchanged by blocks on other qubits
    # only if the block avoids it; here it does not, so compare with the
    # engine's own FP64 probability instead
    p32 = r.GetQubitProbability(32)
    u = rng.random(4096)
    out = r.sample(u)
    assert out.max() < N
    frac = float(np.mean((out >> 32) & 1))
    assert abs(frac - p32) < 4 * math.sqrt(p32 * (1 - p32) / len(u)) + 1e-3, (frac, p32)
    r.close()


def test_32_qubits_complex128_qft_closed_form():
    n = 32
    N = 1 << n
    rng = np.random.default_rng(32)
    r = QubitRegister(n)
    j = (1 << 31) + int(rng.integers(1 << 31))
    r.setToBasisState(j)
    r.qft()
    for k in [int(x) for x in rng.integers(N, size=32)] + [N - 1, 1 << 31]:
        want = np.exp(2j * np.pi * ((j * k) % N) / N) / math.sqrt(N)
        assert abs(r.getBasisStateAmplitude(k) - want) < 1e-12 / math.sqrt(N) * 64, k
    assert abs(r.norm2() - 1) < 1e-12
    r.close()


def test_33_qubits_complex64_qft_closed_form():
    """The QFT passes past 2^32 amplitudes in complex64; the bit reversal's
    tile pairs (4,196,352 at 33 qubits) take two launches."""
    n = 33
    N = 1 << n
    rng = np.random.default_rng(331)
    r = QubitRegister(n, precision="c64")
    j = (1 << 32) + int(rng.integers(1 << 32))
    r.setToBasisState(j)
    r.qft()
    scale = 1 / math.sqrt(N)
    for k in [int(x) for x in rng.integers(N, size=32)] + [0, N - 1, 1 << 32, (1 << 32) - 1]:
        want = np.exp(2j * np.pi * ((j * k) % N) / N) * scale
        assert abs(r.getBasisStateAmplitude(k) - want) < 2e-5 * scale, k
    assert abs(r.norm2() - 1) < 1e-5
    r.close()
