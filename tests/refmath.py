"""Independent numpy references used to pin the oracle and the engine.

- ``embed``: the dense Kronecker embedding of a k-qubit matrix in an n-qubit
  register, the definition the reference gives in SimpleGates.h:211-232
  (1 qubit), :311-333 (2 qubits), :411-434 (3 qubits): element (i, j) is
  M[loc(i), loc(j)] when i and j agree outside the gate's qubits, else 0,
  with loc() reading bit `qubit` as local bit 0, `controllingQubit1` as
  local bit 1 and `controllingQubit2` as local bit 2.
- QFT closed forms: QFT = sqrt(N) * ifft, IQFT = fft / sqrt(N) along the
  transformed qubit range.
"""
import numpy as np


def embed(m: np.ndarray, n: int, qubits) -> np.ndarray:
    m = np.asarray(m, dtype=np.complex128)
    N = 1 << n
    mask = 0
    for q in qubits:
        mask |= 1 << q
    idx = np.arange(N)
    loc = np.zeros(N, dtype=np.int64)
    for b, q in enumerate(qubits):
        loc |= ((idx >> q) & 1) << b
    same = (idx[:, None] | mask) == (idx[None, :] | mask)
    return np.where(same, m[loc[:, None], loc[None, :]], 0)


def apply_dense(psi, m, n, qubits):
    return embed(m, n, qubits) @ psi


def random_state(rng, n):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def random_unitary(rng, d):
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diagonal(r) / np.abs(np.diagonal(r)))


def qft_closed(psi, n, sq=0, eq=None, inverse=False):
    """QFT on qubits [sq, eq] of an n-qubit state via numpy.fft."""
    eq = n - 1 if eq is None else eq
    m = eq - sq + 1
    t = np.asarray(psi).reshape(1 << (n - eq - 1), 1 << m, 1 << sq)
    if inverse:
        t = np.fft.fft(t, axis=1) / np.sqrt(1 << m)
    else:
        t = np.fft.ifft(t, axis=1) * np.sqrt(1 << m)
    return t.reshape(-1)


def cdf_first(psi, prob, dflt):
    """First index with prob <= sequential cumsum(|psi|^2), else dflt.

    |psi|^2 is re*re + im*im (libstdc++'s std::norm for double) and numpy's
    cumsum accumulates sequentially, so this is the reference's scan
    (QubitRegister.h:175-190) to the last bit."""
    psi = np.asarray(psi)
    c = np.cumsum(psi.real * psi.real + psi.imag * psi.imag)
    i = int(np.searchsorted(c, prob, side="left"))
    return i if i < c.shape[0] else dflt


def dense_pauli(x, z, n):
    """2^n x 2^n matrix of the Pauli string with masks (x, z): qubit q is X
    (x bit only), Z (z bit only), Y = [[0, -i], [i, 0]] (both) or I, built
    by Kronecker products with qubit 0 as the least significant bit."""
    I2 = np.eye(2, dtype=np.complex128)
    X2 = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Y2 = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    Z2 = np.array([[1, 0], [0, -1]], dtype=np.complex128)
    m = np.ones((1, 1), dtype=np.complex128)
    for q in range(n):
        xb, zb = (x >> q) & 1, (z >> q) & 1
        m = np.kron(Y2 if xb and zb else X2 if xb else Z2 if zb else I2, m)
    return m


def pauli_ev(psi, x, z):
    """<psi|P|psi> of the Pauli string with masks (x, z) (as dense_pauli)
    without the matrix: (P psi)_s = i^{|x & z|} (-1)^{popcount((s ^ x) & z)}
    psi_{s ^ x}, the row form of the Kronecker product (Y = i X Z per
    qubit), vectorised over all s."""
    psi = np.asarray(psi, dtype=np.complex128)
    s = np.arange(psi.shape[0], dtype=np.uint64)
    t = s ^ np.uint64(x)
    sign = 1.0 - 2.0 * (np.bitwise_count(t & np.uint64(z)) & 1)
    ny = bin(x & z).count("1")
    pp = (1j ** ny) * sign * psi[t]
    return float(np.vdot(psi, pp).real)


def pauli_evs(psi, xs, zs):
    """pauli_ev over many terms, the products conj(psi_s) psi_{s^x} formed
    once per distinct x mask."""
    psi = np.asarray(psi, dtype=np.complex128)
    s = np.arange(psi.shape[0], dtype=np.uint64)
    out = np.zeros(len(xs))
    for x in sorted(set(xs)):
        t = s ^ np.uint64(x)
        prod = np.conj(psi) * psi[t]
        for k, (xk, z) in enumerate(zip(xs, zs)):
            if xk != x:
                continue
            odd = (np.bitwise_count(t & np.uint64(z)) & 1).astype(bool)
            tot = prod.sum() - 2.0 * prod[odd].sum()
            out[k] = float(((1j ** bin(x & z).count("1")) * tot).real)
    return out
