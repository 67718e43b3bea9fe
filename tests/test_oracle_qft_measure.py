"""Pins the oracle's QFT, measurement, sampling and QAOA functions.

QFT: closed form sqrt(N)*ifft / fft/sqrt(N) (numpy) and QFT o IQFT = I.
Measurement: the known answers of MeasurementsTests.cpp:34-167 (collapse
amplitudes), numpy's sequential cumsum for CDF searches, the reference's
defaults (MeasureAll -> 2^n-1, MeasureNoCollapse -> 0) and RepeatedMeasure's
truncation + lower_bound (QubitRegister.h:243-271).
<Z> known answers of Tests.cpp:665-723.
Ising: hand-computed energies of the IsingTests.cpp:14 model and the closed
form of the QAOA cost layer, diag(exp(-i g (sum J z z + sum h z / 2))).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import OracleRegister, ising_state_energy
from qcsim_amd import gates as G
from refmath import cdf_first, qft_closed, random_state

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 10])
def test_qft_closed_form(n):
    rng = np.random.default_rng(n)
    psi = random_state(rng, n)
    r = OracleRegister(n)
    r.set_state(psi)
    r.qft()
    assert np.max(np.abs(r.state() - qft_closed(psi, n))) < 1e-13
    r.set_state(psi)
    r.qft(inverse=True)
    assert np.max(np.abs(r.state() - qft_closed(psi, n, inverse=True))) < 1e-13


def test_qft_subrange_and_roundtrip():
    n = 9
    rng = np.random.default_rng(11)
    psi = random_state(rng, n)
    for sq, eq in [(0, 4), (2, 6), (3, 8), (5, 5), (4, 100)]:
        r = OracleRegister(n)
        r.set_state(psi)
        r.qft(sq, eq)
        e = min(eq, n - 1)
        assert np.max(np.abs(r.state() - qft_closed(psi, n, sq, e))) < 1e-13, (sq, eq)
        r.qft(sq, eq, inverse=True)
        assert np.max(np.abs(r.state() - psi)) < 1e-13


def test_qft_without_swaps_is_bit_reversed():
    n = 6
    rng = np.random.default_rng(2)
    psi = random_state(rng, n)
    r = OracleRegister(n)
    r.set_state(psi)
    r.qft(do_swap=False)
    full = qft_closed(psi, n)
    rev = np.array([int(format(i, f"0{n}b")[::-1], 2) for i in range(1 << n)])
    assert np.max(np.abs(r.state() - full[rev])) < 1e-13


def _measure_state():
    # MeasurementsTests.cpp:34-43: 1/2|00> - i/2|01> + 1/sqrt(2)|11>
    psi = np.array([0.5, -0.5j, 0, 1 / math.sqrt(2)], dtype=complex)
    return psi / np.linalg.norm(psi)


def test_measurement_known_answers():
    with open(os.path.join(GOLDEN, "measurement_known_answers.json")) as f:
        cases = json.load(f)
    for c in cases:
        r = OracleRegister(2)
        r.set_state(_measure_state())
        out = r.measure(c["qubit"], c["qubit"], c["u"])
        assert out == c["outcome"], c
        want = np.array([complex(*z) for z in c["amplitudes"]])
        assert np.max(np.abs(r.state() - want)) < 1e-12, c


def test_cdf_searches_match_sequential_cumsum():
    rng = np.random.default_rng(5)
    for n in (3, 7, 12):
        psi = random_state(rng, n)
        psi[rng.choice(1 << n, size=(1 << n) // 4, replace=False)] = 0  # zero-probability states
        psi /= np.linalg.norm(psi)
        us = np.concatenate([rng.random(200), [0.0, 1 - 2**-53, 0.5]])
        r = OracleRegister(n)
        r.set_state(psi)
        for u in us:
            prob = 1.0 - u
            assert r.measure_no_collapse(u) == cdf_first(psi, prob, 0)
        # MeasureAll default A-1 when prob exceeds the total: scale the state down
        r.set_state(psi * 0.5)
        assert r.measure_all(0.0) == (1 << n) - 1
        r.set_state(psi * 0.5)
        assert r.measure_no_collapse(0.0) == 0


def test_sample_truncation_and_lower_bound():
    rng = np.random.default_rng(9)
    n = 8
    psi = random_state(rng, n)
    r = OracleRegister(n)
    r.set_state(psi)
    u = rng.random(1000)
    got = r.sample(u)
    # restated: CDF truncated after the first value > 1 - eps, then lower_bound
    p = psi.real * psi.real + psi.imag * psi.imag
    c = np.cumsum(p)
    over = np.nonzero(c > 1 - np.finfo(float).eps)[0]
    c = c[: over[0] + 1] if over.size else c
    want = np.searchsorted(c, 1 - u, side="left")
    assert np.array_equal(got, want)
    # range overload maps (s & mask) >> first
    got_r = r.sample(u, 2, 4)
    assert np.array_equal(got_r, (want >> 2) & 7)
    # single-shot shortcut = MeasureNoCollapse
    assert r.sample(u[:1])[0] == cdf_first(psi, 1 - u[0], 0)


def test_expectation_z_known_answers():
    # Tests.cpp:665-723: <Z0> = 1, -1 after X, 0 after H; <Z1> = 0 after H on q1
    def expz(reg, q):
        saved = OracleRegister(reg.n)
        saved.set_state(reg.state())
        reg.apply_gate(G.PauliZGate(), q)
        v = saved.inner(reg)
        reg.set_state(saved.state())
        return v

    r = OracleRegister(3)
    assert abs(expz(r, 0) - 1) < 1e-12
    r.apply_gate(G.PauliXGate(), 0)
    assert abs(expz(r, 0) + 1) < 1e-12
    r.apply_gate(G.HadamardGate(), 0)
    assert abs(expz(r, 0)) < 1e-12
    r = OracleRegister(3)
    r.apply_gate(G.HadamardGate(), 1)
    assert abs(expz(r, 1)) < 1e-12


# IsingTests.cpp:14: J01 = 1.5, J12 = 2, J02 = -3, h = (1.25, 3.25, -2.5)
EDGES = [(0, 1), (1, 2), (0, 2)]
JV = [1.5, 2.0, -3.0]
HV = [1.25, 3.25, -2.5]


def test_ising_energy_hand_values():
    # z = +1 for bit 0 (IsingModel.h:203-206), E = -(sum h z + sum J z z)
    # s=0: -(2 + 0.5) = -2.5; s=7: -(-2 + 0.5) = 1.5; s=1 (z0=-1): -(-0.5 + 3.5) = -3
    assert ising_state_energy(0, 3, EDGES, JV, HV) == pytest.approx(-2.5)
    assert ising_state_energy(7, 3, EDGES, JV, HV) == pytest.approx(1.5)
    assert ising_state_energy(1, 3, EDGES, JV, HV) == pytest.approx(-3.0)
    # exact <H> of a basis state is its energy; of |+++> it is 0
    r = OracleRegister(3)
    r.set_state(np.eye(8)[1])
    assert r.ising_energy(EDGES, JV, HV) == pytest.approx(-3.0)
    r.set_state(np.full(8, 8**-0.5))
    assert abs(r.ising_energy(EDGES, JV, HV)) < 1e-14


def test_qaoa_cost_layer_closed_form():
    # one layer with beta = 0 leaves H^n|0> times the cost diagonal
    n, g = 3, 0.37
    r = OracleRegister(n)
    r.qaoa_exec(EDGES, JV, HV, [g], [0.0])
    s = np.arange(1 << n)
    z = 1 - 2 * ((s[:, None] >> np.arange(n)) & 1)
    phase = sum(J * z[:, i] * z[:, j] for (i, j), J in zip(EDGES, JV)) + 0.5 * (z @ np.array(HV))
    want = np.exp(-1j * g * phase) / math.sqrt(1 << n)
    assert np.max(np.abs(r.state() - want)) < 1e-14
