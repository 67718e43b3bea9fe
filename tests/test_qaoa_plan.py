"""The batched QAOA pass plan (qcs_qaoa_plan, IsingModel.h:557-567's layer
sequence regrouped into passes over the state), replayed in numpy on the CPU:
every plan the library builds must give the same state as the plain circuit
H^n, then per layer exp(-i gamma_l C) and Rx(2 beta_l) on every qubit, for an
arbitrary diagonal C. A plan that applied a layer's cost after one of its
rotations, skipped or repeated a rotation, or used the wrong layer's angle
fails here."""
import numpy as np
import pytest

from qcsim_amd.qaoa import qaoa_plan


def _rx(psi, n, qubits, beta):
    c, s = np.cos(beta), np.sin(beta)
    t = psi.reshape([2] * n)
    for q in qubits:
        ax = n - 1 - q
        a, b = np.take(t, 0, axis=ax), np.take(t, 1, axis=ax)
        t = np.stack([c * a - 1j * s * b, -1j * s * a + c * b], axis=ax)
    return t.reshape(-1)


def _plain(n, C, gam, bet):
    psi = np.full(1 << n, 2 ** (-n / 2), dtype=complex)
    for g, b in zip(gam, bet):
        psi = _rx(psi * np.exp(-1j * g * C), n, range(n), b)
    return psi


def _replay(plan, n, tb, C, gam, bet):
    psi = np.full(1 << n, 2 ** (-n / 2), dtype=complex)
    energy_passes = 0
    for k, (kind, w, mode, qlo, l) in enumerate(plan):
        if kind == "lds":
            return _plain(n, C, gam, bet), 1
        # only the first pass starts from |+...+> (writes without reading)
        assert (k == 0) == (kind == "tile" and mode == 0 and l == 0)
        if kind == "tile":
            if mode == 0:
                psi = psi * np.exp(-1j * gam[l] * C)
            psi = _rx(psi, n, range(tb), bet[l])
        else:
            assert kind == "mixer" and qlo >= tb and qlo + w <= n
            psi = _rx(psi, n, range(qlo, qlo + w), bet[l])
            if mode == 2:
                psi = _rx(psi * np.exp(-1j * gam[l + 1] * C), n, range(qlo, qlo + w), bet[l + 1])
            energy_passes += mode == 1
    return psi, energy_passes


@pytest.mark.parametrize("n,tb,grp", [(10, 12, 6), (14, 8, 6), (14, 6, 4), (15, 6, 5), (16, 6, 4), (13, 7, 4)])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_plan_equals_plain_qaoa(n, tb, grp, p):
    rng = np.random.default_rng(1000 * n + 10 * p + grp)
    C = rng.normal(size=1 << n) * 3
    gam, bet = rng.uniform(0, np.pi, p), rng.uniform(0, np.pi, p)
    plan = qaoa_plan(n, p, tb, grp)
    psi, ne = _replay(plan, n, tb, C, gam, bet)
    np.testing.assert_allclose(psi, _plain(n, C, gam, bet), atol=1e-12)
    assert ne == 1  # the energy is read exactly once, by the last pass
    assert plan[-1][0] == "lds" or plan[-1][2] == 1
    if n > tb:
        ng = -(-(n - tb) // grp)
        assert len(plan) == (p * ng + 1 if ng >= 2 else 2 * p)
        # each layer rotates every qubit exactly once
        for l in range(p):
            seen = []
            for kind, w, mode, qlo, layer in plan:
                if kind == "tile" and layer == l:
                    seen += range(tb)
                if kind == "mixer" and (layer == l or (mode == 2 and layer + 1 == l)):
                    seen += range(qlo, qlo + w)
            assert sorted(seen) == list(range(n))


def test_plan_rejects_bad_configuration():
    with pytest.raises(Exception):
        qaoa_plan(20, 4, 13, 6)
    with pytest.raises(Exception):
        qaoa_plan(20, 4, 12, 3)
