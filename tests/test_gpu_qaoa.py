"""Batched QAOA on the GPU against the oracle's gate-by-gate Exec
(IsingModel.h:557-567) and exact energy; complex64 states, FP64 energies."""
import numpy as np
import pytest

from oracle.oracle import OracleRegister
from qcsim_amd.qaoa import IsingInstance, QaoaBatch, plain_optimize, random_ising

pytestmark = pytest.mark.gpu


def oracle_eval(inst, gammas, betas):
    o = OracleRegister(inst.n)
    o.qaoa_exec(inst.edges, inst.J, inst.h, gammas, betas)
    return o.ising_energy(inst.edges, inst.J, inst.h), o.state()


def make_batch_instances(seed, sizes):
    rng = np.random.default_rng(seed)
    return [random_ising(n, rng) for n in sizes]


@pytest.mark.parametrize("tile_bits,group", [(12, 6), (10, 4), (8, 5), (12, 4), (11, 5), (6, 4)])
def test_batch_matches_oracle(tile_bits, group):
    sizes = [4, 6, 8, 10, 12, 14, 16, 18]
    insts = make_batch_instances(tile_bits * 7 + group, sizes)
    p = 3
    rng = np.random.default_rng(5)
    G = rng.uniform(0, np.pi, (len(insts), p))
    B = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p, tile_bits=tile_bits, group=group)
    E = batch.evaluate(G, B)
    for i, inst in enumerate(insts):
        e_ref, psi_ref = oracle_eval(inst, G[i], B[i])
        scale = 1 + np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h))
        assert abs(E[i] - e_ref) < 2e-5 * scale, (inst.n, E[i], e_ref)
        err = np.max(np.abs(batch.state(i) - psi_ref))
        assert err < 3e-6, (inst.n, err)
    batch.close()


@pytest.mark.parametrize("p", [1, 2, 5])
def test_merged_schedule_depths_match_oracle(p):
    """Depths other than 3 and 4 under the merged pass plan (qcs_qaoa_plan):
    tile bits 6 and width-4 groups give 12..16-qubit instances two or three
    mixer groups (10 qubits: one), so the plan alternates merged passes
    between the last two groups and puts the energy on either of them (p = 1:
    no merged pass)."""
    from qcsim_amd.qaoa import qaoa_plan

    sizes = [6, 10, 12, 14, 16]
    insts = make_batch_instances(100 + p, sizes)
    plans = [qaoa_plan(n, p, 6, 4) for n in sizes]
    assert any(any(st[0] == "mixer" and st[2] == 2 for st in pl) for pl in plans) == (p > 1)
    rng = np.random.default_rng(50 + p)
    G = rng.uniform(0, np.pi, (len(insts), p))
    B = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p, tile_bits=6, group=4)
    E = batch.evaluate(G, B)
    for i, inst in enumerate(insts):
        e_ref, psi_ref = oracle_eval(inst, G[i], B[i])
        scale = 1 + np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h))
        assert abs(E[i] - e_ref) < 2e-5 * scale, (inst.n, E[i], e_ref)
        err = np.max(np.abs(batch.state(i) - psi_ref))
        assert err < 3e-6, (inst.n, err)
    batch.close()


def test_graph_replays_match_eager():
    insts = make_batch_instances(9, [6, 12, 14, 16, 18, 20])
    p = 4
    rng = np.random.default_rng(1)
    g_batch = QaoaBatch(insts, p, use_graph=True)
    e_batch = QaoaBatch(insts, p, tile_bits=12, group=6, use_graph=False)
    for _ in range(4):
        G = rng.uniform(0, np.pi, (len(insts), p))
        B = rng.uniform(0, np.pi, (len(insts), p))
        eg = g_batch.evaluate(G, B)
        ee = e_batch.evaluate(G, B)
        assert np.array_equal(eg, ee)
        eg2 = g_batch.evaluate(G, B)
        assert np.array_equal(eg, eg2)


def test_maxcut_known_answer_with_plain_optimiser():
    """IsingTests.cpp:45-93 (MaxCutTest): 6 nodes, 8 edges with J = -1, h = 0,
    p = 2 starting at gamma = (6, 4.2), beta = (0.7, 0.5). The state picked
    as argmin E(s) * P(s) (IsingModel.h:258-271 with exact probabilities)
    must be one of the minimum-energy states."""
    edges = [(0, 1), (0, 2), (0, 5), (1, 2), (1, 3), (2, 3), (2, 4), (3, 5)]
    inst = IsingInstance(6, edges, -np.ones(len(edges)), np.zeros(6))
    factory = lambda insts, p: QaoaBatch(insts, p)  # noqa: E731
    E_min, (g, b), evals, hist = plain_optimize(factory, inst, [6.0, 4.2], [0.7, 0.5])
    assert evals == len(hist) * 9  # 4p + 1 exact evaluations per iteration
    batch = QaoaBatch([inst], 2)
    batch.evaluate(g[None, :], b[None, :])
    probs = np.abs(batch.state(0)) ** 2
    energies = inst.energies()
    pick = int(np.argmin(energies * probs))
    mins = set(np.nonzero(np.abs(energies - energies.min()) <= 1e-9)[0].tolist())
    assert pick in mins, (pick, mins)


@pytest.mark.parametrize("tile_bits,group", [(12, 6), (11, 5)])
def test_config_sizes_match_oracle(tile_bits, group):
    """The batched-QAOA workload's largest instances at its depth: n = 20, 22, 24
    (9 passes per evaluation with width-6 mixer passes at tile 12)
    at p = 4, states and energies against the oracle's gate-by-gate Exec
    (IsingModel.h:485-567) and exact energy (:106-134)."""
    insts = make_batch_instances(77 + tile_bits, [20, 22, 24])
    p = 4
    rng = np.random.default_rng(6)
    G = rng.uniform(0, np.pi, (len(insts), p))
    B = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p, tile_bits=tile_bits, group=group)
    E = batch.evaluate(G, B)
    for i, inst in enumerate(insts):
        e_ref, psi_ref = oracle_eval(inst, G[i], B[i])
        scale = 1 + np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h))
        err = float(np.max(np.abs(batch.state(i) - psi_ref)))
        print(f"n={inst.n}: energy {E[i]:.6f} vs oracle {e_ref:.6f}, state max-abs {err:.2e}")
        assert abs(E[i] - e_ref) < 2e-5 * scale, (inst.n, E[i], e_ref)
        assert err < 3e-6, (inst.n, err)
    batch.close()


def test_26_qubit_instance_matches_oracle():
    """An instance above 24 qubits: three mixer groups at tile 12 / group 6
    (qubits 12..17, 18..23, 24..25) and a stored order that folds
    both of its high 6-bit fields (bits 18.. and 24..); p = 2, with a 16-qubit
    instance in the same batch."""
    insts = make_batch_instances(2026, [16, 26])
    p = 2
    rng = np.random.default_rng(26)
    G = rng.uniform(0, np.pi, (len(insts), p))
    B = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p, tile_bits=12, group=6)
    E = batch.evaluate(G, B)
    for i, inst in enumerate(insts):
        e_ref, psi_ref = oracle_eval(inst, G[i], B[i])
        scale = 1 + np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h))
        err = float(np.max(np.abs(batch.state(i) - psi_ref)))
        print(f"n={inst.n}: energy {E[i]:.6f} vs oracle {e_ref:.6f}, state max-abs {err:.2e}")
        assert abs(E[i] - e_ref) < 2e-5 * scale, (inst.n, E[i], e_ref)
        assert err < 3e-6, (inst.n, err)
    batch.close()


def test_large_uneven_batch_sample_matches_oracle():
    """A batch of 256 instances with n uniform over 12, 14, ..., 24 at p = 4 (the
    batched-QAOA workload's mix at a quarter of its count), evaluated in one
    persistent launch per stage; a fixed sample of 10 instances (the two
    largest, the two smallest and 6 more) is compared with the oracle."""
    from bench import qaoa_batch_instances

    insts, _ = qaoa_batch_instances(256, seed=41)
    assert len({inst.n for inst in insts}) == 7
    p = 4
    rng = np.random.default_rng(7)
    G = rng.uniform(0, np.pi, (len(insts), p))
    B = rng.uniform(0, np.pi, (len(insts), p))
    batch = QaoaBatch(insts, p)
    E = batch.evaluate(G, B)
    assert np.all(np.isfinite(E))
    order = np.argsort([inst.n for inst in insts], kind="stable")
    sample = sorted(set([int(order[-1]), int(order[-2]), int(order[0]), int(order[1])] + list(range(3, 256, 41))))
    for i in sample:
        inst = insts[i]
        e_ref, psi_ref = oracle_eval(inst, G[i], B[i])
        scale = 1 + np.sum(np.abs(inst.J)) + np.sum(np.abs(inst.h))
        err = float(np.max(np.abs(batch.state(i) - psi_ref)))
        print(f"instance {i} (n={inst.n}): energy error {abs(E[i] - e_ref):.2e}, state max-abs {err:.2e}")
        assert abs(E[i] - e_ref) < 2e-5 * scale, (i, inst.n, E[i], e_ref)
        assert err < 3e-6, (i, inst.n, err)
    batch.close()
