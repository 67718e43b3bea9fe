"""bench.py's host-side contract, without a GPU: more than one GPU is refused
with a message and exit code 2 before any device work; the algorithmic-byte
formula prices each gate family by the fraction of amplitudes it changes;
the workload generators are deterministic; the QAOA byte count follows the
pass plan."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from qcsim_amd import gates as G  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_gate_family():
    n = 10
    full = 2 * 16 * (1 << n)
    assert bench.algorithmic_bytes_1q(n) == full
    assert bench.gate_bytes(G.HadamardGate(), n) == full
    assert bench.gate_bytes(G.SGate(), n) == full / 2  # m00 = 1: only the bit-set half changes
    assert bench.gate_bytes(G.RzGate(0.3), n) == full  # both diagonal entries differ from 1
    assert bench.gate_bytes(G.CNOTGate(), n) == full / 2
    assert bench.gate_bytes(G.ControlledZGate(), n) == full / 4
    assert bench.gate_bytes(G.SwapGate(), n) == full / 2
    assert bench.gate_bytes(G.HadamardGate(), n, amp=8) == full / 2


def test_workload_generators_are_deterministic():
    a = bench.random_layered_circuit(8, 3)
    b = bench.random_layered_circuit(8, 3)
    assert len(a) == 3 * (8 + 4)
    for (ga, qa), (gb, qb) in zip(a, b):
        assert qa == qb
        assert (ga.getRawOperatorMatrix() == gb.getRawOperatorMatrix()).all()
    c2 = bench.random_toffoli_circuit(8, 2)
    assert len(c2) == 2 * (8 + 4 + 1)
    assert sum(1 for g, _ in c2 if g.getQubitsNumber() == 3) == 2


@pytest.mark.parametrize("n,p,tb,grp", [(12, 4, 12, 5), (24, 4, 12, 6), (24, 4, 12, 5), (18, 3, 10, 4)])
def test_qaoa_bytes_follow_the_pass_plan(n, p, tb, grp):
    from qcsim_amd.qaoa import qaoa_plan

    plan = qaoa_plan(n, p, tb, grp)
    amp = 8 * (1 << n)
    # the first pass writes without reading, every later one reads and writes
    assert bench.qaoa_eval_bytes(n, p, tb, grp) == (2 * amp * len(plan) - amp, len(plan))
    if (n, tb, grp) == (24, 12, 6):
        assert len(plan) == 9  # 3 passes in layer 0, 2 in each later one
    if (n, tb, grp) == (24, 12, 5):
        assert len(plan) == 13  # groups 5 + 5 + 2: 4 in layer 0, 3 later
    if (n, tb, grp) == (18, 10, 4):
        assert len(plan) == 7  # groups 4 + 4
