"""The C++ facade (include/QCSim/): reference-shaped programs compile against
it on the CPU, and run correctly on the GPU (tests/cpp/facade_tests.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "facade_tests")


def test_facade_compiles_without_eigen():
    # -B: rebuild even when present, with the repo-owned Eigen subset
    subprocess.run(["make", "-s", "-B", "-C", CPP, "CXXFLAGS=-O1 -std=c++17 -Wall -Werror -DQCS_NO_EIGEN"], check=True)
    assert os.access(BIN, os.X_OK)


@pytest.mark.gpu
def test_facade_programs_on_gpu():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", CPP], check=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL PASSED" in p.stdout
