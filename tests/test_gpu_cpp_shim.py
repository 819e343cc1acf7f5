"""The C++ Solver/updater shims (include/aster_hip_solver.hpp) driven from a
C++ program, checked bit for bit against the oracle inside that program."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_solver")


@pytest.mark.gpu
def test_cpp_solver_shims():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "cpp shim: ok" in r.stdout


def test_cpp_program_is_built():
    assert os.path.exists(BIN)
