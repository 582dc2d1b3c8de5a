"""The public C++ interface on the GPU: tests/cpp/cg_solver_test solves a
BP5 problem by cg::Solve through a solver_t written against the reference's
interface (include/solver.hpp:47-49, include/linearSolver.hpp:49-54) and
compares it with the library's own solve (residual history to 1e-10,
solution to 1e-9), checks linAlg_t, ogs_t and the empty halo_t, and that an
operator's exception reaches the caller."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_cpp_cg_solve_through_solver_t():
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "cg_solver_test")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cg_solver_test ok" in r.stdout
