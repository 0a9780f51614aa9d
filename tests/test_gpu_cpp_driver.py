"""The C++ mirror of the reference interface (feng_amd/cpp: createLinearSystem
with the HIP enumerator, feLinearSystemHIP, solveNewtonRaphson) driven by the
port of exe/example1.cpp's usage, on the reference's own mesh files.  Its
solution must equal the Python driver's over the same C ABI, and its nodal
error against the example's series solution must fall with the degree."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "feng_amd" / "lib" / "example1"
MESHES = ROOT / "tests" / "golden" / "meshes"

pytestmark = pytest.mark.gpu


def _degree(order):
    """A rule exact for the stiffness of the element (degree 2 (order - 1));
    the driver's default 2 under-integrates P3 / P4 into a singular matrix."""
    return 2 if order <= 2 else 2 * order


def _python_driver(mesh_path, order, k=1.0):
    from feng_amd.api import FN_CONST, TRI_P1, TRI_P2, TRI_P3, TRI_P4, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, diffusion_system, solve_newton
    m = Mesh.read_msh(mesh_path)
    S, num = diffusion_system(m, {1: TRI_P1, 2: TRI_P2, 3: TRI_P3, 4: TRI_P4}[order], _degree(order), Fn(FN_CONST, k),
                              Fn(FN_CONST, -1.0))
    solve_newton(S, NLSolverOptions(), rtol=1e-12)
    u = S.get_solution()[:num.nInc]
    S.close()
    return u


@pytest.mark.parametrize("mesh", ["square1.msh", "square2.msh"])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_example1_cpp_matches_python_driver(mesh, order):
    assert EXE.exists(), "build the C++ layer first (python -m feng_amd.build)"
    path = MESHES / mesh
    out = subprocess.run([str(EXE), str(path), str(order), "1", str(_degree(order))], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    u = _python_driver(path, order)
    assert r["n"] == u.size
    assert abs(r["sum_u"] - u.sum()) <= 1e-10 * abs(u.sum())
    assert r["newton"] == 1
    assert r["max_nodal_error"] < 5e-3


def test_example1_error_falls_with_refinement_and_degree():
    errs = {}
    for mesh in ("square1.msh", "square2.msh"):
        for order in (1, 2):
            out = subprocess.run([str(EXE), str(MESHES / mesh), str(order)], capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr
            errs[(mesh, order)] = json.loads(out.stdout.strip().splitlines()[-1])["max_nodal_error"]
    assert errs[("square2.msh", 1)] < errs[("square1.msh", 1)]
    assert errs[("square1.msh", 2)] < errs[("square1.msh", 1)]


def test_example1_default_rule_of_cubic_fields_and_low_rule_refused():
    """Without a degree argument the cubic and quartic runs take degree
    2 order (a degree-2 rule would leave their stiffness singular); a rule
    below 2 (order - 1) is refused before any solve."""
    path = MESHES / "square1.msh"
    out = subprocess.run([str(EXE), str(path), "3"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    u = _python_driver(path, 3)
    assert abs(r["sum_u"] - u.sum()) <= 1e-10 * abs(u.sum())
    bad = subprocess.run([str(EXE), str(path), "4", "1", "2"], capture_output=True, text=True, timeout=120)
    assert bad.returncode != 0 and "below" in bad.stderr
