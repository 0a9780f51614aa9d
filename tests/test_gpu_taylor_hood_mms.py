"""The reference's Taylor-Hood convergence tables through the GPU path.

stokes_MMS.output and navier_stokes_MMS.output (tests/withLinearSolver/ of
the reference), P2-P1 divergence formulation rows: u = (x^4 y^4,
-4/5 x^3 y^5), p = x^2 y^2 on the reference's data/square1-4.msh (copied
to tests/golden/meshes/data/; the scalar tables use data/mmsMeshes/), degree-8
quadrature, u and p essential on "Bord", the manufactured source tabulated at
the quadrature points; Newton with GMRES/ILU(0) on the device; the L2 errors
of u (vector) and p by the device norms.  Printed by the reference with 7
significant digits.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLES = {  # (velocity, pressure) L2 errors, square1..4
    "stokes": (["1.730482e-03", "2.101765e-04", "2.593560e-05", "3.221420e-06"],
               ["1.054679e-02", "2.304575e-03", "5.426010e-04", "1.313177e-04"]),
    "navier_stokes": (["1.730601e-03", "2.102148e-04", "2.593958e-05", "3.221725e-06"],
                      ["1.064397e-02", "2.316230e-03", "5.440602e-04", "1.315003e-04"]),
}


def _u(x, y):
    return x ** 4 * y ** 4, -4.0 / 5.0 * x ** 3 * y ** 5


def _src(x, y, convective):
    mdpdx = (-2.0 * x * y * y, -2.0 * x * x * y)
    lap = (12.0 * (x * x * y ** 4 + x ** 4 * y * y), -4.0 / 5.0 * (6.0 * x * y ** 5 + 20.0 * x ** 3 * y ** 3))
    ugu = (0.0, 0.0)
    if convective:
        u0, u1 = _u(x, y)
        g = ((4.0 * x ** 3 * y ** 4, -12.0 * x * x * y ** 5 / 5.0), (4.0 * x ** 4 * y ** 3, -4.0 * x ** 3 * y ** 4))
        ugu = (u0 * g[0][0] + u1 * g[1][0], u0 * g[0][1] + u1 * g[1][1])
    return tuple(-(-ugu[c] + mdpdx[c] + lap[c]) for c in range(2))


def _close(got, ref):
    return abs(got - float(ref)) <= 1e-6 * float(ref)


@pytest.mark.parametrize("case", ["stokes", "navier_stokes"])
def test_reference_mms_table(case):
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_L2_ERROR, NORM_VECTOR_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton, taylor_hood_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    conv = case == "navier_stokes"
    eu, ep = [], []
    for lvl in range(1, 5):
        print(case, "square", lvl, flush=True)
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        src = [Fn(FN_QP, data=_qp_table(m, 8, lambda x, y, c=c: _src(x, y, conv)[c])) for c in range(2)]
        ubc = lambda x, y, z, comp: np.where(comp == 0, x ** 4 * y ** 4, -4.0 / 5.0 * x ** 3 * y ** 5)
        pbc = lambda x, y, z, comp: x * x * y * y
        S, num = taylor_hood_system(m, 8, 1.0, -1.0 if conv else None, src, ubc, pbc)
        rep = solve_newton(S, NLSolverOptions(1e-12, 1e-12, 1e4, 20), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13,
                           restart=100, max_iter=20000)
        assert rep.converged
        uex = [Fn(FN_QP, data=_qp_table(m, 8, lambda x, y, c=c: _u(x, y)[c])) for c in range(2)]
        pex = Fn(FN_QP, data=_qp_table(m, 8, lambda x, y: x * x * y * y))
        eu.append(S.norm(NORM_VECTOR_L2_ERROR, 0, 8, uex))
        ep.append(S.norm(NORM_L2_ERROR, 1, 8, [pex]))
        S.close()
    ru, rp = TABLES[case]
    assert all(_close(g, r) for g, r in zip(eu, ru)), (eu, ru)
    assert all(_close(g, r) for g, r in zip(ep, rp)), (ep, rp)
