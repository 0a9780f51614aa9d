"""The reference's Taylor-Hood and vector-Laplacian convergence tables through the GPU path.

stokes_MMS.output and navier_stokes_MMS.output (tests/withLinearSolver/ of
the reference), P2-P1 divergence formulation rows: u = (x^4 y^4,
-4/5 x^3 y^5), p = x^2 y^2 on the reference's data/square1-4.msh (copied
to tests/golden/meshes/data/; the scalar tables use data/mmsMeshes/), degree-8
quadrature, u and p essential on "Bord", the manufactured source tabulated at
the quadrature points; Newton with GMRES/ILU(0) on the device; the L2 errors
of u (vector) and p by the device norms.  Both formulations: divergence
(the stress form) and Laplacian (vector diffusion + mixed gradient).  And
vectorFE_vectorLaplacian.output and vectorFE_weakGradient.output (P2 rows).
Printed by the reference with 7
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
    "stokes_laplacian": (["1.729104e-03", "2.114311e-04", "2.643873e-05", "3.316210e-06"],
                         ["1.055463e-02", "2.266154e-03", "5.362661e-04", "1.302836e-04"]),
    "navier_stokes_laplacian": (["1.729202e-03", "2.115442e-04", "2.645172e-05", "3.317293e-06"],
                                ["1.064192e-02", "2.276928e-03", "5.376581e-04", "1.304588e-04"]),
}
# vectorFE_vectorLaplacian.output, Lagrange P2 rows: L2 and H1 (semi-norm) errors
VECTOR_LAPLACIAN = (["1.183067e-03", "1.454727e-04", "1.812249e-05", "2.266050e-06"],
                    ["3.800807e-02", "9.511922e-03", "2.380397e-03", "5.955249e-04"])


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


@pytest.mark.parametrize("case", ["stokes", "navier_stokes", "stokes_laplacian", "navier_stokes_laplacian"])
def test_reference_mms_table(case):
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_L2_ERROR, NORM_VECTOR_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton, taylor_hood_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    conv = case.startswith("navier_stokes")
    eu, ep = [], []
    for lvl in range(1, 5):
        print(case, "square", lvl, flush=True)
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        src = [Fn(FN_QP, data=_qp_table(m, 8, lambda x, y, c=c: _src(x, y, conv)[c])) for c in range(2)]
        ubc = lambda x, y, z, comp: np.where(comp == 0, x ** 4 * y ** 4, -4.0 / 5.0 * x ** 3 * y ** 5)
        pbc = lambda x, y, z, comp: x * x * y * y
        S, num = taylor_hood_system(m, 8, 1.0, -1.0 if conv else None, src, ubc, pbc,
                                    laplacian=case.endswith("laplacian"))
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


def test_reference_vector_laplacian_table():
    """-div(grad u) = f, u = (a x^6 + x^5 y^2, b sin(xy)), a = 0.1234,
    b = 1.2345, strong Dirichlet values, vector P2 (vectorFE.cpp:132-300):
    L2 and vector semi-H1 errors."""
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_VECTOR_L2_ERROR, NORM_VECTOR_SEMI_H1_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton, vector_laplacian_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    a, b = 0.1234, 1.2345
    usol = (lambda x, y: a * x ** 6 + x ** 5 * y * y, lambda x, y: b * np.sin(x * y))
    grad = (lambda x, y: 6 * a * x ** 5 + 5 * x ** 4 * y * y, lambda x, y: y * b * np.cos(x * y),
            lambda x, y: 2 * x ** 5 * y, lambda x, y: x * b * np.cos(x * y))
    src = (lambda x, y: 30 * a * x ** 4 + 20 * x ** 3 * y * y + 2 * x ** 5,
           lambda x, y: -b * np.sin(x * y) * (x * x + y * y))
    el2, eh1 = [], []
    for lvl in range(1, 5):
        print("vector laplacian square", lvl, flush=True)
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        S, num = vector_laplacian_system(m, 8, [Fn(FN_QP, data=_qp_table(m, 8, f)) for f in src],
                                         lambda x, y, z, c: np.where(c == 0, usol[0](x, y), usol[1](x, y)))
        rep = solve_newton(S, NLSolverOptions(1e-12, 1e-12, 1e4, 10), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13,
                           restart=100, max_iter=20000)
        assert rep.converged
        el2.append(S.norm(NORM_VECTOR_L2_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, f)) for f in usol]))
        eh1.append(S.norm(NORM_VECTOR_SEMI_H1_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, f)) for f in grad]))
        S.close()
    assert all(_close(g, r) for g, r in zip(el2, VECTOR_LAPLACIAN[0])), (el2, VECTOR_LAPLACIAN[0])
    assert all(_close(g, r) for g, r in zip(eh1, VECTOR_LAPLACIAN[1])), (eh1, VECTOR_LAPLACIAN[1])


# vectorFE_weakGradient.output, Lagrange P2 rows
WEAK_GRADIENT = ["3.578664e-03", "4.543448e-04", "5.759571e-05", "7.267833e-06"]


def test_reference_weak_gradient_table():
    """u = grad f in the weak sense (VectorMass + GradSource), vector P2,
    strong Dirichlet values, a..d = 0.1234, 1.2345, 2.3456, 3.4567
    (vectorFE.cpp:19-90)."""
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_VECTOR_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton, weak_gradient_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    a, b, c, d = 0.1234, 1.2345, 2.3456, 3.4567
    usol = (lambda x, y: a * (x * y + np.cos(x)) + b * y + 3 * c * x * x,
            lambda x, y: a / 2 * x * x + b * x + 3 * d * y * y + 2 * d * y * np.cos(y * y))
    f = lambda x, y: -(a / 2 * x * x * y + a * np.sin(x) + b * x * y + c * x ** 3 + d * y ** 3 + d * np.sin(y * y))
    err = []
    for lvl in range(1, 5):
        print("weak gradient square", lvl, flush=True)
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        S, num = weak_gradient_system(m, 8, Fn(FN_QP, data=_qp_table(m, 8, f)),
                                      lambda x, y, z, cc: np.where(cc == 0, usol[0](x, y), usol[1](x, y)))
        rep = solve_newton(S, NLSolverOptions(1e-12, 1e-12, 1e4, 10), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13,
                           restart=100, max_iter=20000)
        assert rep.converged
        err.append(S.norm(NORM_VECTOR_L2_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, g)) for g in usol]))
        S.close()
    assert all(_close(g, r) for g, r in zip(err, WEAK_GRADIENT)), (err, WEAK_GRADIENT)
