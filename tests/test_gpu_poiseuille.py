"""The reference's Navier-Stokes Poiseuille test (navier_stokes.cpp:175-300,
373-378) on the device: its mesh data/poiseuille0.msh with the physical
groups Domain / Inlet / Outlet / NoSlip, Taylor-Hood P2-P1, U essential on
Inlet (the exact profile) and NoSlip (0), the outlet's second velocity
component constrained to 0 (uOutlet->setEssentialComponent(1, true)), the
divergence formulation (VectorConvectiveAcceleration (1), MixedDivergence
{p, u} (1), DivergenceNewtonianStress {u, p} (-1, 1)), degree-8 quadrature,
Newton with the reference's options.  The Poiseuille flow u = (y (H - y) / 2,
0), p = L - x (H = 1, L = 5, dp/dx = -1) lies in the discrete spaces, so the
reference asserts L2 errors below 1e-13 for u and p; the device's Newton with
GMRES(50) and the block Schur preconditioner meets the same bound."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_poiseuille_exact_on_the_reference_mesh():
    from feng_amd import api as A
    from feng_amd.api import FN_CONST, FN_QP, KSP_GMRES, NORM_L2_ERROR, NORM_VECTOR_L2_ERROR, PC_FIELDSPLIT, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    print("poiseuille", flush=True)
    m = Mesh.read_msh(GOLDEN / "meshes" / "data" / "poiseuille0.msh")
    num = A.Numbering(m, [A.TRI_VP2, A.TRI_P1], [["Inlet", "NoSlip"], []])
    S = A.LinearSystemHIP(0)
    S.upload_mesh(m)
    fu = S.add_field(A.TRI_VP2, num.adr[0])
    fp = S.add_field(A.TRI_P1, num.adr[1])
    S.set_sizes(num.nInc, num.nDofs)
    deg = 8
    S.add_form(A.VECTOR_CONVECTIVE_ACCELERATION, [fu], deg, [Fn(FN_CONST, 1.0)])
    S.add_form(A.MIXED_DIVERGENCE, [fp, fu], deg, [Fn(FN_CONST, 1.0)])
    S.add_form(A.DIV_NEWTONIAN_STRESS, [fu, fp], deg, [Fn(FN_CONST, -1.0), Fn(FN_CONST, 1.0)])
    S.build_csr()
    H, L, dpdx = 1.0, 5.0, -1.0
    usol = (lambda x, y: -dpdx / 2.0 * y * (H - y), lambda x, y: 0.0 * x)
    psol = lambda x, y: -dpdx * (L - x)
    S.set_solution(A.essential_values(num, [lambda x, y, z, c: np.where(c == 0, usol[0](x, y), 0.0), None]))
    d = num.dof_xyz.reshape(-1, 3)
    outlet_v = np.flatnonzero((np.arange(num.nDofs) < num.nInc) & (num.dof_field == 0) & (num.dof_comp == 1)
                              & np.isclose(d[:, 0], L))
    assert outlet_v.size > 0
    S.set_constrained_dofs(outlet_v)
    steps = []
    rep = solve_newton(S, NLSolverOptions(1e-10, 1e-10, 1e4, 10), progress=lambda it, r: steps.append(r),
                       ksp=KSP_GMRES, pc=PC_FIELDSPLIT, rtol=1e-13, restart=50, max_iter=5000)
    assert rep.converged and all(r.converged for r in steps)
    eu = S.norm(NORM_VECTOR_L2_ERROR, fu, deg, [Fn(FN_QP, data=_qp_table(m, deg, f)) for f in usol])
    ep = S.norm(NORM_L2_ERROR, fp, deg, [Fn(FN_QP, data=_qp_table(m, deg, psol))])
    print(f"Newton {rep.iterations}, GMRES {[r.iterations for r in steps]}, L2 error u {eu:.2e}, p {ep:.2e}",
          flush=True)
    assert eu < 1e-13 and ep < 1e-13
    S.close()
