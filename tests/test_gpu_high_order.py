"""Cubic and quartic Lagrange triangles (FE_TRI_P3 / FE_TRI_P4, feSpaceTriP3
and feSpaceTriP4) on the device: the P3 and P4 rows of the reference's golden
tables scalarFE_diffusion.output, scalarFE_nonlineardiffusion.output,
scalarFE_advdiff.output and scalarFE_advdiffreact.output through the fused
scalar sweep, Newton with GMRES/ILU(0) and the device L2 norm; the assembled
systems against the oracle; norms of polynomial fields the spaces hold
exactly; the dimension-2 rows (P1 to P4) of interpolationScalar.output
through the device norm."""
import math

import numpy as np
import pytest

from tests.test_oracle_high_order import DEGREE, DIFFUSION_HO, NLDIFF_HO

pytestmark = pytest.mark.gpu

def _newton(S, symmetric=False, rtol=1e-13, tol=1e-12):
    """Newton with AMG-PCG (symmetric systems) or GMRES/ILU(0), to a relative
    residual these systems reach (1e-14 is below the FP64 floor of the P3/P4
    diffusion matrices: the Krylov solvers stagnate there; the
    advection-diffusion ones reach it).  A Newton tolerance below the first
    correction's accuracy adds refinement steps with the FP64 residual."""
    from feng_amd.api import KSP_CG, KSP_GMRES, PC_AMG, PC_ILU0
    from feng_amd.problems import NLSolverOptions, solve_newton
    if symmetric:
        kw = dict(ksp=KSP_CG, pc=PC_AMG, rtol=rtol, max_iter=5000)
    else:
        kw = dict(ksp=KSP_GMRES, pc=PC_ILU0, rtol=rtol, restart=200, max_iter=5000)
    rep = solve_newton(S, NLSolverOptions(tol, tol, 1e4, 20), **kw)
    assert rep.converged
    return rep


def _matches(err, ref):
    """The 7 printed digits, or 5e-15 absolute: an error below 1e-9 carries
    the rounding of the assembly and the solve (different summation orders
    here and in the reference's direct solver) in its 7th digit."""
    return all(abs(e - r) <= max(1e-6 * r, 5e-15) for e, r in zip(err, ref))


@pytest.mark.parametrize("nonlinear", [False, True])
@pytest.mark.parametrize("kind", [5, 6])
def test_reference_diffusion_rows(kind, nonlinear):
    from feng_amd.api import FN_CONST, FN_POLY_U, FN_QP, FN_SINSIN, NORM_L2_ERROR, Fn, Mesh
    from feng_amd.problems import diffusion_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_nonlinear_diffusion import DK_POLY, K_POLY, source
    from tests.test_oracle_time import _qp_table
    print("high-order diffusion", kind, nonlinear, flush=True)
    bc = lambda x, y, z, c: np.sin(np.pi * x) * np.sin(np.pi * y)
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        if nonlinear:
            S, num = diffusion_system(m, kind, 8, Fn(FN_POLY_U, data=K_POLY),
                                      Fn(FN_QP, data=_qp_table(m, 8, lambda x, y: source(x, y))), bc,
                                      dk=Fn(FN_POLY_U, data=DK_POLY))
        else:
            S, num = diffusion_system(m, kind, 8, Fn(FN_CONST, 1.0), Fn(FN_SINSIN, -2 * math.pi ** 2), bc)
        _newton(S, symmetric=not nonlinear)
        err.append(S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_SINSIN, 1.0)]))
        S.close()
    assert _matches(err, (NLDIFF_HO if nonlinear else DIFFUSION_HO)[DEGREE[kind]]), err


@pytest.mark.parametrize("react", [False, True])
@pytest.mark.parametrize("kind", [5, 6])
def test_reference_advdiff_rows(kind, react):
    from feng_amd.api import FN_QP, NORM_L2_ERROR, Fn, Mesh
    from tests.conftest import GOLDEN
    from tests.test_gpu_advection import _system
    from tests.test_oracle_advection import ADVDIFF, u_exact
    from tests.test_oracle_time import _qp_table
    print("high-order advdiff", kind, react, flush=True)
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        S, num = _system(m, kind, react)
        _newton(S, rtol=1e-14, tol=1e-14)
        err.append(S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, lambda x, y: u_exact(x, y)))]))
        S.close()
    assert _matches(err, ADVDIFF[(react, kind)]), err


@pytest.mark.parametrize("kind", [5, 6])
def test_high_order_assembly_matches_oracle(oracle, kind):
    from feng_amd.api import (ADVECTION, DIFFUSION, FN_CONST, FN_POLY_U, FN_QP, MASS, NONLINEAR_DIFFUSION, SOURCE,
                              TRANSIENT_MASS, Fn, LinearSystemHIP, Mesh, Numbering)
    from tests.test_oracle_nonlinear_diffusion import DK_POLY, K_POLY
    from tests.test_oracle_time import _qp_table
    print("high-order assembly", kind, flush=True)
    m = Mesh.square(7, 0.2, 6)
    deg = 7
    nl = {5: 10, 6: 15}[kind]
    num = Numbering(m, [kind], [1])
    n = num.nInc
    rng = np.random.default_rng(kind)
    u = rng.normal(size=num.nDofs)
    past = rng.normal(size=num.nDofs)
    tab = _qp_table(m, deg, lambda x, y: np.cos(x + 2 * y))
    S = LinearSystemHIP(0)
    S.upload_mesh(m)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(n, num.nDofs)
    S.add_form(DIFFUSION, [f], deg, [Fn(FN_CONST, 1.3)])
    S.add_form(NONLINEAR_DIFFUSION, [f], deg, [Fn(FN_POLY_U, data=K_POLY), Fn(FN_POLY_U, data=DK_POLY)])
    S.add_form(ADVECTION, [f], deg, [Fn(FN_CONST, 0.7), Fn(FN_CONST, -1.1)])
    S.add_form(MASS, [f], deg, [Fn(FN_CONST, 2.0)])
    S.add_form(TRANSIENT_MASS, [f], deg, [Fn(FN_CONST, 0.5)])
    S.add_form(SOURCE, [f], deg, [Fn(FN_QP, data=tab)])
    S.build_csr()
    S.set_solution(u)
    S.set_past_solution(1, past)
    S.set_bdf([40.0, -40.0])
    S.setToZero()
    S.assemble(3)
    val, rhs = S.get_system()
    ia, ja = S.get_csr()
    x, c = m.xyz, m.conn
    a = num.adr[0]
    v0, r0 = oracle.assemble_diffusion(kind, x, c, a, n, ia, ja, deg, (0, 1.3, 0.0, None), (5, 0.0, 0.0, tab), u)
    oracle.assemble_nonlinear_diffusion(kind, x, c, a, n, ia, ja, deg, K_POLY, DK_POLY, u, v0, r0)
    oracle.assemble_advection(kind, x, c, a, n, ia, ja, deg, [(0, 0.7, 0.0, None), (0, -1.1, 0.0, None)], u, v0, r0)
    oracle.assemble_mass_power(kind, x, c, a, n, ia, ja, deg, (0, 2.0, 0.0, None), 1.0, u, v0, r0)
    oracle.assemble_transient_mass(kind, x, c, a, n, ia, ja, deg, (0, 0.5, 0.0, None), 40.0, 40.0 * (u - past), v0,
                                   r0)
    assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0)
    assert ja.size == ia[-1] and nl * nl * m.E >= ja.size
    S.close()


@pytest.mark.parametrize("kind", [5, 6])
def test_norms_of_polynomial_fields(kind):
    from feng_amd.api import FN_CONST, FN_QP, NORM_AREA, NORM_L2, NORM_L2_ERROR, NORM_SEMI_H1, Fn, Mesh
    from feng_amd.problems import diffusion_system
    from tests.test_oracle_time import _qp_table
    n = DEGREE[kind]
    m = Mesh.square(6, 0.2, 2)
    S, num = diffusion_system(m, kind, 8, Fn(FN_CONST, 1.0), Fn(FN_CONST, 0.0))
    xyz = num.dof_xyz.reshape(-1, 3)
    # x^n + y: held exactly by the space; its L2 norm and gradient norm on the unit square
    S.set_solution(xyz[:, 0] ** n + xyz[:, 1])
    l2sq = 1 / (2 * n + 1) + 2 / ((n + 1) * 2) + 1 / 3
    grad2 = n * n / (2 * n - 1) + 1
    assert abs(S.norm(NORM_AREA, 0, 8) - 1) <= 1e-13
    assert abs(S.norm(NORM_L2, 0, 8) - math.sqrt(l2sq)) <= 1e-12
    assert abs(S.norm(NORM_SEMI_H1, 0, 8) - math.sqrt(grad2)) <= 1e-12
    tab = _qp_table(m, 8, lambda x, y: x ** n + y)
    assert S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_QP, data=tab)]) <= 1e-13
    S.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_reference_interpolation_rows(order):
    """interpolationScalar.output, dimension-2 rows: the L2 error of the
    nodal interpolant of x^6 + y^6 by the device norm (degree-12 rule)."""
    from feng_amd.api import FN_CONST, FN_QP, NORM_L2_ERROR, Fn, Mesh
    from feng_amd.problems import diffusion_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_high_order import INTERP, KIND_OF
    from tests.test_oracle_time import _qp_table
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        S, num = diffusion_system(m, KIND_OF[order], 12, Fn(FN_CONST, 1.0), Fn(FN_CONST, 0.0))
        xyz = num.dof_xyz.reshape(-1, 3)
        S.set_solution(xyz[:, 0] ** 6 + xyz[:, 1] ** 6)
        err.append(S.norm(NORM_L2_ERROR, 0, 12, [Fn(FN_QP, data=_qp_table(m, 12, lambda x, y: x ** 6 + y ** 6))]))
        S.close()
    assert _matches(err, INTERP[order]), err


def test_reference_vector_interpolation_p2_row():
    """interpolationVector.output, dimension 2, vector P2 row: the device
    VECTOR_L2_ERROR of the nodal interpolant of (sin xy, cos xy) on
    data/square1-4.msh (degree-12 rule), on the Taylor-Hood velocity field."""
    from feng_amd.api import FN_QP, NORM_VECTOR_L2_ERROR, Fn, Mesh
    from feng_amd.problems import navier_stokes_system
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    ref = [1.478830e-04, 1.848170e-05, 2.310065e-06, 2.887533e-07]
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / "data" / f"square{lvl}.msh")
        S, num = navier_stokes_system(m, 40.0, 12)
        xyz = num.dof_xyz.reshape(-1, 3)
        xy = xyz[:, 0] * xyz[:, 1]
        vel = num.dof_field == 0
        u = np.zeros(num.nDofs)
        u[vel] = np.where(num.dof_comp[vel] == 0, np.sin(xy[vel]), np.cos(xy[vel]))
        S.set_solution(u)
        err.append(S.norm(NORM_VECTOR_L2_ERROR, 0, 12, [Fn(FN_QP, data=_qp_table(m, 12, lambda x, y: np.sin(x * y))),
                                                        Fn(FN_QP, data=_qp_table(m, 12, lambda x, y: np.cos(x * y)))]))
        S.close()
    assert _matches(err, ref), err


@pytest.mark.parametrize("kind", [5, 6])
def test_cubic_and_quartic_diffusion_on_the_matrix_cores(oracle, kind):
    """Diffusion + source on cubic / quartic triangles takes the MFMA sweep
    (GEMM form, 7 / 15 tiles of 16 entries per element, asmDiffusionMfma):
    values and RHS equal the oracle's element loop to 1e-12 and the scalar
    sweep (asm_kernel 0) to 1e-13; its FP32 instance (element matrices
    computed in FP32 on the matrix cores) is within 1e-6 of the FP64 values."""
    from feng_amd.api import DIFFUSION, FN_ELEM, FN_QP, PREC_FP32_IR, SOURCE, Fn, LinearSystemHIP, Mesh, Numbering
    from tests.test_oracle_time import _qp_table
    print("high-order MFMA diffusion", kind, flush=True)
    m = Mesh.square(9, 0.2, 3)
    deg = 2 * (kind - 3)  # 4 (P3), 6 (P4): exact for the stiffness of a straight element
    num = Numbering(m, [kind], [1])
    n = num.nInc
    k = m.coeff_blocks(3, 0.1, 10.0, 7)
    tab = _qp_table(m, deg, lambda x, y: np.sin(3 * x) * np.cos(2 * y))
    S = LinearSystemHIP(0)
    S.upload_mesh(m)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(n, num.nDofs)
    S.add_form(DIFFUSION, [f], deg, [Fn(FN_ELEM, data=k)])
    S.add_form(SOURCE, [f], deg, [Fn(FN_QP, data=tab)])
    S.build_csr()
    u = np.random.default_rng(kind).normal(size=num.nDofs)
    S.set_solution(u)
    out = {}
    for ak in (1, 0):
        S.set_tuning("asm_kernel", ak)
        S.setToZero()
        S.assemble(3)
        out[ak] = S.get_system()
    ia, ja = S.get_csr()
    v0, r0 = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, deg, (4, 0.0, 0.0, k),
                                       (5, 0.0, 0.0, tab), u)
    (vm, rm), (vs, rs) = out[1], out[0]
    print("MFMA vs oracle", np.linalg.norm(vm - v0) / np.linalg.norm(v0), "vs scalar",
          np.linalg.norm(vm - vs) / np.linalg.norm(vs), flush=True)
    assert np.linalg.norm(vm - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.linalg.norm(rm - r0) <= 1e-12 * np.linalg.norm(r0)
    assert np.linalg.norm(vm - vs) <= 1e-13 * np.linalg.norm(vs)
    assert np.linalg.norm(rm - rs) <= 1e-13 * np.linalg.norm(rs)
    S.set_tuning("asm_kernel", 1)
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    v32 = S.get_system32()
    err = np.abs(v32.astype(np.float64) - v0).max() / np.abs(v0).max()
    print("FP32 MFMA values vs FP64 oracle", err, flush=True)
    assert err <= 1e-6 and not np.array_equal(v32, vm.astype(np.float32))
    S.close()
