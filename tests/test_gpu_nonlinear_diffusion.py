"""NONLINEAR_DIFFUSION on the device (feSysElm.cpp:681-760): the reference's
golden table scalarFE_nonlineardiffusion.output (P1, P2 on
mmsMeshes/square1-4.msh, k(u) = 1 + u^2 as FN_POLY_U coefficients) through
the fused scalar sweep, Newton with GMRES/ILU(0) (the Jacobian is not
symmetric) and the device L2 norm; the assembled system against the oracle on
triangles and tetrahedra, with every kernel option."""
import numpy as np
import pytest

from tests.test_oracle_nonlinear_diffusion import DK_POLY, K_POLY, NLDIFF, matches, oracle_nldiff, source, u_exact

pytestmark = pytest.mark.gpu


def _system(m, kind, deg=8, src=None):
    from feng_amd.api import FN_POLY_U, FN_QP, Fn
    from feng_amd.problems import diffusion_system
    from tests.test_oracle_time import _qp_table
    if src is None:
        src = Fn(FN_QP, data=_qp_table(m, deg, lambda x, y: source(x, y)))
    return diffusion_system(m, kind, deg, Fn(FN_POLY_U, data=K_POLY), src, lambda x, y, z, c: u_exact(x, y),
                            dk=Fn(FN_POLY_U, data=DK_POLY))


@pytest.mark.parametrize("kind", [1, 2])
def test_reference_nonlinear_diffusion_table(oracle, kind):
    from feng_amd.api import FN_SINSIN, KSP_GMRES, NORM_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton
    from tests.conftest import GOLDEN
    print("nonlinear diffusion", kind, flush=True)
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        S, num = _system(m, kind)
        rep = solve_newton(S, NLSolverOptions(1e-12, 1e-12, 1e4, 20), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13,
                           restart=100, max_iter=20000)
        assert rep.converged and rep.iterations > 2  # a nonlinear problem
        err.append(S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_SINSIN, 1.0)]))
        if lvl == 2:
            _, uo, _ = oracle_nldiff(oracle, m, kind)
            u = S.get_solution()
            assert np.abs(u - uo).max() <= 1e-10 * np.abs(uo).max()
        S.close()
    assert matches(err, NLDIFF[kind]), err


def _oracle_system(oracle, m, kind, num, ia, ja, u, deg, kpoly, src):
    v0, r0 = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, deg, (0, 0.0, 0.0, None),
                                       src, u)
    oracle.assemble_nonlinear_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, deg, kpoly,
                                        [0.0] if len(kpoly) == 1 else DK_POLY, u, v0, r0)
    return v0, r0


@pytest.mark.parametrize("kind", [1, 2, 4])
def test_nonlinear_diffusion_assembly_matches_oracle(oracle, kind):
    from feng_amd.api import FN_CONST, FN_QP, FeError, Fn, Mesh
    from tests.test_oracle_time import _qp_table
    print("nonlinear diffusion assembly", kind, flush=True)
    deg = 5 if kind != 4 else 2  # the tetrahedron rules go to degree 2
    m = Mesh.cube(5) if kind == 4 else Mesh.square(9, 0.2, 6)
    if kind == 4:
        srcfn, src = Fn(FN_CONST, 1.5), (0, 1.5, 0.0, None)
    else:
        tab = _qp_table(m, deg, lambda x, y: source(x, y))
        srcfn, src = Fn(FN_QP, data=tab), (5, 0.0, 0.0, tab)
    S, num = _system(m, kind, deg, srcfn)
    rng = np.random.default_rng(kind)
    u = S.get_solution()
    u[:num.nInc] = rng.normal(size=num.nInc)
    S.set_solution(u)
    ia, ja = S.get_csr()
    v0, r0 = _oracle_system(oracle, m, kind, num, ia, ja, u, deg, K_POLY, src)
    # the lane-per-row and MFMA options do not apply to this form: the sweep
    # runs the scalar kernel whatever they say
    for key, value in (("scalar_lanes", 0), ("scalar_lanes", 1), ("asm_kernel", 1), ("asm_kernel", 0)):
        S.set_tuning(key, value)
        S.setToZero()
        S.assemble(3)
        val, rhs = S.get_system()
        assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0), (kind, key, value)
        assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0), (kind, key, value)
    # new polynomials through fe_hip_update_coeff (form 0, the nonlinear
    # diffusion): k = 2, dk/du = 0, the linear form
    S.update_coeff(0, 0, [2.0, 0.0, 0.0])
    S.update_coeff(0, 1, [0.0, 0.0])
    S.setToZero()
    S.assemble(3)
    val, rhs = S.get_system()
    v1, r1 = _oracle_system(oracle, m, kind, num, ia, ja, u, deg, [2.0], src)
    assert np.linalg.norm(val - v1) <= 1e-12 * np.linalg.norm(v1)
    assert np.linalg.norm(rhs - r1) <= 1e-12 * np.linalg.norm(r1)
    with pytest.raises(FeError):
        S.update_coeff(0, 0, [1.0, 2.0])  # the polynomial keeps its length
    S.close()


def test_poly_coefficient_only_for_nonlinear_diffusion():
    from feng_amd.api import (DIFFUSION, FN_CONST, FN_POLY_U, NONLINEAR_DIFFUSION, FeError, Fn, LinearSystemHIP, Mesh,
                              Numbering)
    m = Mesh.square(4)
    num = Numbering(m, [1], [1])
    S = LinearSystemHIP(0)
    S.upload_mesh(m)
    f = S.add_field(1, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    with pytest.raises(FeError):
        S.add_form(DIFFUSION, [f], 2, [Fn(FN_POLY_U, data=[1.0, 0.0, 1.0])])
    with pytest.raises(FeError):
        S.add_form(NONLINEAR_DIFFUSION, [f], 2, [Fn(FN_POLY_U, data=np.ones(9)), Fn(FN_CONST, 0.0)])
    S.close()
