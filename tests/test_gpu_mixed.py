"""Mixed scalar forms on the device (MIXED_MASS, MIXED_MASS_POWER,
MIXED_GRADGRAD, feSysElm.cpp:283-403, :886-976): the reference's golden
table scalarFE_mixedDiffusion.output (P1 to P4 on mmsMeshes/square1-4.msh,
both fields) through the device assembly and the device norm; the assembled
coupled system against the oracle on P1, P2, P3 triangles and P1 tets.

The coupled system's diagonal blocks are mass matrices (O(h^2)) against O(1)
grad-grad couplings: ILU(0) pivots on them and the device GMRES/ILU(0)
converges only on the coarsest mesh with P1 / P2, where it is checked; the
table's Newton steps solve on the device with the direct solve (FE_KSP_DIRECT,
dense LU with partial pivoting), the role of the reference's direct solver;
the direct solve equals a host sparse LU of the same system."""
import numpy as np
import pytest

from tests.test_oracle_mixed import KIND_OF, MIXED, NF, matches, u_exact, u_source, v_exact, v_source

pytestmark = pytest.mark.gpu


def _system(m, kind, deg=8):
    from feng_amd.api import FN_QP, Fn
    from feng_amd.problems import mixed_diffusion_system
    from tests.test_oracle_time import _qp_table
    return mixed_diffusion_system(m, kind, deg, Fn(FN_QP, data=_qp_table(m, deg, u_source)),
                                  Fn(FN_QP, data=_qp_table(m, deg, v_source)),
                                  lambda x, y, z, c: u_exact(x, y), lambda x, y, z, c: v_exact(x, y))


def _direct_newton(S, tol=1e-13):
    """Newton on the device (feNonLinearSolver's loop), every step solved by
    the device direct solve."""
    from feng_amd.api import KSP_DIRECT
    from feng_amd.problems import NLSolverOptions, solve_newton
    rep = solve_newton(S, NLSolverOptions(tol, tol, 1e4, 20), ksp=KSP_DIRECT, rtol=1e-12)
    assert rep.converged and all(r.converged for r in rep.linear)
    return rep.iterations


def test_direct_solve_equals_host_sparse_lu():
    """One Newton system of the P2 mixed problem on square3.msh (1922
    unknowns, nonsymmetric): the device direct solve equals scipy's sparse LU
    to 1e-12 and its FP64 residual is at round-off."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from feng_amd.api import KSP_DIRECT, Mesh
    from tests.conftest import GOLDEN
    m = Mesh.read_msh(GOLDEN / "meshes" / "square3.msh")
    S, num = _system(m, KIND_OF[2])
    n = num.nInc
    S.setToZero()
    S.assemble(3)
    ia, ja = S.get_csr()
    val, rhs = S.get_system()
    A = sp.csr_matrix((val, ja, ia), shape=(n, n))
    assert abs(A - A.T).max() > 1e-8 * abs(A).max()  # a nonsymmetric system
    r, du = S.solve(ksp=KSP_DIRECT, rtol=1e-12, want_du=True)
    ref = spla.spsolve(A.tocsc(), rhs)
    print("direct vs host LU", np.abs(du - ref).max() / np.abs(ref).max(), "relres", r.relres, flush=True)
    assert r.converged and r.iterations == 1 and r.relres <= 1e-13
    assert np.abs(du - ref).max() <= 1e-12 * np.abs(ref).max()
    S.close()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_reference_mixed_diffusion_table(order):
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    print("mixed diffusion", order, flush=True)
    kind = KIND_OF[order]
    err = {0: [], 1: []}
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        S, num = _system(m, kind)
        assert _direct_newton(S) > 2  # a nonlinear problem
        for f, ex in ((0, u_exact), (1, v_exact)):
            err[f].append(S.norm(NORM_L2_ERROR, f, 8, [Fn(FN_QP, data=_qp_table(m, 8, ex))]))
        S.close()
        if lvl == 1 and order <= 2:
            # the device Newton / GMRES-ILU(0) path where ILU(0) is stable
            S, num = _system(m, kind)
            rep = solve_newton(S, NLSolverOptions(1e-13, 1e-13, 1e4, 20), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-14,
                               restart=200, max_iter=5000)
            assert rep.converged and rep.iterations > 2
            e0 = S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, u_exact))])
            assert abs(e0 - err[0][-1]) <= max(1e-9 * err[0][-1], 5e-15)
            S.close()
    assert matches(err[0], MIXED[(order, 0)]), err[0]
    assert matches(err[1], MIXED[(order, 1)]), err[1]


@pytest.mark.parametrize("kind", [1, 2, 5, 4])
def test_mixed_assembly_matches_oracle(oracle, kind):
    from feng_amd.api import FN_CONST, FN_QP, Fn, Mesh
    from feng_amd.problems import mixed_diffusion_system
    from tests.test_oracle_time import _qp_table
    print("mixed assembly", kind, flush=True)
    deg = 2 if kind == 4 else 6
    m = Mesh.cube(4) if kind == 4 else Mesh.square(7, 0.2, 9)
    if kind == 4:
        su, sv = Fn(FN_CONST, 0.3), Fn(FN_CONST, -0.6)
        osu, osv = (0, 0.3, 0.0, None), (0, -0.6, 0.0, None)
    else:
        tu, tv = _qp_table(m, deg, u_source), _qp_table(m, deg, v_source)
        su, sv = Fn(FN_QP, data=tu), Fn(FN_QP, data=tv)
        osu, osv = (5, 0.0, 0.0, tu), (5, 0.0, 0.0, tv)
    S, num = mixed_diffusion_system(m, kind, deg, su, sv)
    n, nf = num.nInc, NF.get(kind, 4)
    au, av = num.adr
    u = S.get_solution()
    u[:n] = np.random.default_rng(kind).normal(size=n)
    S.set_solution(u)
    S.setToZero()
    S.assemble(3)
    val, rhs = S.get_system()
    ia, ja = S.get_csr()
    X, T = m.xyz, m.conn
    v0, r0 = oracle.assemble_diffusion(kind, X, T, au, n, ia, ja, deg, (0, 0.0, 0.0, None), osu, u)
    v2, r2 = oracle.assemble_diffusion(kind, X, T, av, n, ia, ja, deg, (0, 0.0, 0.0, None), osv, u)
    v0 += v2
    r0 += r2
    oracle.assemble_mass_power(kind, X, T, au, n, ia, ja, deg, (0, 1.0, 0.0, None), 1.0, u, v0, r0)
    oracle.assemble_mixed(kind, X, T, au, av, n, ia, ja, deg, 0.0, 1.0, 3.0, -1.0, u, v0, r0)
    oracle.assemble_mass_power(kind, X, T, av, n, ia, ja, deg, (0, 1.0, 0.0, None), 1.0, u, v0, r0)
    oracle.assemble_mixed(kind, X, T, av, au, n, ia, ja, deg, 1.0, 0.0, 1.0, -1.0, u, v0, r0)
    assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0)
    S.close()
