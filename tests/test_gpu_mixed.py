"""Mixed scalar forms on the device (MIXED_MASS, MIXED_MASS_POWER,
MIXED_GRADGRAD, feSysElm.cpp:283-403, :886-976): the reference's golden
table scalarFE_mixedDiffusion.output (P1 to P4 on mmsMeshes/square1-4.msh,
both fields) through the device assembly and the device norm; the assembled
coupled system against the oracle on P1, P2, P3 triangles and P1 tets.

The coupled system's diagonal blocks are mass matrices (O(h^2)) against O(1)
grad-grad couplings: ILU(0) pivots on them and the device GMRES/ILU(0)
converges only on the coarsest mesh with P1 / P2, where it is checked; the table's
Newton steps solve the device-assembled systems with a host direct solver,
in place of the reference's direct solver, which has no device counterpart
here."""
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


def _host_newton(S, n, tol=1e-13):
    """Newton on the device-assembled residual and Jacobian, each step solved
    by a host direct solver."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    ia, ja = S.get_csr()
    for it in range(20):
        S.setToZero()
        S.assemble(3)
        val, rhs = S.get_system()
        du = spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
        u = S.get_solution()
        u[:n] += du
        S.set_solution(u)
        if np.abs(du).max() <= tol:
            return it + 1
    raise AssertionError("Newton did not converge")


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
        assert _host_newton(S, num.nInc) > 2  # a nonlinear problem
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
