"""ADVECTION on the device: the reference's golden tables scalarFE_advdiff.output
and scalarFE_advdiffreact.output (P1, P2 on mmsMeshes/square1-4.msh) through
the fused scalar sweep, Newton with GMRES/ILU(0) (the matrix is not
symmetric) and the device L2 norm; assembled system against the oracle."""
import numpy as np
import pytest

from tests.test_oracle_advection import ADVDIFF, ALPHA, CX, CY, K, matches, oracle_advdiff, source, u_exact

pytestmark = pytest.mark.gpu


def _system(m, kind, react, deg=8):
    from feng_amd.api import FN_CONST, FN_QP, Fn
    from feng_amd.problems import advection_diffusion_system
    from tests.test_oracle_time import _qp_table
    src = Fn(FN_QP, data=_qp_table(m, deg, lambda x, y: source(x, y, react)))
    return advection_diffusion_system(m, kind, deg, [Fn(FN_CONST, CX), Fn(FN_CONST, CY)], Fn(FN_CONST, K), src,
                                      Fn(FN_CONST, ALPHA) if react else None,
                                      lambda x, y, z, c: u_exact(x, y))


@pytest.mark.parametrize("react", [False, True])
@pytest.mark.parametrize("kind", [1, 2])
def test_reference_advdiff_table(oracle, kind, react):
    from feng_amd.api import FN_QP, KSP_GMRES, NORM_L2_ERROR, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, solve_newton
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    print("advdiff", kind, react, flush=True)
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        S, num = _system(m, kind, react)
        rep = solve_newton(S, NLSolverOptions(1e-12, 1e-12, 1e4, 10), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13,
                           restart=100, max_iter=20000)
        assert rep.converged
        err.append(S.norm(NORM_L2_ERROR, 0, 8, [Fn(FN_QP, data=_qp_table(m, 8, lambda x, y: u_exact(x, y)))]))
        if lvl == 2:
            _, uo = oracle_advdiff(oracle, m, kind, react)
            u = S.get_solution()
            assert np.abs(u - uo).max() <= 1e-10 * np.abs(uo).max()
        S.close()
    assert matches(err, ADVDIFF[(react, kind)]), err


def test_advection_assembly_matches_oracle(oracle):
    import scipy.sparse as sp
    from feng_amd.api import Mesh
    from tests.test_oracle_time import _qp_table
    m = Mesh.square(9, 0.2, 6)
    for kind in (1, 2):
        S, num = _system(m, kind, True, deg=5)
        rng = np.random.default_rng(kind)
        u = S.get_solution()
        u[:num.nInc] = rng.normal(size=num.nInc)
        S.set_solution(u)
        for lanes in (0, 1):
            S.set_tuning("scalar_lanes", lanes)
            S.setToZero()
            S.assemble(3)
            val, rhs = S.get_system()
            ia, ja = S.get_csr()
            v0, r0 = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, 5,
                                               (0, K, 0.0, None),
                                               (5, 0.0, 0.0, _qp_table(m, 5, lambda x, y: source(x, y, True))), u)
            oracle.assemble_advection(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, 5,
                                      [(0, CX, 0.0, None), (0, CY, 0.0, None)], u, v0, r0)
            oracle.assemble_mass_power(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, 5, (0, ALPHA, 0.0, None),
                                       1.0, u, v0, r0)
            assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0), (kind, lanes)
            assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0), (kind, lanes)
        S.close()
