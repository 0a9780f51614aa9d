"""Pins of the advection oracle (or_assemble_advection, or_l2_error_qp).

- Be = -Ae u for the linear advection form (matrix and residual agree), and
  a constant field gives zero advective residual on interior rows (the
  conservative form integrates c . grad(phi_i) of an interior function to 0).
- The oracle reproduces the reference's golden tables scalarFE_advdiff.output
  and scalarFE_advdiffreact.output (P1 to P4 on mmsMeshes/square1-4.msh,
  degree-8 quadrature, u = sin(xy), velocity (1, 2), k = 1, reaction 7).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ADVDIFF = {  # scalarFE_advdiff.output / scalarFE_advdiffreact.output, P1 and P2 rows
    (False, 1): [6.627960e-03, 1.642586e-03, 4.096968e-04, 1.023639e-04],
    (False, 2): [8.535503e-05, 1.069797e-05, 1.337817e-06, 1.672420e-07],
    (True, 1): [6.049672e-03, 1.478301e-03, 3.672435e-04, 9.166052e-05],
    (True, 2): [8.511754e-05, 1.068925e-05, 1.337531e-06, 1.672330e-07],
    # P3 (kind 5) and P4 (kind 6) rows
    (False, 5): [3.656498e-06, 2.270226e-07, 1.406515e-08, 8.740671e-10],
    (False, 6): [1.850028e-07, 5.817233e-09, 1.814484e-10, 5.656129e-12],
    (True, 5): [3.644954e-06, 2.268197e-07, 1.405997e-08, 8.738565e-10],
    (True, 6): [1.846470e-07, 5.814960e-09, 1.814328e-10, 5.656016e-12],
}
CX, CY, K, ALPHA = 1.0, 2.0, 1.0, 7.0


def u_exact(x, y):
    return np.sin(x * y)


def source(x, y, react):
    s = (CX * y + CY * x) * np.cos(x * y) + K * (x * x + y * y) * np.sin(x * y)
    if react:
        s = s + ALPHA * np.sin(x * y)
    return -s


def matches(err, ref):
    return all(abs(e - r) <= 1e-6 * r for e, r in zip(err, ref))


def oracle_advdiff(oracle, mesh, kind, react):
    from feng_amd.api import Numbering, essential_values
    from tests.test_oracle_time import _qp_table
    num = Numbering(mesh, [kind], [1])
    n = num.nInc
    nf = {1: 3, 2: 6, 5: 10, 6: 15}[kind]
    ia, ja = oracle.csr(n, [(num.adr[0], nf, num.adr[0], nf)])
    u = essential_values(num, [lambda x, y, z, c: u_exact(x, y)])
    src = _qp_table(mesh, 8, lambda x, y: source(x, y, react))
    for _ in range(2):
        val, rhs = oracle.assemble_diffusion(kind, mesh.xyz, mesh.conn, num.adr[0], n, ia, ja, 8,
                                             (0, K, 0.0, None), (5, 0.0, 0.0, src), u)
        oracle.assemble_advection(kind, mesh.xyz, mesh.conn, num.adr[0], n, ia, ja, 8,
                                  [(0, CX, 0.0, None), (0, CY, 0.0, None)], u, val, rhs)
        if react:
            oracle.assemble_mass_power(kind, mesh.xyz, mesh.conn, num.adr[0], n, ia, ja, 8, (0, ALPHA, 0.0, None),
                                       1.0, u, val, rhs)
        u[:n] += spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
    return num, u


@pytest.mark.parametrize("react", [False, True])
@pytest.mark.parametrize("kind", [1, 2, 5, 6])
def test_oracle_reproduces_reference_advdiff_tables(oracle, kind, react):
    from feng_amd.api import Mesh
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        num, u = oracle_advdiff(oracle, m, kind, react)
        err.append(oracle.l2_error_qp(kind, m.xyz, m.conn, num.adr[0], u, 8,
                                      _qp_table(m, 8, lambda x, y: u_exact(x, y))))
    assert matches(err, ADVDIFF[(react, kind)]), err


def test_advection_matrix_and_residual_agree(oracle):
    from feng_amd.api import Mesh, Numbering
    m = Mesh.square(6, 0.2, 3)
    num = Numbering(m, [2], [0])
    n = num.nInc
    ia, ja = oracle.csr(n, [(num.adr[0], 6, num.adr[0], 6)])
    rng = np.random.default_rng(0)
    u = rng.normal(size=n)
    val, rhs = np.zeros(ja.size), np.zeros(n)
    oracle.assemble_advection(2, m.xyz, m.conn, num.adr[0], n, ia, ja, 5, [(0, CX, 0.0, None), (0, CY, 0.0, None)],
                              u, val, rhs)
    A = sp.csr_matrix((val, ja, ia), shape=(n, n))
    assert np.abs(rhs + A @ u).max() <= 1e-13 * np.abs(rhs).max()
    # a constant field: zero advective residual on rows whose function vanishes on the boundary
    num_b = Numbering(m, [2], [1])
    ib, jb = oracle.csr(num_b.nInc, [(num_b.adr[0], 6, num_b.adr[0], 6)])
    rb = np.zeros(num_b.nInc)
    oracle.assemble_advection(2, m.xyz, m.conn, num_b.adr[0], num_b.nInc, ib, jb, 5,
                              [(0, CX, 0.0, None), (0, CY, 0.0, None)], np.ones(num_b.nDofs), None, rb)
    assert np.abs(rb).max() <= 1e-14
