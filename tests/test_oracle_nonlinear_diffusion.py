"""Pins of the nonlinear diffusion oracle (or_assemble_nonlinear_diffusion,
feSysElm.cpp:681-760).

- A constant conductivity reduces it to the linear diffusion form
  (matrix and residual equal or_assemble_diffusion's).
- The matrix is the exact Jacobian of the residual: Ae = -dBe/du, checked by
  central differences for k(u) = 1 + u^2 (a transposed phi_i / phi_j or a
  dropped dk term fails it).
- Newton on the oracle reproduces the reference's golden table
  scalarFE_nonlineardiffusion.output (P1 and P2 rows, mmsMeshes/square1-4.msh,
  degree-8 quadrature, u = sin(pi x) sin(pi y), k(u) = 1 + u^2).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

NLDIFF = {  # scalarFE_nonlineardiffusion.output, P1 and P2 rows
    1: [6.911232e-02, 1.827478e-02, 4.643887e-03, 1.165997e-03],
    2: [4.304667e-03, 5.475869e-04, 6.872638e-05, 8.600158e-06],
}
K_POLY, DK_POLY = [1.0, 0.0, 1.0], [0.0, 2.0]


def u_exact(x, y):
    return np.sin(np.pi * x) * np.sin(np.pi * y)


def source(x, y):
    """div((1 + u^2) grad u) of the exact solution (the reference's source)."""
    cx, cy, sx, sy = np.cos(np.pi * x), np.cos(np.pi * y), np.sin(np.pi * x), np.sin(np.pi * y)
    return 2 * np.pi ** 2 * (cx * cx * sx * sy ** 3 + cy * cy * sx ** 3 * sy - sx * sy * (sx * sx * sy * sy + 1))


def matches(err, ref):
    return all(abs(e - r) <= 1e-6 * r for e, r in zip(err, ref))


def oracle_nldiff(oracle, mesh, kind, tol=1e-12, max_newton=20):
    from feng_amd.api import Numbering, essential_values
    from tests.test_oracle_time import _qp_table
    num = Numbering(mesh, [kind], [1])
    n = num.nInc
    nf = {1: 3, 2: 6}[kind]
    ia, ja = oracle.csr(n, [(num.adr[0], nf, num.adr[0], nf)])
    u = essential_values(num, [lambda x, y, z, c: u_exact(x, y)])
    src = _qp_table(mesh, 8, lambda x, y: source(x, y))
    for it in range(max_newton):
        val, rhs = oracle.assemble_diffusion(kind, mesh.xyz, mesh.conn, num.adr[0], n, ia, ja, 8,
                                             (0, 0.0, 0.0, None), (5, 0.0, 0.0, src), u)
        oracle.assemble_nonlinear_diffusion(kind, mesh.xyz, mesh.conn, num.adr[0], n, ia, ja, 8, K_POLY, DK_POLY,
                                            u, val, rhs)
        du = spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
        u[:n] += du
        if np.abs(du).max() <= tol:
            break
    return num, u, it + 1


def _square_case(kind):
    from feng_amd.api import Mesh, Numbering
    m = Mesh.square(5, 0.2, 11)
    num = Numbering(m, [kind], [1])
    nf = {1: 3, 2: 6}[kind]
    return m, num, nf


@pytest.mark.parametrize("kind", [1, 2])
def test_constant_conductivity_is_linear_diffusion(oracle, kind):
    m, num, nf = _square_case(kind)
    n = num.nInc
    ia, ja = oracle.csr(n, [(num.adr[0], nf, num.adr[0], nf)])
    u = np.random.default_rng(1).normal(size=num.nDofs)
    v1, r1 = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, 6, (0, 2.5, 0.0, None),
                                       (0, 0.0, 0.0, None), u)
    v2, r2 = np.zeros(ja.size), np.zeros(n)
    oracle.assemble_nonlinear_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, 6, [2.5], [0.0], u, v2, r2)
    assert np.abs(v1 - v2).max() <= 1e-12 * np.abs(v1).max()
    assert np.abs(r1 - r2).max() <= 1e-12 * np.abs(r1).max()


@pytest.mark.parametrize("kind", [1, 2])
def test_matrix_is_the_residual_jacobian(oracle, kind):
    m, num, nf = _square_case(kind)
    n = num.nInc
    ia, ja = oracle.csr(n, [(num.adr[0], nf, num.adr[0], nf)])
    u = np.random.default_rng(2).normal(size=num.nDofs)
    val, r0 = np.zeros(ja.size), np.zeros(n)
    oracle.assemble_nonlinear_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, 6, K_POLY, DK_POLY, u, val, r0)
    A = sp.csr_matrix((val, ja, ia), shape=(n, n)).toarray()
    h = 1e-6
    for j in range(0, n, max(1, n // 7)):
        up, um = u.copy(), u.copy()
        up[j] += h
        um[j] -= h
        rp, rm = np.zeros(n), np.zeros(n)
        oracle.assemble_nonlinear_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, 6, K_POLY, DK_POLY, up,
                                            None, rp)
        oracle.assemble_nonlinear_diffusion(kind, m.xyz, m.conn, num.adr[0], n, ia, ja, 6, K_POLY, DK_POLY, um,
                                            None, rm)
        col = -(rp - rm) / (2 * h)
        assert np.abs(col - A[:, j]).max() <= 1e-6 * np.abs(A[:, j]).max(), j
    # the dk term makes the Jacobian non-symmetric
    assert np.abs(A - A.T).max() > 1e-3


@pytest.mark.parametrize("kind", [1, 2])
def test_oracle_reproduces_reference_nonlinear_diffusion_table(oracle, kind):
    from feng_amd.api import Mesh
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    err = []
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        num, u, its = oracle_nldiff(oracle, m, kind)
        assert its < 20
        err.append(oracle.l2_error_qp(kind, m.xyz, m.conn, num.adr[0], u, 8,
                                      _qp_table(m, 8, lambda x, y: u_exact(x, y))))
    assert matches(err, NLDIFF[kind]), err
