"""Pins of the mixed scalar forms of the oracle (or_assemble_mixed: mixed
mass, mixed mass power, mixed grad-grad, feSysElm.cpp:283-403, :886-976).

- With the test and trial maps of one field they are the mass, mass-power
  and diffusion forms (or_assemble_mass_power, or_assemble_diffusion).
- With two fields, the matrix is the exact Jacobian of the residual with
  respect to the trial field (central differences; the test field's values
  do not enter), so a swapped test / trial role or a wrong power fails.
- Newton on the oracle reproduces the reference's golden table
  scalarFE_mixedDiffusion.output (u + v^3 - lap v = f, u + v - lap u = g, P1
  to P4 on mmsMeshes/square1-4.msh, degree-8 quadrature, both fields).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MIXED = {  # scalarFE_mixedDiffusion.output: (degree, field) -> errors
    (1, 0): [8.367722e-02, 2.264353e-02, 5.782545e-03, 1.453525e-03],
    (1, 1): [7.400104e-02, 2.021458e-02, 5.179422e-03, 1.303132e-03],
    (2, 0): [4.400867e-03, 5.508090e-04, 6.882950e-05, 8.603396e-06],
    (2, 1): [4.444817e-03, 5.523949e-04, 6.887708e-05, 8.604779e-06],
    (3, 0): [3.367211e-04, 2.000302e-05, 1.215996e-06, 7.501904e-08],
    (3, 1): [3.420727e-04, 2.032149e-05, 1.230033e-06, 7.561815e-08],
    (4, 0): [2.057034e-05, 6.589477e-07, 2.073064e-08, 6.485224e-10],
    (4, 1): [2.065756e-05, 6.598122e-07, 2.074137e-08, 6.486780e-10],
}
KIND_OF = {1: 1, 2: 2, 3: 5, 4: 6}
NF = {1: 3, 2: 6, 5: 10, 6: 15}


def u_exact(x, y):
    return np.sin(np.pi * x) * np.sin(np.pi * y)


def v_exact(x, y):
    return np.cos(np.pi * x) * np.cos(np.pi * y)


def u_source(x, y):
    u, v = u_exact(x, y), v_exact(x, y)
    return -(u + v ** 3 - 2 * np.pi ** 2 * v)


def v_source(x, y):
    u, v = u_exact(x, y), v_exact(x, y)
    return -(u + v - 2 * np.pi ** 2 * u)


def matches(err, ref):
    return all(abs(e - r) <= max(1e-6 * r, 5e-15) for e, r in zip(err, ref))


def oracle_mixed(oracle, mesh, kind, tol=1e-13):
    """The reference's form stack (scalarFE.cpp:455-475) assembled by the
    oracle, Newton with direct solves."""
    from feng_amd.api import Numbering, essential_values
    from tests.test_oracle_time import _qp_table
    num = Numbering(mesh, [kind, kind], [1, 1])
    n, nf = num.nInc, NF[kind]
    au, av = num.adr
    ia, ja = oracle.csr(n, [(au, nf, au, nf), (au, nf, av, nf), (av, nf, av, nf), (av, nf, au, nf)])
    u = essential_values(num, [lambda x, y, z, c: u_exact(x, y), lambda x, y, z, c: v_exact(x, y)])
    su = _qp_table(mesh, 8, u_source)
    sv = _qp_table(mesh, 8, v_source)
    X, T = mesh.xyz, mesh.conn
    for it in range(20):
        val, rhs = oracle.assemble_diffusion(kind, X, T, au, n, ia, ja, 8, (0, 0.0, 0.0, None), (5, 0.0, 0.0, su), u)
        v2, r2 = oracle.assemble_diffusion(kind, X, T, av, n, ia, ja, 8, (0, 0.0, 0.0, None), (5, 0.0, 0.0, sv), u)
        val += v2
        rhs += r2
        oracle.assemble_mass_power(kind, X, T, au, n, ia, ja, 8, (0, 1.0, 0.0, None), 1.0, u, val, rhs)
        oracle.assemble_mixed(kind, X, T, au, av, n, ia, ja, 8, 0.0, 1.0, 3.0, -1.0, u, val, rhs)
        oracle.assemble_mass_power(kind, X, T, av, n, ia, ja, 8, (0, 1.0, 0.0, None), 1.0, u, val, rhs)
        oracle.assemble_mixed(kind, X, T, av, au, n, ia, ja, 8, 1.0, 0.0, 1.0, -1.0, u, val, rhs)
        du = spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
        u[:n] += du
        if np.abs(du).max() <= tol:
            break
    return num, u


def test_same_field_reduces_to_mass_power_and_diffusion(oracle):
    from feng_amd.api import Mesh, Numbering
    m = Mesh.square(5, 0.2, 7)
    for kind in (1, 2, 5):
        num = Numbering(m, [kind], [1])
        n, nf, a = num.nInc, NF[kind], num.adr[0]
        ia, ja = oracle.csr(n, [(a, nf, a, nf)])
        u = np.random.default_rng(kind).normal(size=num.nDofs)
        v1, r1 = np.zeros(ja.size), np.zeros(n)
        oracle.assemble_mixed(kind, m.xyz, m.conn, a, a, n, ia, ja, 6, 0.7, 1.3, 3.0, 2.5, u, v1, r1)
        v0, r0 = oracle.assemble_diffusion(kind, m.xyz, m.conn, a, n, ia, ja, 6, (0, 2.5, 0.0, None),
                                           (0, 0.0, 0.0, None), u)
        oracle.assemble_mass_power(kind, m.xyz, m.conn, a, n, ia, ja, 6, (0, 0.7, 0.0, None), 1.0, u, v0, r0)
        oracle.assemble_mass_power(kind, m.xyz, m.conn, a, n, ia, ja, 6, (0, 1.3, 0.0, None), 3.0, u, v0, r0)
        assert np.abs(v1 - v0).max() <= 1e-12 * np.abs(v0).max()
        assert np.abs(r1 - r0).max() <= 1e-12 * np.abs(r0).max()


def test_matrix_is_the_jacobian_in_the_trial_field(oracle):
    from feng_amd.api import Mesh, Numbering
    m = Mesh.square(4, 0.2, 5)
    kind = 2
    num = Numbering(m, [kind, kind], [1, 1])
    n, nf = num.nInc, NF[kind]
    au, av = num.adr
    ia, ja = oracle.csr(n, [(au, nf, av, nf)] + [(av, nf, av, nf)])
    u = np.random.default_rng(3).normal(size=num.nDofs)
    val, r0 = np.zeros(ja.size), np.zeros(n)
    oracle.assemble_mixed(kind, m.xyz, m.conn, au, av, n, ia, ja, 6, 0.4, 1.0, 3.0, -1.0, u, val, r0)
    A = sp.csr_matrix((val, ja, ia), shape=(n, n)).toarray()
    rows_u = np.unique(au[au < n])
    h = 1e-6
    for j in np.unique(av[av < n])[::3]:
        up, um = u.copy(), u.copy()
        up[j] += h
        um[j] -= h
        rp, rm = np.zeros(n), np.zeros(n)
        oracle.assemble_mixed(kind, m.xyz, m.conn, au, av, n, ia, ja, 6, 0.4, 1.0, 3.0, -1.0, up, None, rp)
        oracle.assemble_mixed(kind, m.xyz, m.conn, au, av, n, ia, ja, 6, 0.4, 1.0, 3.0, -1.0, um, None, rm)
        col = -(rp - rm) / (2 * h)
        assert np.abs(col - A[:, j]).max() <= 1e-6 * max(1.0, np.abs(A[:, j]).max()), j
    # rows are the test field's, columns the trial field's; the test values do not enter
    assert np.all(np.isin(np.flatnonzero(np.abs(A).sum(axis=1)), rows_u))
    u2 = u.copy()
    u2[rows_u] += 1.0
    r2 = np.zeros(n)
    oracle.assemble_mixed(kind, m.xyz, m.conn, au, av, n, ia, ja, 6, 0.4, 1.0, 3.0, -1.0, u2, None, r2)
    assert np.array_equal(r2, r0)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_oracle_reproduces_reference_mixed_diffusion_table(oracle, order):
    from feng_amd.api import Mesh
    from tests.conftest import GOLDEN
    from tests.test_oracle_time import _qp_table
    kind = KIND_OF[order]
    err = {0: [], 1: []}
    for lvl in range(1, 5):
        m = Mesh.read_msh(GOLDEN / "meshes" / f"square{lvl}.msh")
        num, u = oracle_mixed(oracle, m, kind)
        for f, ex in ((0, u_exact), (1, v_exact)):
            err[f].append(oracle.l2_error_qp(kind, m.xyz, m.conn, num.adr[f], u, 8, _qp_table(m, 8, ex)))
    assert matches(err[0], MIXED[(order, 0)]), err[0]
    assert matches(err[1], MIXED[(order, 1)]), err[1]
