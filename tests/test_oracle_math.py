"""Closed-form pins of the oracle where no reference vector exists:
quadrature exactness, textbook P1 stiffness, solver identities."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp

from tests.oracle_bind import P


@pytest.mark.parametrize("deg", range(1, 13))
def test_triangle_rule_integrates_monomials_exactly(oracle, deg):
    # int_T r^a s^b = a! b! / (a+b+2)!  on the reference triangle
    r, s, w = (np.zeros(36) for _ in range(3))
    nq = oracle.L.or_quad_tri(deg, P(r, C.c_double), P(s, C.c_double), P(w, C.c_double))
    assert nq > 0
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            approx = float(np.sum(w[:nq] * r[:nq] ** a * s[:nq] ** b))
            assert abs(approx - exact) <= 1e-13


def test_tet_rules_truncated_weights(oracle):
    r, s, t, w = (np.zeros(8) for _ in range(4))
    nq = oracle.L.or_quad_tet(2, *[P(x, C.c_double) for x in (r, s, t, w)])
    assert nq == 4
    # the reference's 15-digit literals: sum of weights is 1/6 to ~1e-15
    assert abs(w[:nq].sum() - 1.0 / 6.0) < 1e-14
    # degree-2 exactness at the precision of the 12-digit points
    assert abs(np.sum(w[:nq] * r[:nq] ** 2) - 2.0 / 120.0) < 1e-12


def test_p1_element_is_textbook_stiffness(oracle):
    rng = np.random.default_rng(0)
    for _ in range(5):
        xy = rng.random((3, 2))
        if (xy[1, 0] - xy[0, 0]) * (xy[2, 1] - xy[0, 1]) - (xy[1, 1] - xy[0, 1]) * (xy[2, 0] - xy[0, 0]) < 0:
            xy = xy[[0, 2, 1]]
        coords = np.zeros(9)
        coords.reshape(3, 3)[:, :2] = xy
        from tests.oracle_bind import OrFn
        Ae = np.zeros(9)
        oracle.L.or_element_diffusion(1, P(coords, C.c_double), 2, OrFn(0, 2.5, 0.0, None), 0, P(Ae, C.c_double))
        # K_ij = k |T| grad(l_i) . grad(l_j), barycentric gradients from the inverse of [1 x y]
        M = np.hstack([np.ones((3, 1)), xy])
        G = np.linalg.inv(M)[1:, :]  # 2 x 3
        area = 0.5 * abs(np.linalg.det(M))
        K = 2.5 * area * G.T @ G
        assert np.abs(Ae.reshape(3, 3) - K).max() <= 1e-13 * np.abs(K).max()


def _poisson1d(n):
    A = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64), A


def test_cg_matches_dense_solve(oracle):
    ia, ja, val, A = _poisson1d(50)
    b = np.sin(np.arange(50.0))
    x, it, rr = oracle.cg_jacobi(ia, ja, val, b, 1e-12, 1000)
    xd = np.linalg.solve(A.toarray(), b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    assert rr <= 1e-12


def test_ic0_is_exact_cholesky_without_fill(oracle):
    ia, ja, val, A = _poisson1d(30)
    l, rc = oracle.ic0(ia, ja, val)
    assert rc == 0
    L = sp.csr_matrix((l, ja, ia), shape=A.shape).toarray()
    assert np.abs(np.triu(L, 1)).max() == 0.0
    assert np.abs(L @ L.T - A.toarray()).max() <= 1e-12
    b = np.cos(np.arange(30.0))
    x, it, rr = oracle.pcg_ic0(ia, ja, val, l, b, 1e-12, 100)
    assert it == 1 and rr <= 1e-12


def test_ilu0_exact_on_tridiagonal_and_gmres(oracle):
    n = 40
    A = sp.diags([-np.ones(n - 1) * 1.3, 3.0 * np.ones(n), -np.ones(n - 1) * 0.7], [-1, 0, 1], format="csr")
    A.sort_indices()
    ia, ja, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    lu, rc = oracle.ilu0(ia, ja, val)
    assert rc == 0
    LU = sp.csr_matrix((lu, ja, ia), shape=A.shape).toarray()
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.abs(L @ U - A.toarray()).max() <= 1e-12
    b = np.arange(1.0, n + 1)
    x, it, rr = oracle.gmres_ilu0(ia, ja, val, lu, b, 30, 1e-12, 100)
    assert it <= 2 and rr <= 1e-12
    assert np.abs(A @ x - b).max() <= 1e-10


def test_gmres_nonsymmetric_convergence(oracle):
    rng = np.random.default_rng(1)
    n = 200
    A = sp.random(n, n, density=0.03, random_state=1, format="csr") + sp.eye(n) * 4.0
    A = A.tocsr()
    A.sort_indices()
    ia, ja, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    lu, rc = oracle.ilu0(ia, ja, val)
    b = rng.random(n)
    x, it, rr = oracle.gmres_ilu0(ia, ja, val, lu, b, 10, 1e-10, 500)
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)


def _ilu0_textbook(A):
    """ILU(0) by the IKJ elimination restricted to the pattern (Saad,
    Iterative Methods for Sparse Linear Systems, alg. 10.4), dense storage."""
    n = A.shape[0]
    M = A.toarray().astype(float)
    P = A.toarray() != 0
    for i in range(1, n):
        for k in range(i):
            if not P[i, k]:
                continue
            M[i, k] /= M[k, k]
            for j in range(k + 1, n):
                if P[i, j]:
                    M[i, j] -= M[i, k] * M[k, j]
    return M


def _fe_matrices():
    from tests.test_oracle_mpnn import grid_matrix
    rng = np.random.default_rng(5)
    ia, ja, val = grid_matrix(9, 2)
    A2 = sp.csr_matrix((val, ja, ia), shape=(len(ia) - 1,) * 2)
    # a nonsymmetric matrix on the same pattern (an advection-like skew part)
    S = A2.copy()
    S.data = rng.normal(size=S.nnz) * 0.2 * np.abs(A2.data).max()
    N2 = (A2 + (S - S.T)).tocsr()
    from feng_amd.api import Mesh, Numbering, TET_P1
    m = Mesh.cube(3)
    num = Numbering(m, [TET_P1], [0])
    rows, cols, vals = [], [], []
    conn = m.conn.reshape(-1, 4)
    for e in range(conn.shape[0]):
        d = num.adr[0][4 * e:4 * e + 4]
        K = rng.normal(size=(4, 4))
        K = K @ K.T + 4 * np.eye(4)
        for a in range(4):
            for b in range(4):
                rows.append(d[a]); cols.append(d[b]); vals.append(K[a, b])
    A3 = sp.csr_matrix((vals, (rows, cols)), shape=(num.nDofs,) * 2)
    out = []
    for A in (A2, N2, A3):
        A = A.tocsr()
        A.sum_duplicates()
        A.sort_indices()
        out.append(A)
    return out


def test_ilu0_factor_reproduces_the_matrix_on_its_pattern(oracle):
    """The defining property of ILU(0): (L U)_ij = a_ij on the pattern of A
    (L unit lower, U upper, both on that pattern), with fill outside it, on a
    2D FE matrix, a nonsymmetric one on the same pattern and a 3D tet matrix;
    and the factor equals an independent textbook ILU(0) (IKJ elimination
    restricted to the pattern)."""
    for A in _fe_matrices():
        n = A.shape[0]
        ia, ja, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
        lu, rc = oracle.ilu0(ia, ja, val)
        assert rc == 0
        LU = sp.csr_matrix((lu, ja, ia), shape=A.shape).toarray()
        L = np.tril(LU, -1) + np.eye(n)
        U = np.triu(LU)
        prod = L @ U
        P = A.toarray() != 0
        scale = np.abs(A.toarray()).max()
        assert np.abs(prod[P] - A.toarray()[P]).max() <= 1e-12 * scale
        assert np.abs(prod[~P]).max() > 1e-6 * scale  # ILU(0) is not the exact LU here
        ref = _ilu0_textbook(A)
        assert np.abs(LU[P] - ref[P]).max() <= 1e-12 * np.abs(ref[P]).max()


def test_fp64_residual_floor_of_the_high_contrast_p2_problem(oracle):
    """Why the mixed-precision workload (P2, k log-uniform in [1e-2, 1e2] on
    16 x 16 blocks) is solved to rel-res 1e-8 and not 1e-10: the FP64 solution
    of a sparse direct LU, its residual evaluated in extended precision, has a
    relative residual that grows about 4x per refinement (the floor scales as
    || |A| |x| || / ||b||, ~h^-2): 3e-12, 1.4e-11, 7e-11 at 32^2, 64^2, 128^2
    -- extrapolated, ~1e-8 at 1024^2."""
    import scipy.sparse.linalg as spla
    from feng_amd.api import Mesh, Numbering, TRI_P2
    rel = []
    for N in (32, 64, 128):
        m = Mesh.square(N)
        k = m.coeff_blocks(16, 1e-2, 1e2, 3)
        num = Numbering(m, [TRI_P2], [1])
        n = num.nInc
        ia, ja = oracle.csr(n, [(num.adr[0], 6, num.adr[0], 6)])
        val, rhs = oracle.assemble_diffusion(TRI_P2, m.xyz, m.conn, num.adr[0], n, ia, ja, 2, (4, 0.0, 0.0, k),
                                             (0, -1.0, 0.0, None), np.zeros(num.nDofs))
        A = sp.csr_matrix((val, ja, ia), shape=(n, n))
        x = spla.spsolve(A.tocsc(), rhs)
        Al = sp.csr_matrix((val.astype(np.longdouble), ja, ia), shape=(n, n))
        r = rhs.astype(np.longdouble) - Al @ x.astype(np.longdouble)
        rel.append(float(np.sqrt((r * r).sum()) / np.sqrt((rhs.astype(np.longdouble) ** 2).sum())))
    print("FP64 direct-solve relative residuals at 32^2, 64^2, 128^2:", rel)
    assert rel[1] > 3 * rel[0] and rel[2] > 3 * rel[1]
    assert rel[2] > 3e-11
