"""Pins of the learned-preconditioner oracle (oracle/fe_oracle_mpnn.py).

The reference has no learned preconditioner, so nothing of the reference pins
these functions; they are pinned by mathematics instead:
  - the hand-written gradient against central finite differences of the
    forward, on every parameter group;
  - the zero-head net reproduces the IC(0) split factor G = I - L_strict,
    D = u_ii bit for bit (ILU(0) from the C oracle, feLinearSystemPETSc's
    -pc_type icc counterpart);
  - an exact inverse factor (full pattern, G = L^-1) gives zero loss, the
    truncated one does not;
  - G^T D^-1 G is SPD for arbitrary weights (so PCG with it cannot break down).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fe_oracle_mpnn as mp
from tests.oracle_bind import load_oracle


def grid_matrix(N=7, seed=3):
    """SPD 9-point-pattern matrix on an N x N grid: weighted graph Laplacian
    with random positive edge weights plus a positive diagonal shift."""
    rng = np.random.default_rng(seed)
    idx = np.arange(N * N).reshape(N, N)
    rows, cols, w = [], [], []
    for di, dj in ((0, 1), (1, 0), (1, 1), (1, -1)):
        for i in range(N):
            for j in range(N):
                i2, j2 = i + di, j + dj
                if 0 <= i2 < N and 0 <= j2 < N:
                    rows.append(idx[i, j]); cols.append(idx[i2, j2]); w.append(rng.uniform(0.2, 3.0))
    rows, cols, w = np.array(rows), np.array(cols), np.array(w)
    n = N * N
    L = sp.coo_matrix((-w, (rows, cols)), shape=(n, n))
    L = L + L.T
    deg = -np.asarray(L.sum(axis=1)).ravel()
    A = (L + sp.diags(deg + rng.uniform(0.05, 0.5, n))).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


@pytest.fixture(scope="module")
def graph():
    ia, ja, val = grid_matrix()
    lu, rc = load_oracle().ilu0(ia, ja, val)
    assert rc == 0
    return mp.Graph(ia, ja, val, lu)


def test_param_layout():
    assert mp.param_count() == 16 * 32 + 32 + 3 * (64 * 32 + 32) + 2 * 32 * 8 + 32 + 1
    th = mp.init_params(5)
    assert np.allclose(mp.pack(mp.unpack(th)), th)
    assert np.all(mp.unpack(th)["b1"] == 0) and np.all(mp.unpack(th)["wt"] == 0)


def test_gradient_matches_finite_differences(graph):
    rng = np.random.default_rng(0)
    th = mp.init_params(7)
    # larger heads so that every term of the loss carries weight
    d = mp.unpack(th)
    d["P"] *= 20.0
    d["Q"] *= 20.0
    d["wt"] = rng.normal(0, 0.3, d["wt"].shape)
    th = mp.pack(d)
    kw = dict(K=4, seed=11, nu=2, round_bf16=False)
    L0, g, _ = mp.loss_and_grad(graph, th, **kw)
    assert L0 > 0
    o = 0
    checked = 0
    for name, shape in mp.LAYOUT:
        c = int(np.prod(shape))
        for q in rng.choice(c, size=min(c, 6), replace=False):
            k = o + q
            h = 1e-6 * max(1.0, abs(th[k]))
            tp = th.copy(); tp[k] += h
            tm = th.copy(); tm[k] -= h
            fd = (mp.loss_and_grad(graph, tp, want_grad=False, **kw)[0] -
                  mp.loss_and_grad(graph, tm, want_grad=False, **kw)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * np.abs(g).max(), (name, q, fd, g[k])
            checked += 1
        o += c
    assert checked >= 60


def test_zero_heads_reproduce_ic0_split(graph):
    th = mp.init_params(2)
    d = mp.unpack(th)
    d["P"][:] = 0.0
    th = mp.pack(d)
    f = mp.forward(graph, th)
    assert np.array_equal(f["g"], graph.g0) and np.array_equal(f["d"], graph.d0)
    G, dinv = mp.preconditioner(graph, f)
    # G = I - L_strict and D = u_ii in the original coordinates
    ia = np.asarray(graph.Ahat.indptr)
    n = graph.n
    lu_rows = graph.rows
    L = sp.csr_matrix((np.where(graph.lower, graph.g0 * graph.s[graph.cols] / graph.s[lu_rows], 0.0),
                       graph.cols, ia), shape=(n, n))
    assert abs(G - (L + sp.identity(n))).max() < 1e-14
    assert np.allclose(dinv * graph.a_ii * graph.d0, 1.0, rtol=1e-15, atol=0)


def test_exact_inverse_factor_gives_zero_loss():
    rng = np.random.default_rng(4)
    n = 9
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    S = sp.csr_matrix(A)
    S.sort_indices()
    ia, ja, val = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()
    lu, rc = load_oracle().ilu0(ia, ja, val)
    g = mp.Graph(ia, ja, val, lu)
    xi = mp.smoothed_probes(g, 3, 1, 0, 0.0)
    # full pattern: ILU(0) is the exact LDL^T of A^; G = L^-1 makes A~ = I
    Ah = g.Ahat.toarray()
    Lc = np.linalg.cholesky(Ah)
    dd = np.diag(Lc) ** 2
    Lunit = Lc / np.diag(Lc)[None, :]
    Ginv = np.linalg.inv(Lunit)
    gvals = np.where(g.lower, Ginv[g.rows, g.cols], 0.0)
    assert mp.factor_loss(g, gvals, dd, xi) < 1e-24
    assert np.allclose(g.d0, dd, rtol=1e-12)
    assert mp.factor_loss(g, g.g0, g.d0, xi) > 1e-4


def test_corrected_preconditioner_is_spd(graph):
    th = mp.init_params(9)
    d = mp.unpack(th)
    d["P"] *= 200.0
    d["wt"] = np.linspace(-1, 1, 32)
    f = mp.forward(graph, mp.pack(d))
    G, dinv = mp.preconditioner(graph, f)
    M = (G.T @ sp.diags(dinv) @ G).toarray()
    assert np.allclose(M, M.T)
    assert np.linalg.eigvalsh(M).min() > 0


def test_probes_and_bf16_rounding():
    xi = mp.probes(1000, 8, 3)
    assert set(np.unique(xi)) == {-1.0, 1.0}
    assert abs(xi.mean()) < 0.05
    assert not np.array_equal(xi, mp.probes(1000, 8, 4))
    # bf16: 8 significant bits, round to nearest even
    assert mp.bf16(np.float32(1.0 + 2 ** -8))[()] == 1.0
    assert mp.bf16(np.float32(1.0 + 3 * 2 ** -8))[()] == 1.0 + 2 ** -6
    assert mp.bf16(np.float32(1.0 + 2 ** -7))[()] == 1.0 + 2 ** -7
