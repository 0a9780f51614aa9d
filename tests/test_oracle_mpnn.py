"""Pins of the learned-preconditioner oracle (oracle/fe_oracle_mpnn.py).

The reference has no learned preconditioner, so nothing of the reference pins
these functions; they are pinned by mathematics instead:
  - the hand-written gradient against central finite differences of the
    forward, on every parameter group;
  - the zero-head net keeps the IC(0) split factor G = I - L_strict (ILU(0)
    of the C oracle, the counterpart of feLinearSystemPETSc's -pc_type icc)
    and takes the Kaporin-optimal pivots diag(G A G^T), computed densely;
  - the FSAI factor satisfies its optimality condition (A^ g^_i)_j = 0 on the
    strict lower pattern, and on a full pattern it is the exact inverse
    factor: K = 1, so the loss equals log det A^ / n;
  - G^T D^-1 G is SPD for arbitrary weights (PCG with it cannot break down).
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


def graph_of(ia, ja, val):
    lu, rc = load_oracle().ilu0(ia, ja, val)
    assert rc == 0
    return mp.Graph(ia, ja, val, lu)


@pytest.fixture(scope="module")
def graph():
    return graph_of(*grid_matrix())


def dense_spd(n=9, seed=4):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(n, n))
    S = sp.csr_matrix(B @ B.T + n * np.eye(n))
    S.sort_indices()
    return S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()


def test_param_layout():
    assert mp.param_count() == 16 * 32 + 32 + 3 * (64 * 32 + 32) + 2 * 32 * 8 + 32 * 5 + 5
    th = mp.init_params(5)
    assert np.allclose(mp.pack(mp.unpack(th)), th)
    d = mp.unpack(th)
    assert np.all(d["b1"] == 0) and np.all(d["Hs"] == 0) and np.all(d["hb"] == 0)


def _perturbed(seed=7):
    rng = np.random.default_rng(seed)
    d = mp.unpack(mp.init_params(seed))
    d["P"] *= 20.0
    d["Q"] *= 20.0
    d["Hs"] = rng.normal(0, 0.05, d["Hs"].shape)
    d["hb"] = np.array([0.1, 0.02, -0.03, 0.3, 0.2])
    return mp.pack(d)


def test_gradient_matches_finite_differences(graph):
    rng = np.random.default_rng(0)
    th = _perturbed()
    L0, g, _ = mp.loss_and_grad(graph, th, round_bf16=False)
    o = 0
    checked = 0
    for name, shape in mp.LAYOUT:
        c = int(np.prod(shape))
        for q in rng.choice(c, size=min(c, 6), replace=False):
            k = o + q
            h = 1e-6 * max(1.0, abs(th[k]))
            tp = th.copy(); tp[k] += h
            tm = th.copy(); tm[k] -= h
            fd = (mp.loss_and_grad(graph, tp, round_bf16=False, want_grad=False)[0] -
                  mp.loss_and_grad(graph, tm, round_bf16=False, want_grad=False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * np.abs(g).max(), (name, q, fd, g[k])
            checked += 1
        o += c
    assert checked >= 50


def test_zero_heads_keep_ic0_factor(graph):
    d = mp.unpack(mp.init_params(2))
    d["P"][:] = 0.0
    f = mp.forward(graph, mp.pack(d))
    assert np.array_equal(f["g"], graph.g0)
    G = graph.factor(graph.g0).toarray()
    Bd = np.diag(G @ graph.Ahat.toarray() @ G.T)
    assert np.allclose(f["d"], Bd, rtol=1e-13, atol=0)
    Gm, dinv = mp.preconditioner(graph, f)
    # G = I - L_strict in the original coordinates
    n = graph.n
    L = sp.csr_matrix((np.where(graph.lower, graph.g0 * graph.s[graph.cols] / graph.s[graph.rows], 0.0),
                       graph.cols, graph.ia), shape=(n, n))
    assert abs(Gm - (L + sp.identity(n))).max() < 1e-14


def test_fsai_optimality_and_full_pattern_identity(graph):
    gf = mp.fsai(graph)
    B, w = mp.row_products(graph, gf)
    assert np.abs(w).max() < 1e-12
    assert np.log(B).mean() < np.log(mp.row_products(graph, graph.g0)[0]).mean()
    # full pattern: FSAI = exact inverse factor, K = 1, L = log det(A^) / n
    g2 = graph_of(*dense_spd())
    gf2 = mp.fsai(g2)
    B2, _ = mp.row_products(g2, gf2)
    loss_fsai = float(np.log(B2).mean())  # tau = 0
    sign, logdet = np.linalg.slogdet(g2.Ahat.toarray())
    assert sign > 0 and abs(loss_fsai - logdet / g2.n) < 1e-12
    G = g2.factor(gf2).toarray()
    At = np.diag(B2 ** -0.5) @ G @ g2.Ahat.toarray() @ G.T @ np.diag(B2 ** -0.5)
    assert np.allclose(At, np.eye(g2.n), atol=1e-12)


def test_corrected_preconditioner_is_spd(graph):
    d = mp.unpack(_perturbed(9))
    d["P"] *= 10.0
    d["hb"][3] = 2.0
    f = mp.forward(graph, mp.pack(d))
    G, dinv = mp.preconditioner(graph, f)
    M = (G.T @ sp.diags(dinv) @ G).toarray()
    assert np.allclose(M, M.T)
    assert np.linalg.eigvalsh(M).min() > 0


def test_bf16_rounding():
    assert mp.bf16(np.float32(1.0 + 2 ** -8))[()] == 1.0
    assert mp.bf16(np.float32(1.0 + 3 * 2 ** -8))[()] == 1.0 + 2 ** -6
    assert mp.bf16(np.float32(1.0 + 2 ** -7))[()] == 1.0 + 2 ** -7
