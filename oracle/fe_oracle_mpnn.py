"""CPU oracle of the learned IC(0)-pattern preconditioner (message-passing net).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker;
the product path (feng_amd/csrc/fe_learned.hip) never calls it.

The reference has no learned preconditioner: its PCG back end is PETSc's
KSPCG with -pc_type icc (feLinearSystemPETSc.cpp:900-1031).  The network
corrects the factor of the split IC(0) preconditioner
    M^-1 = G^T D^-1 G,   G = I - L_strict (unit lower),  D = IC(0) pivots,
that the device path applies.  This module restates, in float64 numpy with
optional rounding at the points where the device kernels store bfloat16, the
whole computation: features, the 3-layer 32-wide message-passing net, the
corrected factor, the Kaporin loss and its gradient (written out by hand, in
the order of the device backward kernels).  Pins (tests/test_oracle_mpnn.py):
central finite differences of its own forward (the gradient), the zero-head
net reproduces the IC(0) factor exactly, the per-row minimiser of the loss is
the FSAI factor (closed form by dense solves), and G^T D^-1 G is SPD.

Normalised coordinates.  With s_i = a_ii^-1/2 and S = diag(s):
    A^ = S A S (unit diagonal),  G^ = S G S^-1 (g^_ij = g_ij s_i / s_j),
    D^ = S D S,
so that G^T D^-1 G = S G^^T D^^-1 G^ S.

Net (H = 32 hidden, parameters in one float32 vector, layout `LAYOUT`):
    x_i  = [1, d^0_i, sum_j!=i |a^_ij|, sum_j!=i a^_ij, deg_i / 16,
            sum_j<i g^0_ij^2, sum_j<i w0_ij^2, 0 ... 0]   (16 features)
    h0   = relu(x W0 + b0)
    m_l  = A^_off h_{l-1};  h_l = relu([h_{l-1}, m_l] W_l + b_l),  l = 1..3
    p = h3 P, q = h3 Q (8 wide), [tau, s, r, sigma, rho] = h3 Hs + hb.
Corrected factor, on the strict lower entries (i, j), j < i:
    w0_ij = (A^ g^0_i)_j   (row i of G^0 times A^, at column j)
    g^_ij = g^0_ij - (sigma_i + rho_j) w0_ij + a^_ij (p_i . q_j + s_i + r_j)
    d^_i  = B_ii exp(clamp(tau_i, -4, 4)),  B_ii = g^_i^T A^ g^_i.
The w0 term moves g^ along the negative gradient of the loss at the IC(0)
factor, so the net learns step lengths; the a^ term adds learned corrections.

Loss: Kaporin's condition number of A~ = D^^-1/2 G^ A^ G^^T D^^-1/2,
    K = (tr A~ / n)^n / det A~,  which bounds the CG iteration count
(Kaporin, Numer. Linear Algebra Appl. 1 (1994)); up to the constant
-log det A^ / n,
    L = log K / n = log((1/n) sum_i e^-tau_i) + mean(tau) + mean(log B_ii).
Its minimiser over the pattern is the factorized sparse approximate inverse
(FSAI) of A^ (Kolotilina & Yeremin 1993), with tau = 0.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

H = 32
NIN = 16
NPQ = 8
NS = 5  # scalar heads: tau, s, r, sigma, rho
TAU_CLAMP = 4.0

# (name, shape) in storage order
LAYOUT = [("W0", (NIN, H)), ("b0", (H,)),
          ("W1", (2 * H, H)), ("b1", (H,)),
          ("W2", (2 * H, H)), ("b2", (H,)),
          ("W3", (2 * H, H)), ("b3", (H,)),
          ("P", (H, NPQ)), ("Q", (H, NPQ)),
          ("Hs", (H, NS)), ("hb", (NS,))]


def param_count() -> int:
    return int(sum(np.prod(s) for _, s in LAYOUT))


def unpack(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        c = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + c], np.float64).reshape(shape)
        o += c
    return out


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in LAYOUT])


def splitmix64(x):
    """SplitMix64 finaliser of x + golden gamma (uint64 numpy arrays)."""
    x = (np.asarray(x, np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def init_params(seed: int) -> np.ndarray:
    """Uniform init: weights U(-r, r) with r = sqrt(6 / fan_in) (heads P, Q
    scaled by 0.05 so the untrained net stays close to IC(0)), biases and the
    scalar heads zero.  u_q = (splitmix64(seed * 2^32 + q) >> 11) 2^-53."""
    theta = np.zeros(param_count(), np.float64)
    o = 0
    with np.errstate(over="ignore"):
        for name, shape in LAYOUT:
            c = int(np.prod(shape))
            if name in ("W0", "W1", "W2", "W3", "P", "Q"):
                fan_in = shape[0]
                r = np.sqrt(6.0 / fan_in) * (0.05 if name in ("P", "Q") else 1.0)
                q = np.arange(o, o + c, dtype=np.uint64)
                key = (np.uint64(seed & 0xFFFFFFFF) << np.uint64(32)) + q
                u = (splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
                theta[o:o + c] = r * (2.0 * u - 1.0)
            o += c
    return theta.astype(np.float32).astype(np.float64)


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Matrix graph quantities of one SPD CSR matrix and its IC(0) factor
    (lu on the same pattern: strict lower = L, upper incl. diagonal = U)."""

    def __init__(self, ia, ja, val, lu):
        ia = np.asarray(ia, np.int64)
        ja = np.asarray(ja, np.int64)
        val = np.asarray(val, np.float64)
        lu = np.asarray(lu, np.float64)
        n = len(ia) - 1
        self.n = n
        self.ia = ia
        rows = np.repeat(np.arange(n), np.diff(ia))
        self.rows, self.cols = rows, ja
        diag_mask = rows == ja
        dpos = np.nonzero(diag_mask)[0]
        assert len(dpos) == n, "every row must store its diagonal"
        a_ii = val[dpos]
        self.s = 1.0 / np.sqrt(a_ii)
        self.ahat = val * self.s[rows] * self.s[ja]
        self.lower = rows > ja
        off = ~diag_mask
        self.Aoff = sp.csr_matrix((np.where(off, self.ahat, 0.0), ja, ia), shape=(n, n))
        self.Ahat = sp.csr_matrix((self.ahat, ja, ia), shape=(n, n))
        self.d0 = lu[dpos] / a_ii
        self.g0 = np.where(self.lower, -lu * self.s[rows] / self.s[ja], 0.0)
        self.deg = np.bincount(rows[off], minlength=n).astype(np.float64)
        self.a_ii = a_ii
        _, self.w0 = row_products(self, self.g0)

    def features(self, round_bf16=True):
        n = self.n
        X = np.zeros((n, NIN))
        X[:, 0] = 1.0
        X[:, 1] = self.d0
        off = self.rows != self.cols
        X[:, 2] = np.bincount(self.rows[off], weights=np.abs(self.ahat[off]), minlength=n)
        X[:, 3] = np.bincount(self.rows[off], weights=self.ahat[off], minlength=n)
        X[:, 4] = self.deg / 16.0
        X[:, 5] = np.bincount(self.rows, weights=self.g0 ** 2, minlength=n)
        X[:, 6] = np.bincount(self.rows, weights=self.w0 ** 2, minlength=n)
        return bf16(X) if round_bf16 else X

    def factor(self, g):
        """Lower CSR matrix G^ = I + strict-lower values g."""
        n = self.n
        m = self.lower
        return (sp.csr_matrix((g[m], (self.rows[m], self.cols[m])), shape=(n, n)) + sp.identity(n)).tocsr()


def row_products(graph: Graph, g):
    """B_ii = g^_i^T A^ g^_i for every row of G^ = I + strict lower g, and
    w_e = (A^ g^_i)_j on every strict lower entry e = (i, j) (0 elsewhere)."""
    G = graph.factor(g)
    AG = (G @ graph.Ahat).tocsr()  # row i holds (A^ g^_i)^T (A^ symmetric)
    B = np.asarray(AG.multiply(G).sum(axis=1)).ravel()
    w = np.asarray(AG[graph.rows, graph.cols]).ravel()
    return B, np.where(graph.lower, w, 0.0)


def forward(graph: Graph, theta, round_bf16=True):
    """Net forward and corrected factor.  Returns every intermediate the
    device kernels store (X, h0..h3, m1..m3, p, q, heads) and g, B, d."""
    R = bf16 if round_bf16 else (lambda v: np.asarray(v, np.float64))
    W = unpack(theta)
    out = {}
    X = graph.features(round_bf16)
    out["X"] = X
    z = X @ R(W["W0"]) + W["b0"]
    h = R(np.maximum(z, 0.0))
    out["h0"] = h
    for l in (1, 2, 3):
        m = R(graph.Aoff @ h)
        xin = np.concatenate([h, m], axis=1)
        z = xin @ R(W[f"W{l}"]) + W[f"b{l}"]
        out[f"m{l}"] = m
        h = R(np.maximum(z, 0.0))
        out[f"h{l}"] = h
    out["p"] = h @ W["P"]
    out["q"] = h @ W["Q"]
    hs = h @ W["Hs"] + W["hb"]
    out["heads"] = hs
    tau, sv, rv, sig, rho = (hs[:, k] for k in range(NS))
    rows, cols = graph.rows, graph.cols
    c = np.einsum("ek,ek->e", out["p"][rows], out["q"][cols]) + sv[rows] + rv[cols]
    g = graph.g0 - (sig[rows] + rho[cols]) * graph.w0 + graph.ahat * c
    out["g"] = np.where(graph.lower, g, 0.0)
    B, w = row_products(graph, out["g"])
    out["B"], out["w"] = B, w
    out["d"] = B * np.exp(np.clip(tau, -TAU_CLAMP, TAU_CLAMP))
    return out


def loss_and_grad(graph: Graph, theta, round_bf16=True, want_grad=True):
    """Kaporin loss L = log(mean e^-tau) + mean(tau) + mean(log B) and its
    gradient with respect to theta, backward in the device kernels' order:
    edges -> heads -> layers 3..1 (dense, then aggregation) -> layer 0."""
    R = bf16 if round_bf16 else (lambda v: np.asarray(v, np.float64))
    W = unpack(theta)
    f = forward(graph, theta, round_bf16)
    n = graph.n
    hs = f["heads"]
    tau = np.clip(hs[:, 0], -TAU_CLAMP, TAU_CLAMP)
    et = np.exp(-tau)
    S = et.sum()
    B, w = f["B"], f["w"]
    loss = float(np.log(S / n) + tau.mean() + np.log(B).mean())
    if not want_grad:
        return loss, None, f
    rows, cols, low = graph.rows, graph.cols, graph.lower
    ghs = np.zeros((n, NS))
    active = (hs[:, 0] > -TAU_CLAMP) & (hs[:, 0] < TAU_CLAMP)
    ghs[:, 0] = np.where(active, (1.0 - n * et / S) / n, 0.0)
    ge = np.where(low, 2.0 * w / B[rows] / n, 0.0)      # dL/dg^_ij
    gc = ge * graph.ahat                                # dL/dc_ij
    gsr = -ge * graph.w0                                # dL/d(sigma_i + rho_j)
    gp = np.zeros((n, NPQ))
    gq = np.zeros((n, NPQ))
    np.add.at(gp, rows, gc[:, None] * f["q"][cols])
    np.add.at(gq, cols, gc[:, None] * f["p"][rows])
    ghs[:, 1] = np.bincount(rows, weights=gc, minlength=n)
    ghs[:, 2] = np.bincount(cols, weights=gc, minlength=n)
    ghs[:, 3] = np.bincount(rows, weights=gsr, minlength=n)
    ghs[:, 4] = np.bincount(cols, weights=gsr, minlength=n)
    grads = {k: np.zeros_like(v) for k, v in W.items()}
    h3 = f["h3"]
    grads["P"] = h3.T @ gp
    grads["Q"] = h3.T @ gq
    grads["Hs"] = h3.T @ ghs
    grads["hb"] = ghs.sum(axis=0)
    gh = gp @ W["P"].T + gq @ W["Q"].T + ghs @ W["Hs"].T
    for l in (3, 2, 1):
        hprev = f[f"h{l - 1}"]
        gz = R(np.where(f[f"h{l}"] > 0, gh, 0.0))
        xin = np.concatenate([hprev, f[f"m{l}"]], axis=1)
        grads[f"W{l}"] = xin.T @ gz
        grads[f"b{l}"] = gz.sum(axis=0)
        gx = gz @ R(W[f"W{l}"]).T
        gh = gx[:, :H] + graph.Aoff @ gx[:, H:]
    gz = R(np.where(f["h0"] > 0, gh, 0.0))
    grads["W0"] = f["X"].T @ gz
    grads["b0"] = gz.sum(axis=0)
    return loss, pack(grads), f


def fsai(graph: Graph):
    """Closed-form minimiser of mean(log B) over the strict lower pattern:
    row i solves A^[P_i, P_i] y = e_i (P_i = {j <= i in the pattern}) and
    g^_i = y / y_i (Kolotilina & Yeremin 1993)."""
    g = np.zeros(len(graph.cols))
    Ah = graph.Ahat.tocsr()
    for i in range(graph.n):
        P = np.arange(graph.ia[i], graph.ia[i + 1])
        P = P[graph.cols[P] <= i]
        c = graph.cols[P]
        e = np.zeros(len(c))
        e[-1] = 1.0
        y = np.linalg.solve(Ah[c][:, c].toarray(), e)
        g[P[:-1]] = y[:-1] / y[-1]
    return np.where(graph.lower, g, 0.0)


def preconditioner(graph: Graph, f):
    """(G, dinv) of the corrected split preconditioner in the original
    coordinates: g_ij = g^_ij s_j / s_i on the strict lower entries (unit
    diagonal), dinv_i = 1 / (d^_i a_ii)."""
    g = np.where(graph.lower, f["g"] * graph.s[graph.cols] / graph.s[graph.rows], 0.0)
    dinv = 1.0 / (f["d"] * graph.a_ii)
    return graph.factor(g), dinv


def adam(theta, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Alg. 1) on float32 master weights."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
