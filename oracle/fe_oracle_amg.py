"""CPU restatement of the aggregation multigrid preconditioner (FE_PC_AMG).

Test infrastructure only: tests/, the smoke check and the bench CPU baseline
may import it, as the checker; the product path never does.

The reference reaches algebraic multigrid only through PETSc's runtime
options (`-pc_type gamg` read by the KSP set-up in
feLinearSystemPETSc.cpp:900-1031), so nothing of the reference pins these
functions.  They are pinned by mathematics in tests/test_oracle_amg.py:
  - the aggregation is the greedy maximal independent set in descending
    hashed priority, checked against Luby's parallel rounds with the same
    priorities (a different algorithm with the same fixed point), and the
    independence / maximality / root-adjacency invariants;
  - the Galerkin operator equals P^T A P computed by sparse matrix products;
  - the two-level cycle equals the textbook error propagation
    E = (I - S A)(I - oc P A_c^-1 P^T A)(I - S A), M^-1 = (I - E) A^-1,
    formed densely; a single level is the exact inverse;
  - the multilevel cycle is symmetric positive definite, also when coarsening
    stalls on a block-diagonal matrix and the coarsest level is swept.
"""
import numpy as np
import scipy.sparse as sp

DENSE = 64  # the coarsest level is inverted densely up to this many rows
STALL = 0.9  # coarsening stops when a level keeps more than this fraction of its rows
SWEEPS = 4   # l1-Jacobi sweeps that solve a stalled coarsest level
MASK = (1 << 64) - 1


def mix64(i):
    """SplitMix64 finaliser of the row index (the hashed priority)."""
    x = (int(i) + 0x9E3779B97F4A7C15) & MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def priorities(n):
    return np.array([mix64(i) for i in range(n)], dtype=np.uint64)


def aggregate(ia, ja):
    """Greedy maximal independent set of the graph of the CSR pattern in
    descending priority (ties by row), roots numbered in row order; every
    other row joins the largest-numbered neighbouring root.  Returns (agg, nc)."""
    n = len(ia) - 1
    pri = priorities(n)
    order = sorted(range(n), key=lambda i: (-int(pri[i]), i))
    state = np.zeros(n, np.int8)
    for i in order:
        if state[i] != 0:
            continue
        state[i] = 1
        nb = ja[ia[i]:ia[i + 1]]
        state[nb[state[nb] == 0]] = -1
    rid = np.full(n, -1, np.int64)
    roots = np.flatnonzero(state == 1)
    rid[roots] = np.arange(len(roots))
    agg = rid.copy()
    for i in np.flatnonzero(state != 1):
        agg[i] = rid[ja[ia[i]:ia[i + 1]]].max()
    return agg, len(roots)


def tentative(agg, nc):
    n = len(agg)
    return sp.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, nc))


def galerkin(A, agg, nc):
    """A_c[I, J] = sum of a_ij over i in I, j in J (explicit zeros kept)."""
    A = A.tocoo()
    C = sp.coo_matrix((A.data, (agg[A.row], agg[A.col])), shape=(nc, nc)).tocsr()
    C.sum_duplicates()
    C.sort_indices()
    return C


def l1_scale(A, c):
    s = np.asarray(abs(A).sum(axis=1)).ravel()
    return np.where(s > 0, c / np.where(s > 0, s, 1.0), 0.0)


class Level:
    pass


def hierarchy(A, c=1.8, oc=1.8, dense=DENSE, max_levels=20):
    """Levels of the CSR matrix A (scipy CSR with sorted indices)."""
    levels = []
    A = A.tocsr()
    while True:
        L = Level()
        L.A = A
        L.l1 = l1_scale(A, c)
        L.c = c
        levels.append(L)
        if A.shape[0] <= dense:
            L.Ainv = np.linalg.inv(A.toarray())
            break
        agg, nc = aggregate(A.indptr, A.indices)
        if len(levels) >= max_levels or nc > STALL * A.shape[0]:
            L.sweeps = SWEEPS  # stalled: the coarsest level is swept, not inverted
            break
        L.agg, L.nc = agg, nc
        A = galerkin(A, L.agg, L.nc)
    for L in levels:
        L.oc = oc
    return levels


def coarse_solve(L, b):
    """Dense inverse, or (stalled coarsening) SWEEPS l1-Jacobi sweeps from zero
    with weights 1 / sum_j |a_ij|."""
    if hasattr(L, "Ainv"):
        return L.Ainv @ b
    w = L.l1 / L.c
    x = w * b
    for _ in range(L.sweeps - 1):
        x = x + w * (b - L.A @ x)
    return x


def vcycle(levels, b, l=0):
    """u = S b; b_c = P^T (b - A u); e = V(b_c); y = u + oc P e; x = y + S (b - A y)."""
    L = levels[l]
    if l == len(levels) - 1:
        return coarse_solve(L, b)
    u = L.l1 * b
    r = b - L.A @ u
    bc = np.bincount(L.agg, weights=r, minlength=L.nc)
    e = vcycle(levels, bc, l + 1)
    y = u + L.oc * e[L.agg]
    return y + L.l1 * (b - L.A @ y)
