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


# ---------------------------------------------------------------------------
# W-cycle of the coarse levels (the single-workgroup kernel of the nested
# multigrid PCG, fe_amg_pcg.hip).  Levels 0 .. R-1 take the V-cycle above;
# each level l of R .. D-1 is solved by `gamma` cycles (the second on the
# residual b - A e of the first, e = e1 + C(b - A e1)); level D is solved with
# its exact inverse.  rounded=True models the kernel's data: the FP32 values
# and l1 weights (the FP32 rounding of the FP64 ones), the weights w rounded
# to bfloat16 and the row-scaled values h_ij = a_ij w_i rounded to binary16;
# the operator is then A~ = diag(w)^-1 H and every formula of the cycle uses
# H and w:  r = b - (H x) / w,  e = y + w b - H y.  Pinned in
# tests/test_oracle_amg.py: gamma = 1 with exact data is the V-cycle, the
# two-cycle solve has the error propagation (I - C A~)^2 of its single cycle
# C, and the bfloat16 rounding is the nearest-even one.

def round_bf16(x):
    """float32 values rounded to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def _wdata(L, rounded):
    """(H, w): row-scaled values as a CSR matrix and the weights, float64."""
    key = "_wd%d" % int(rounded)
    if not hasattr(L, key):
        A = L.A.tocsr()
        rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
        if rounded:
            w32 = round_bf16(L.l1.astype(np.float32))
            h = (A.data.astype(np.float32) * w32[rows]).astype(np.float16).astype(np.float64)
            w = w32.astype(np.float64)
        else:
            w = L.l1.copy()
            h = A.data * w[rows]
        setattr(L, key, (sp.csr_matrix((h, A.indices, A.indptr), shape=A.shape), w))
    return getattr(L, key)


def _wsolve(levels, l, b, D, rounded, gamma):
    L = levels[l]
    if l == D:
        if not hasattr(L, "AinvW"):
            L.AinvW = np.linalg.inv(L.A.toarray())
        return L.AinvW @ b
    H, w = _wdata(L, rounded)

    def cycle(rhs):
        x = w * rhs
        r = rhs - (H @ x) / w
        ec = _wsolve(levels, l + 1, np.bincount(L.agg, weights=r, minlength=L.nc), D, rounded, gamma)
        y = x + L.oc * ec[L.agg]
        return y + w * rhs - H @ y

    e = cycle(b)
    for _ in range(gamma - 1):
        e = e + cycle(b - (H @ e) / w)
    return e


def wcycle(levels, b, R, D, rounded=True, gamma=2):
    """V-cycle on levels 0 .. R-1 whose level-R coarse problem is solved by
    gamma cycles per level down to the exact level D (R <= D)."""
    def down(l, rhs):
        if l == R:
            return _wsolve(levels, l, rhs, D, rounded, gamma)
        L = levels[l]
        u = L.l1 * rhs
        r = rhs - L.A @ u
        e = down(l + 1, np.bincount(L.agg, weights=r, minlength=L.nc))
        y = u + L.oc * e[L.agg]
        return y + L.l1 * (rhs - L.A @ y)
    return down(0, np.asarray(b, dtype=np.float64))
