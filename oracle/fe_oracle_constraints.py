"""CPU restatement of the constraints the reference applies to an assembled
system before solving (feNonLinearSolver.cpp:94-95).

Test infrastructure only: tests/, the smoke check and the bench CPU baseline
may import it, as the checker; the product path never does.

- csr_add_periodic: the (slave, master) entries of periodic pairs of
  unknowns added to the pattern (feCompressedRowStorage.cpp:96-108).
- constrain_essential_components: the loops of
  feLinearSystemMklPardiso.cpp:1039-1094, row by row in ascending order:
  every occurrence of the row's DOF in ja set to 0, then the row set to 0
  with 1 on its diagonal, then its RHS entry to 0.
- apply_periodicity: feLinearSystemMklPardiso.cpp:1099-1127 over the
  master -> slave map in ascending master order: the slave row set to 0,
  1 at the slave, -1 at the master, RHS 0.

Pinned in tests/test_oracle_constraints.py: with the outlet's tangential
velocity constrained, the Taylor-Hood Newton solution of Poiseuille flow is
the exact flow (quadratic velocity, linear pressure, both in the discrete
spaces); periodic slave rows hold exactly the two entries and the solved
slave values equal their masters.
"""
import numpy as np


def csr_add_periodic(ia, ja, n, pairs):
    """pairs: (master, slave) DOFs; entries enter only when both are < n."""
    rows = [list(ja[ia[i]:ia[i + 1]]) for i in range(n)]
    for m, s in pairs:
        if s < n and m < n:
            rows[s].append(m)
    ia2 = np.zeros(n + 1, np.int32)
    out = []
    for i, r in enumerate(rows):
        u = sorted(set(r))
        out.extend(u)
        ia2[i + 1] = ia2[i] + len(u)
    return ia2, np.array(out, np.int32)


def constrain_essential_components(ia, ja, val, rhs, dofs, n):
    """In place on val / rhs; dofs >= n are skipped."""
    rows = sorted(set(int(d) for d in dofs if d < n))
    for r in rows:
        val[ja == r] = 0.0
        for p in range(ia[r], ia[r + 1]):
            val[p] = 1.0 if ja[p] == r else 0.0
        rhs[r] = 0.0


def apply_periodicity(ia, ja, val, rhs, pairs, n):
    """In place; pairs: (master, slave), one slave per master."""
    for m, s in sorted(pairs):
        if s < n and m < n:
            for p in range(ia[s], ia[s + 1]):
                v = 0.0
                if ja[p] == s:
                    v = 1.0
                if ja[p] == m:
                    v = -1.0
                val[p] = v
            rhs[s] = 0.0
