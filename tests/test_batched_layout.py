"""Host layout of the batched system of both element degrees
(feng_amd.batched.union_layout): problem p's unknowns are the rows
[rows[p], rows[p+1]); each field maps its own degree's elements through the
problem's numbering (feNumber.cpp:677-843) and the other degree's elements
to one extra essential DOF.  The union CSR pattern built by the oracle
(feCompressedRowStorage.cpp:15-133 restated) from both fields' maps is
block-diagonal, and each block is the pattern of the problem alone."""
import numpy as np

from feng_amd import api as A
from feng_amd.batched import union_layout

PROBLEMS = [(3, A.TRI_P1, 1), (4, A.TRI_P2, 2), (2, A.TRI_P2, 3), (5, A.TRI_P1, 4), (3, A.TRI_P2, 5)]


def test_union_layout_partitions_rows_and_maps_the_other_degree_to_one_dof():
    U = union_layout(PROBLEMS)
    rows, NU, dummy = U["rows"], U["NU"], U["dummy"]
    assert rows[0] == 0 and rows[-1] == NU and np.all(np.diff(rows) == U["nInc"])
    ess = np.concatenate([[0], np.cumsum([nm.nDofs - nm.nInc for nm in U["nums"]])]) + NU
    assert dummy == ess[-1]
    e0 = 0
    for p, (N_, kind, _) in enumerate(PROBLEMS):
        m, nm = U["meshes"][p], U["nums"][p]
        for k, nl in ((A.TRI_P1, 3), (A.TRI_P2, 6)):
            a = U["adr"][k][e0 * nl:(e0 + m.E) * nl]
            if k != kind:
                assert np.all(a == dummy)
                continue
            assert np.array_equal(a, U["maps"][p][nm.adr[0]])
            unk = a[a < NU]
            assert unk.min() == rows[p] and unk.max() == rows[p + 1] - 1
            es = a[a >= NU]
            assert np.all((es >= ess[p]) & (es < ess[p + 1]))
        e0 += m.E
    assert e0 * 3 == U["adr"][A.TRI_P1].size and e0 * 6 == U["adr"][A.TRI_P2].size
    assert U["V"] == sum(m.V for m in U["meshes"]) and U["xyz"].size == 3 * U["V"]
    assert U["conn"].min() >= 0 and U["conn"].max() < U["V"] and U["coeff"].size == e0


def test_union_pattern_is_block_diagonal_with_each_problems_own_pattern(oracle):
    U = union_layout(PROBLEMS)
    rows, NU = U["rows"], U["NU"]
    a1, a2 = U["adr"][A.TRI_P1], U["adr"][A.TRI_P2]
    ia, ja = oracle.csr(NU, [(a1, 3, a1, 3), (a2, 6, a2, 6)])
    for p, (N_, kind, _) in enumerate(PROBLEMS):
        nm = U["nums"][p]
        nl = nm.adr[0].size // U["meshes"][p].E
        ia0, ja0 = oracle.csr(nm.nInc, [(nm.adr[0], nl, nm.adr[0], nl)])
        r0, r1 = int(rows[p]), int(rows[p + 1])
        assert np.array_equal(ia[r0:r1 + 1] - ia[r0], ia0)
        assert np.array_equal(ja[ia[r0]:ia[r1]] - r0, ja0)
