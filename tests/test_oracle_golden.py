"""The CPU oracle against vectors produced by the reference itself
(tests/golden/*.npz, written by tests/golden/make_fixtures.py from the
reference's mesh reader, numbering, CSR builder and element kernels)."""
import numpy as np
import pytest

from tests.cases import DIFF_CASES, NS_CASES, bedge_from, coeffs_of, kind_of, load, relerr


@pytest.mark.parametrize("name", DIFF_CASES + NS_CASES)
def test_numbering_and_csr_bit_exact(oracle, name):
    g = load(name)
    conn, nvpe = g["conn"], g["nvpe"]
    V = g["xyz"].size // 3
    kind = kind_of(g)
    if nvpe == 3:
        edges, ne = oracle.tri_edges(conn)
        assert np.array_equal(edges, g["edges"])
        bedge = bedge_from(conn, edges, g["bconn"])
        bnv = 2
    else:
        edges, ne, bedge, bnv = None, 0, None, 3
    if kind == 3:
        kinds, ess = [3, 1], [1, 1]
    else:
        kinds, ess = [kind], [1]
    n, nd, adr = oracle.number(V, conn, nvpe, edges, ne, g["bconn"], bnv, bedge, kinds, ess)
    assert (n, nd) == (g["n"], g["ndofs"])
    for f in range(len(kinds)):
        assert np.array_equal(adr[f], g[f"adr{f}"])
    if kind == 3:
        u, p = adr
        forms = [(u, 12, u, 12), (p, 3, u, 12),
                 (u, 12, np.hstack([u.reshape(-1, 12), p.reshape(-1, 3)]).ravel(), 15)]
    else:
        nl = adr[0].size // (conn.size // nvpe)
        forms = [(adr[0], nl, adr[0], nl)]
    ia, ja = oracle.csr(n, forms)
    assert np.array_equal(ia, g["ia"]) and np.array_equal(ja, g["ja"])


@pytest.mark.parametrize("name", DIFF_CASES)
def test_diffusion_assembly_matches_reference(oracle, name):
    g = load(name)
    kind = kind_of(g)
    k, s = coeffs_of(g)
    val, rhs = oracle.assemble_diffusion(kind, g["xyz"], g["conn"], g["adr0"], g["n"], g["ia"], g["ja"], g["deg"],
                                         k, s, g["u0"])
    assert relerr(val, g["val"]) <= 1e-12
    assert relerr(rhs, g["rhs"]) <= 1e-12


def test_coloring_is_a_valid_vertex_coloring(oracle):
    g = load("sq8_p2_sinsin")
    col, nc = oracle.color(g["conn"], 3, g["xyz"].size // 3)
    # same partition as the reference's colorElements (fixture 'color')
    assert np.array_equal(col, g["color"])
    conn = g["conn"].reshape(-1, 3)
    for c in range(nc):
        verts = conn[col == c].ravel()
        assert np.unique(verts).size == verts.size


@pytest.mark.parametrize("name", NS_CASES)
def test_navier_stokes_assembly_matches_reference(oracle, name):
    """Taylor-Hood P2-P1 with the Kovasznay boundary data (Re = 40): the four
    NS forms (convective, divergence, source, Newtonian stress) assembled by
    the oracle against the reference's matrix and residual."""
    import ctypes as C
    from tests.oracle_bind import P
    g = load(name)
    n = g["n"]
    val = np.zeros(g["ja"].size)
    rhs = np.zeros(n)
    rc = oracle.L.or_assemble_ns(g["xyz"].size // 3, P(g["xyz"], C.c_double), g["conn"].size // 3,
                                 P(g["conn"], C.c_int), P(g["adr0"], C.c_int), P(g["adr1"], C.c_int), n,
                                 P(g["ia"], C.c_int), P(g["ja"], C.c_int), g["deg"], 1.0 / float(g["extra"]),
                                 P(g["u0"], C.c_double), None, P(val, C.c_double), P(rhs, C.c_double))
    assert rc == 0
    assert relerr(val, g["val"]) <= 1e-12
    assert relerr(rhs, g["rhs"]) <= 1e-12
