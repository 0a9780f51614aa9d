"""Pins of the constraint oracle (oracle/fe_oracle_constraints.py).

- Poiseuille flow through the unit square (Taylor-Hood, Re = 1, no essential
  spaces: inlet and walls constrained in both velocity components, the
  outlet in its tangential component only, as the reference's Poiseuille
  tests set uOutlet->setEssentialComponent(1, true)): the quadratic velocity
  and linear pressure lie in the discrete spaces, so Newton on the
  constrained oracle system returns the exact flow; without the outlet
  constraint the traction-free outlet makes it inexact.
- Periodic pairs: slave rows hold exactly 1 at the slave and -1 at its last
  master, other rows are untouched, the solution's slave values equal their
  masters.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fe_oracle_constraints as oc
from tests.oracle_bind import P


def poiseuille_case(N=6):
    """Mesh, numbering, constrained DOFs and their values, exact fields."""
    from feng_amd.api import Mesh, Numbering, TRI_P1, TRI_VP2
    m = Mesh.square(N)
    num = Numbering(m, [TRI_VP2, TRI_P1], [0, 0])
    xyz = num.dof_xyz.reshape(-1, 3)
    x, y = xyz[:, 0], xyz[:, 1]
    vel = num.dof_field == 0
    comp = num.dof_comp
    eps = 1e-12
    wall = (y < eps) | (y > 1 - eps) | (x < eps)
    outlet = (x > 1 - eps) & ~wall
    cons = np.flatnonzero(vel & (wall | (outlet & (comp == 1))))
    ux = 0.5 * y * (1 - y)
    exact = np.where(vel, np.where(comp == 0, ux, 0.0), 1.0 - x)
    u0 = np.zeros(num.nDofs)
    u0[cons] = exact[cons]
    return m, num, cons, u0, exact, vel, outlet


def oracle_ns_system(oracle, m, num, Re, deg, extra_pairs=None):
    uU, pP = num.adr
    forms = [(uU, 12, uU, 12), (pP, 3, uU, 12),
             (uU, 12, np.hstack([uU.reshape(-1, 12), pP.reshape(-1, 3)]).ravel(), 15)]
    ia, ja = oracle.csr(num.nInc, forms)
    if extra_pairs is not None:
        ia, ja = oc.csr_add_periodic(ia, ja, num.nInc, extra_pairs)
    n = num.nInc

    def assemble(u):
        val, rhs = np.zeros(ja.size), np.zeros(n)
        rc = oracle.L.or_assemble_ns(m.V, P(m.xyz, C.c_double), m.E, P(m.conn, C.c_int), P(uU, C.c_int),
                                     P(pP, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg, 1.0 / Re,
                                     P(np.ascontiguousarray(u), C.c_double), None, P(val, C.c_double),
                                     P(rhs, C.c_double))
        assert rc == 0
        return val, rhs
    return ia, ja, assemble


def oracle_newton(ia, ja, assemble, u0, n, cons, iters=4):
    u = u0.copy()
    for _ in range(iters):
        val, rhs = assemble(u)
        oc.constrain_essential_components(ia, ja, val, rhs, cons, n)
        du = spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
        u[:n] += du
    return u


@pytest.mark.parametrize("constrain_outlet", [True, False])
def test_poiseuille_is_exact_with_the_outlet_component_constrained(oracle, constrain_outlet):
    m, num, cons, u0, exact, vel, outlet = poiseuille_case()
    if not constrain_outlet:
        keep = ~np.isin(cons, np.flatnonzero(outlet))
        cons = cons[keep]
    ia, ja, assemble = oracle_ns_system(oracle, m, num, 1.0, 5)
    u = oracle_newton(ia, ja, assemble, u0, num.nInc, cons)
    err = np.abs(u - exact).max()
    if constrain_outlet:
        assert err <= 1e-10, err
    else:
        assert err > 1e-4, err


def test_constrained_rows_and_columns(oracle):
    m, num, cons, u0, exact, vel, outlet = poiseuille_case(4)
    ia, ja, assemble = oracle_ns_system(oracle, m, num, 1.0, 5)
    val, rhs = assemble(u0)
    v0, r0 = val.copy(), rhs.copy()
    oc.constrain_essential_components(ia, ja, val, rhs, cons, num.nInc)
    rows = np.repeat(np.arange(num.nInc), np.diff(ia))
    isc = np.zeros(num.nInc, bool)
    isc[cons] = True
    assert np.all(val[isc[rows]] == (ja[isc[rows]] == rows[isc[rows]]))
    assert np.all(val[~isc[rows] & isc[ja]] == 0)
    free = ~isc[rows] & ~isc[ja]
    assert np.array_equal(val[free], v0[free])
    assert np.all(rhs[isc] == 0) and np.array_equal(rhs[~isc], r0[~isc])


def periodic_pairs(num):
    """P1 DOFs on x = 0 (masters) and x = 1 (slaves) at the same height."""
    xyz = num.dof_xyz.reshape(-1, 3)
    left = np.flatnonzero(xyz[:, 0] < 1e-12)
    right = np.flatnonzero(xyz[:, 0] > 1 - 1e-12)
    right = right[np.argsort(xyz[right, 1])]
    left = left[np.argsort(xyz[left, 1])]
    return left, right


def test_periodic_rows_and_solution(oracle):
    from feng_amd.api import Mesh, Numbering, TRI_P1
    m = Mesh.square(6, 0.0, 0)
    num = Numbering(m, [TRI_P1], [0])
    master, slave = periodic_pairs(num)
    n = num.nInc
    ia0, ja0 = oracle.csr(n, [(num.adr[0], 3, num.adr[0], 3)])
    ia, ja = oc.csr_add_periodic(ia0, ja0, n, list(zip(master, slave)))
    val, rhs = oracle.assemble_diffusion(1, m.xyz, m.conn, num.adr[0], n, ia, ja, 2, (0, 1.0, 0.0, None),
                                         (1, -3.0, 0.0, None), np.zeros(num.nDofs))
    oracle.assemble_transient_mass(1, m.xyz, m.conn, num.adr[0], n, ia, ja, 2, (0, 1.0, 0.0, None), 10.0,
                                   np.zeros(num.nDofs), val, rhs)
    v0 = val.copy()
    oc.apply_periodicity(ia, ja, val, rhs, list(zip(master, slave)), n)
    rows = np.repeat(np.arange(n), np.diff(ia))
    iss = np.zeros(n, bool)
    iss[slave] = True
    assert np.array_equal(val[~iss[rows]], v0[~iss[rows]])
    mof = dict(zip(slave, master))
    for s_ in slave:
        seg = slice(ia[s_], ia[s_ + 1])
        nz = {int(j): v for j, v in zip(ja[seg], val[seg]) if v != 0}
        assert nz == {int(s_): 1.0, int(mof[s_]): -1.0}
        assert rhs[s_] == 0
    x = spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(n, n)).tocsc(), rhs)
    assert np.abs(x[slave] - x[master]).max() <= 1e-13 * np.abs(x).max()
