"""Constraints on the device system (fe_constraints.hip) against the oracle
(oracle/fe_oracle_constraints.py, pinned in tests/test_oracle_constraints.py).

- Poiseuille flow with the outlet's tangential velocity constrained (the
  reference's uOutlet->setEssentialComponent(1, true)): the constrained
  Taylor-Hood system equals the oracle's (constrained entries exactly 0 / 1,
  others to 1e-12), and Newton with GMRES/ILU(0) on the device returns the
  exact flow.
- Periodic pairs: the CSR pattern gains the (slave, master) entries exactly
  as the oracle's; after applyPeriodicity the system equals the oracle's and
  the solved slave values equal their masters.
"""
import numpy as np
import pytest

from oracle import fe_oracle_constraints as oc
from tests.test_oracle_constraints import oracle_newton, oracle_ns_system, periodic_pairs, poiseuille_case

pytestmark = pytest.mark.gpu


def test_poiseuille_constrained_system_and_exact_solution(oracle):
    from feng_amd.api import KSP_GMRES, PC_ILU0
    from feng_amd.problems import NLSolverOptions, navier_stokes_system, solve_newton
    print("poiseuille", flush=True)
    m, num, cons, u0, exact, vel, outlet = poiseuille_case(8)
    S, num2 = navier_stokes_system(m, 1.0, 5, essential=False)
    assert num2.nInc == num.nDofs
    S.set_constrained_dofs(cons)
    S.set_solution(u0)
    S.setToZero()
    S.assemble(3)
    S.constrainEssentialComponents()
    val, rhs = S.get_system()
    ia, ja, assemble = oracle_ns_system(oracle, m, num, 1.0, 5)
    v0, r0 = assemble(u0)
    oc.constrain_essential_components(ia, ja, v0, r0, cons, num.nInc)
    ia_d, ja_d = S.get_csr()
    assert np.array_equal(ia_d, ia) and np.array_equal(ja_d, ja)
    hard = (v0 == 0) | (v0 == 1)
    rows = np.repeat(np.arange(num.nInc), np.diff(ia))
    isc = np.zeros(num.nInc, bool)
    isc[cons] = True
    touched = isc[rows] | isc[ja]
    assert np.array_equal(val[touched], v0[touched]) and hard[touched].all()
    assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.array_equal(rhs[isc], r0[isc]) and np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0)
    rep = solve_newton(S, NLSolverOptions(1e-10, 1e-12, 1e4, 8), ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-12,
                       restart=50, max_iter=20000)
    assert rep.converged
    u = S.get_solution()
    assert np.abs(u - exact).max() <= 1e-9, np.abs(u - exact).max()
    uo = oracle_newton(ia, ja, assemble, u0, num.nInc, cons)
    assert np.abs(u - uo).max() <= 1e-9
    S.close()


def test_periodic_pattern_system_and_solution(oracle):
    from feng_amd.api import FN_CONST, FN_SINSIN, KSP_GMRES, PC_ILU0, Fn, Mesh, Numbering, TRI_P1
    from feng_amd.problems import heat_system
    print("periodic", flush=True)
    m = Mesh.square(12, 0.0, 0)
    num0 = Numbering(m, [TRI_P1], [0])
    master, slave = periodic_pairs(num0)
    S, num, forms = heat_system(m, TRI_P1, 2, Fn(FN_CONST, 1.0), Fn(FN_CONST, 1.0), Fn(FN_SINSIN, -3.0),
                                essential=False, periodic=(master, slave))
    n = num.nInc
    ia0, ja0 = oracle.csr(n, [(num.adr[0], 3, num.adr[0], 3)])
    ia, ja = oc.csr_add_periodic(ia0, ja0, n, list(zip(master, slave)))
    ia_d, ja_d = S.get_csr()
    assert np.array_equal(ia_d, ia) and np.array_equal(ja_d, ja) and ja.size > ja0.size
    S.set_bdf([10.0])
    S.setToZero()
    S.assemble(3)
    S.applyPeriodicity()
    val, rhs = S.get_system()
    v0, r0 = oracle.assemble_diffusion(1, m.xyz, m.conn, num.adr[0], n, ia, ja, 2, (0, 1.0, 0.0, None),
                                       (1, -3.0, 0.0, None), np.zeros(num.nDofs))
    oracle.assemble_transient_mass(1, m.xyz, m.conn, num.adr[0], n, ia, ja, 2, (0, 1.0, 0.0, None), 10.0,
                                   np.zeros(num.nDofs), v0, r0)
    oc.apply_periodicity(ia, ja, v0, r0, list(zip(master, slave)), n)
    assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0)
    rows = np.repeat(np.arange(n), np.diff(ia))
    iss = np.zeros(n, bool)
    iss[slave] = True
    assert np.array_equal(val[iss[rows]], v0[iss[rows]]) and np.all(rhs[slave] == 0)
    r, du = S.solve(ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-12, restart=50, max_iter=5000, want_du=True)
    assert r.converged
    assert np.abs(du[slave] - du[master]).max() <= 1e-10 * np.abs(du).max()
    S.close()


def test_periodic_channel_with_pairs_from_the_numbering(oracle):
    """Inlet -> Outlet periodicity on the reference's Poiseuille mesh (P2,
    walls essential, an x-dependent source): the pairs of
    fe_numbering_periodic (feNumber::applyPeriodicity) enter the device
    pattern and system exactly as the oracle's restatement of the reference's
    CSR and applyPeriodicity loops build them; the device direct solve ties
    every slave to its master and equals the oracle's system solved on the
    host."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from feng_amd.api import FN_CONST, FN_SINSIN, KSP_DIRECT, Fn, Mesh, Numbering, TRI_P2
    from feng_amd.problems import heat_system
    from tests.conftest import GOLDEN
    print("periodic channel", flush=True)
    m = Mesh.read_msh(GOLDEN / "meshes" / "data" / "poiseuille0.msh")
    master, slave = Numbering(m, [TRI_P2], [["NoSlip"]]).periodic_dofs(0, "Inlet", "Outlet", (5.0, 0.0))
    S, num, forms = heat_system(m, TRI_P2, 4, Fn(FN_CONST, 1.0), Fn(FN_CONST, 1.0), Fn(FN_SINSIN, -3.0),
                                essential=["NoSlip"], periodic=(master, slave))
    n = num.nInc
    pairs = list(zip(master, slave))
    inner = [(a, b) for a, b in pairs if a < n and b < n]
    assert len(inner) >= 3  # the walls' corner DOFs are essential, the others tie
    ia0, ja0 = oracle.csr(n, [(num.adr[0], 6, num.adr[0], 6)])
    ia, ja = oc.csr_add_periodic(ia0, ja0, n, pairs)
    ia_d, ja_d = S.get_csr()
    assert np.array_equal(ia_d, ia) and np.array_equal(ja_d, ja) and ja.size > ja0.size
    S.set_bdf([10.0])
    S.setToZero()
    S.assemble(3)
    S.applyPeriodicity()
    val, rhs = S.get_system()
    v0, r0 = oracle.assemble_diffusion(2, m.xyz, m.conn, num.adr[0], n, ia, ja, 4, (0, 1.0, 0.0, None),
                                       (1, -3.0, 0.0, None), np.zeros(num.nDofs))
    oracle.assemble_transient_mass(2, m.xyz, m.conn, num.adr[0], n, ia, ja, 4, (0, 1.0, 0.0, None), 10.0,
                                   np.zeros(num.nDofs), v0, r0)
    oc.apply_periodicity(ia, ja, v0, r0, pairs, n)
    assert np.linalg.norm(val - v0) <= 1e-12 * np.linalg.norm(v0)
    assert np.linalg.norm(rhs - r0) <= 1e-12 * np.linalg.norm(r0)
    r, du = S.solve(ksp=KSP_DIRECT, rtol=1e-12, want_du=True)
    ref = spla.spsolve(sp.csr_matrix((v0, ja, ia), shape=(n, n)).tocsc(), r0)
    assert r.converged and np.abs(du - ref).max() <= 1e-10 * np.abs(ref).max()
    ms, ss = np.array(inner).T
    assert np.abs(du[ss] - du[ms]).max() <= 1e-12 * np.abs(du).max()
    S.close()
