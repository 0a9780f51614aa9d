"""Transient path on the GPU: TRANSIENT_MASS assembled on the device with the
BDF history gathered per element, driven by feng_amd.timeint.

- The reference's golden tables scalarFEunsteady_linearODE.output (BDF1;
  BDF2 with the exact-solution starter), to the 7 printed digits, and
  scalarFEunsteady_nonlinearODE.output (u' = -u^2 through the mass-power
  form) to 7 digits or 5e-11 absolute (the reference's Newton tolerance).
- Heat equations (P2 triangles with block coefficients and a time-dependent
  quadrature-point source and Dirichlet data; P1 tets with AMG-PCG) match
  the BDF oracle (oracle/fe_oracle_time.py over the C oracle's assembly)
  at every step to 1e-10 relative: the same scheme, the device solving to
  rtol 1e-13 where the oracle solves directly.
- The device history: rotation and set/get of past solutions.
- TRANSIENT_VECTOR_MASS fused into the Taylor-Hood sweep: one residual +
  matrix at a random state and BDF2 history equals the oracle to 1e-12, and
  BDF2 steps of the unsteady Navier-Stokes stack (Newton with GMRES/ILU(0)
  on the device, Newton with direct solves in the oracle) agree to 1e-9.
"""
import numpy as np
import pytest

from oracle import fe_oracle_time as ot
from tests.test_oracle_time import (LINEAR_ODE_TABLE, NONLINEAR_ODE_TABLE, OracleScalarTransient, _qp_table,
                                    table_matches)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["BDF1", "BDF2"])
def test_linear_ode_reference_table(scheme):
    from feng_amd.api import FN_CONST, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, heat_system
    from feng_amd.timeint import BDF2_START_EXACT, TimeIntegrator, linf_nodal_error
    print("linear ODE", scheme, flush=True)
    m = Mesh.square(3)
    S, num, forms = heat_system(m, 1, 8, Fn(FN_CONST, 1.0), None, Fn(FN_CONST, 0.0), essential=False)
    err = []
    for nsteps in (10, 20, 40, 80, 160):
        ti = TimeIntegrator(S, 1 if scheme == "BDF1" else 2, 0.0, 1.0, nsteps, np.zeros(num.nDofs),
                            nl=NLSolverOptions(1e-10, 1e-10, 1e4, 10),
                            coefficients=[(forms["source"], 0, lambda t: -4.0 * t ** 3)],
                            norms=[linf_nodal_error(num, lambda t, x, y, z: t ** 4 + 0 * x)],
                            exact=lambda t: np.full(num.nDofs, t ** 4), start=BDF2_START_EXACT,
                            solve_kw=dict(rtol=1e-13))
        post = ti.make_steps()
        assert len(post[0]) == nsteps + 1 and abs(post[0][-1] - 1.0) < 1e-12
        err.append(max(post[1]))
    S.close()
    assert table_matches(err, LINEAR_ODE_TABLE[scheme]), err


@pytest.mark.parametrize("scheme", ["BDF1", "BDF2"])
def test_nonlinear_ode_reference_table(scheme):
    """u' = -u^2 through TRANSIENT_MASS + the MASS form with exponent 2
    (feSysElm_MassPower), Newton on the device."""
    from feng_amd.api import FN_CONST, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, heat_system
    from feng_amd.timeint import BDF2_START_EXACT, TimeIntegrator, linf_nodal_error
    print("nonlinear ODE", scheme, flush=True)
    m = Mesh.square(3)
    S, num, forms = heat_system(m, 1, 8, Fn(FN_CONST, 1.0), None, None, essential=False,
                                react=(Fn(FN_CONST, 1.0), 2.0))
    err = []
    for nsteps in (10, 20, 40, 80, 160):
        ti = TimeIntegrator(S, 1 if scheme == "BDF1" else 2, 0.0, 1.0, nsteps, np.ones(num.nDofs),
                            nl=NLSolverOptions(1e-13, 1e-13, 1e4, 10),
                            norms=[linf_nodal_error(num, lambda t, x, y, z: 1.0 / (1.0 + t) + 0 * x)],
                            exact=lambda t: np.full(num.nDofs, 1.0 / (1.0 + t)), start=BDF2_START_EXACT,
                            solve_kw=dict(rtol=1e-14))
        post = ti.make_steps()
        assert all(r.newton_iterations <= 6 for r in ti.reports)
        err.append(max(post[1]))
    S.close()
    assert table_matches(err, NONLINEAR_ODE_TABLE[scheme], atol=5e-11), err


def _heat_case(case, oracle):
    from feng_amd.api import FN_BLOCKS, FN_CONST, FN_ELEM, FN_QP, PC_AMG, PC_JACOBI, Fn, Mesh
    from feng_amd.problems import heat_system
    if case == "p2_blocks_bdf2":
        m = Mesh.square(10, jitter=0.2, seed=3)
        kind, deg, order, start, pc = 2, 5, 2, "BDF1", PC_JACOBI
        kdat = m.coeff_blocks(4, 1e-1, 1e1, 2)
        mass = (FN_BLOCKS, 3.0)
        base = _qp_table(m, deg, lambda x, y: np.sin(3 * x) * np.cos(2 * y) - 1.0)
    else:  # p1 tets, BDF1, AMG
        m = Mesh.cube(8)
        kind, deg, order, start, pc = 4, 2, 1, None, PC_AMG
        kdat = m.coeff_grf(0.2, 1.0, 64, 4)
        mass = (FN_CONST, 2.0)
        base = None
    k = Fn(FN_ELEM, data=kdat)
    S, num, forms = heat_system(m, kind, deg, Fn(mass[0], mass[1]), k,
                                Fn(FN_QP, data=base) if base is not None else Fn(FN_CONST, -1.0))
    xyz = num.dof_xyz.reshape(-1, 3)
    g = np.cos(xyz[:, 0] + 0.5 * xyz[:, 1]) + xyz[:, 2]
    u0 = g.copy()
    ess = lambda t: (1.0 + t * t) * g[num.nInc:]
    if base is not None:
        src_dev = lambda t: (1.0 + t) * base
        src_or = lambda t: (5, 0.0, 0.0, (1.0 + t) * base)
    else:
        src_dev = lambda t: -1.0 - t
        src_or = lambda t: (0, -1.0 - t, 0.0, None)
    prob = OracleScalarTransient(oracle, m, kind, deg, (4, 0.0, 0.0, kdat), (mass[0], mass[1], 0.0, None),
                                 src_or, 1)
    assert prob.n == num.nInc and prob.nDofs == num.nDofs
    return S, num, forms, dict(order=order, start=start, pc=pc, u0=u0, ess=ess, src_dev=src_dev, prob=prob)


@pytest.mark.parametrize("case", ["p2_blocks_bdf2", "p1_tets_bdf1_amg"])
def test_heat_equation_matches_oracle(oracle, case):
    from feng_amd.problems import NLSolverOptions
    from feng_amd.timeint import TimeIntegrator
    print("heat", case, flush=True)
    S, num, forms, c = _heat_case(case, oracle)
    nsteps = 4
    sols = []
    ti = TimeIntegrator(S, c["order"], 0.0, 0.5, nsteps, c["u0"], nl=NLSolverOptions(1e-10, 1e-10, 1e4, 10),
                        coefficients=[(forms["source"], 0, c["src_dev"])], essential=c["ess"],
                        norms=[lambda S_, t: sols.append(S_.get_solution()) or 0.0],
                        start=c["start"] or "BDF1", solve_kw=dict(rtol=1e-13, pc=c["pc"]))
    post = ti.make_steps()
    prob = c["prob"]
    times, ref = ot.integrate(prob.assemble, c["u0"], prob.n, c["order"], 0.0, 0.5, nsteps,
                              start=c["start"] or "BDF1", essential=c["ess"])
    assert np.allclose(post[0], times, rtol=0, atol=1e-14)
    assert len(sols) == len(ref) == nsteps + 1
    for t, u, r in zip(times, sols, ref):
        assert np.linalg.norm(u - r) <= 1e-10 * np.linalg.norm(r), (t, np.linalg.norm(u - r) / np.linalg.norm(r))
    # the solution moved: the comparison is not of two initial conditions
    assert np.linalg.norm(sols[-1] - sols[0]) > 1e-2 * np.linalg.norm(sols[0])
    S.close()


def test_history_rotation_and_transfer():
    from feng_amd.api import FN_CONST, Fn, Mesh
    from feng_amd.problems import heat_system
    m = Mesh.square(5)
    S, num, forms = heat_system(m, 2, 2, Fn(FN_CONST, 1.0), Fn(FN_CONST, 1.0), Fn(FN_CONST, 0.0))
    rng = np.random.default_rng(0)
    v = [rng.normal(size=num.nDofs) for _ in range(6)]
    for j in range(1, 5):
        S.set_past_solution(j, v[j])
    S.set_solution(v[0])
    S.rotate_solutions()
    # sol[1] = current, sol[j] = old sol[j - 1]; the current solution is kept
    assert np.array_equal(S.get_past_solution(1), v[0])
    for j in range(2, 5):
        assert np.array_equal(S.get_past_solution(j), v[j - 1])
    assert np.array_equal(S.get_solution(), v[0])
    S.set_essential(v[5][num.nInc:])
    u = S.get_solution()
    assert np.array_equal(u[:num.nInc], v[0][:num.nInc]) and np.array_equal(u[num.nInc:], v[5][num.nInc:])
    S.close()


def _ns_oracle(oracle, m, num, Re, deg, mass):
    import ctypes as C
    import scipy.sparse as sp
    from tests.oracle_bind import P
    uU, pP = num.adr
    forms = [(uU, 12, uU, 12), (pP, 3, uU, 12),
             (uU, 12, np.hstack([uU.reshape(-1, 12), pP.reshape(-1, 3)]).ravel(), 15)]
    ia, ja = oracle.csr(num.nInc, forms)
    n = num.nInc

    def assemble(u, udot, c0, t):
        val, rhs = np.zeros(ja.size), np.zeros(n)
        rc = oracle.L.or_assemble_ns(m.V, P(m.xyz, C.c_double), m.E, P(m.conn, C.c_int), P(uU, C.c_int),
                                     P(pP, C.c_int), n, P(ia, C.c_int), P(ja, C.c_int), deg, 1.0 / Re,
                                     P(np.ascontiguousarray(u), C.c_double), None, P(val, C.c_double),
                                     P(rhs, C.c_double))
        assert rc == 0
        oracle.assemble_vector_transient_mass(m.xyz, m.conn, uU, n, ia, ja, deg, mass, c0, udot, val, rhs)
        return sp.csr_matrix((val, ja, ia), shape=(n, n)), rhs
    return ia, ja, assemble


def test_vector_mass_assembly_matches_oracle(oracle):
    """One BDF2 residual + matrix of the unsteady Taylor-Hood stack at a
    random state and history: device vs oracle."""
    from feng_amd.api import FN_CONST, Fn, Mesh
    from feng_amd.problems import navier_stokes_system
    m = Mesh.square(12, 0.2, 1)
    S, num = navier_stokes_system(m, 40.0, 5, mass=Fn(FN_CONST, 1.5))
    ia0, ja0, assemble = _ns_oracle(oracle, m, num, 40.0, 5, 1.5)
    ia, ja = S.get_csr()
    assert np.array_equal(ia, ia0) and np.array_equal(ja, ja0)
    rng = np.random.default_rng(5)
    u = S.get_solution()
    u[:num.nInc] = 0.3 * rng.normal(size=num.nInc)
    h1, h2 = u + 0.1 * rng.normal(size=u.size), u + 0.2 * rng.normal(size=u.size)
    cn = [1.0 / 0.1 + 1.0 / 0.2, -1.0 / 0.1 - 1.0 / 0.1, 0.1 / 0.1 / 0.2]
    S.set_solution(u)
    S.set_past_solution(1, h1)
    S.set_past_solution(2, h2)
    S.set_bdf(cn)
    S.setToZero()
    S.assemble(3)
    val, rhs = S.get_system()
    A0, b0 = assemble(u, cn[0] * u + cn[1] * h1 + cn[2] * h2, cn[0], 0.0)
    assert np.linalg.norm(val - A0.data) <= 1e-12 * np.linalg.norm(A0.data)
    assert np.linalg.norm(rhs - b0) <= 1e-12 * np.linalg.norm(b0)
    S.close()


def test_unsteady_navier_stokes_matches_oracle(oracle):
    """BDF2 (BDF1 starter) on the unsteady Taylor-Hood stack from rest with
    Kovasznay boundary values: the device Newton + GMRES/ILU(0) steps equal
    the oracle's Newton + direct solves at every step."""
    from feng_amd.api import FN_CONST, KSP_GMRES, PC_ILU0, Fn, Mesh
    from feng_amd.problems import NLSolverOptions, navier_stokes_system
    from feng_amd.timeint import TimeIntegrator
    print("unsteady NS", flush=True)
    m = Mesh.square(10, 0.2, 2)
    # this stack's momentum rows carry -(u . grad u - div sigma): the mass
    # coefficient -1 keeps u_t with the same sign (the reference's unsteady
    # stack flips convection and stress instead, vectorFEunsteady.cpp:560-572)
    S, num = navier_stokes_system(m, 40.0, 5, mass=Fn(FN_CONST, -1.0))
    _, _, assemble = _ns_oracle(oracle, m, num, 40.0, 5, -1.0)
    u0 = S.get_solution()
    sols = []
    nl = NLSolverOptions(1e-11, 1e-12, 1e4, 12)
    ti = TimeIntegrator(S, 2, 0.0, 0.2, 4, u0, nl=nl, norms=[lambda S_, t: sols.append(S_.get_solution()) or 0.0],
                        solve_kw=dict(ksp=KSP_GMRES, pc=PC_ILU0, rtol=1e-13, restart=50, max_iter=20000))
    ti.make_steps()
    times, ref = ot.integrate(assemble, u0, num.nInc, 2, 0.0, 0.2, 4, start="BDF1", tol=1e-11, max_newton=12)
    assert len(sols) == len(ref) == 5
    for t, u, r in zip(times, sols, ref):
        assert np.linalg.norm(u - r) <= 1e-9 * np.linalg.norm(r), (t, np.linalg.norm(u - r) / np.linalg.norm(r))
    assert np.linalg.norm(sols[-1][:num.nInc]) > 1e-2 * np.linalg.norm(sols[-1])
    S.close()
