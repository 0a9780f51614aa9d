"""CPU restatement of the BDF time integration (BDF1 / BDF2) of the reference.

Test infrastructure only: tests/, the smoke check and the bench CPU baseline
may import it, as the checker; the product path never does.

Follows feTimeIntegration.cpp:73-94 (time update), :383-437 (BDF2 started by
a BDF1 step over a tenth of the step, then a variable-step BDF2 step),
:449-551 (makeSteps of BDF1 / BDF2, the exact-solution starter at :484-500)
and feSolutionContainer.cpp:82-99 (BDF coefficients), :291-305 (rotation),
:338-348 (u_dot = sum_j cn[j] sol[j]).  The Newton loop is the one of
feNonLinearSolver.cpp:64-142 with a sparse direct solve.

Pinned in tests/test_oracle_time.py: the coefficients differentiate
polynomials of the scheme's order exactly on variable steps; the integrator
reproduces the reference's golden tables scalarFEunsteady_linearODE.output
(BDF1, and BDF2 with the exact-solution starter); with the BDF1 starter a
solution linear in time is integrated exactly.
"""
import numpy as np
import scipy.sparse.linalg as spla


def bdf_coefficients(order, dts):
    """cn of u_dot = sum_j cn[j] u(t_j), t_0 the new time; dts[0] = t_0 - t_1,
    dts[1] = t_1 - t_2."""
    if order == 0:
        return [0.0]
    if order == 1:
        return [1.0 / dts[0], -1.0 / dts[0]]
    h0, h1 = dts[0], dts[1]
    return [1.0 / h0 + 1.0 / (h0 + h1), -1.0 / h0 - 1.0 / h1, h0 / h1 / (h0 + h1)]


def integrate(assemble, u0, n, order, t0, t1, nsteps, start="BDF1", exact=None, essential=None,
              tol=1e-10, max_newton=10):
    """Integrates from t0 to t1 in nsteps steps of BDF order 1 or 2.

    assemble(u, udot, c0, t) -> (A, b): the Jacobian (n x n sparse) and the
    residual b = -NL(u) of the unknown rows; u holds nDofs values, the first n
    unknown.  essential(t) -> the nDofs - n essential values at t (None: kept).
    Returns (times, solutions) at every post-processing point (the initial
    condition first)."""
    u = np.array(u0, dtype=np.float64)
    hist = []          # hist[j - 1] = sol[j]
    state = {"t": t0, "d": [0.0, 0.0]}
    dt = (t1 - t0) / nsteps
    times, sols = [t0], [u.copy()]

    def advance(h):
        state["t"] += h
        state["d"] = [h, state["d"][0]]
        hist.insert(0, u.copy())
        del hist[4:]

    def newton(cn):
        t = state["t"]
        if essential is not None:
            u[n:] = essential(t)
        for it in range(max_newton + 1):
            udot = cn[0] * u
            for j in range(1, len(cn)):
                udot = udot + cn[j] * hist[j - 1]
            A, b = assemble(u, udot, cn[0], t)
            if it > 0 and np.abs(b).max(initial=0.0) <= tol:
                return
            du = spla.spsolve(A.tocsc(), b)
            u[:n] += du
            if np.abs(b).max(initial=0.0) <= tol or np.abs(du).max(initial=0.0) <= tol:
                return

    def step(h, k):
        advance(h)
        newton(bdf_coefficients(k, state["d"]))

    first = 0
    if order == 2:
        if start == "InitialCondition":
            advance(dt)
            u[:] = exact(state["t"])
        else:
            h1 = 0.1 * dt
            h2 = dt - h1
            step(h1, 1)
            advance(h2)
            newton(bdf_coefficients(2, [h2, h1]))
        times.append(state["t"])
        sols.append(u.copy())
        first = 1
    for _ in range(first, nsteps):
        step(dt, order)
        times.append(state["t"])
        sols.append(u.copy())
    return times, sols
