"""BDF time integration over the device linear system.

The reference's TimeIntegrator family (feTimeIntegration.cpp:73-551) with the
solution history of feSolutionContainer / BDFContainer
(feSolutionContainer.cpp:82-99,291-305,338-348).  The history vectors, the
time-derivative gather u_dot = sum_j cn[j] sol[j] and the TRANSIENT_MASS
element terms live on the device (fe_hip_set_bdf, fe_hip_rotate_solutions,
fe_hip_set_past_solution); this module only sequences the steps: time update,
rotation, BDF coefficients, essential values and time-dependent coefficients
at the new time, Newton (problems.solve_newton), post-processing.

Time-dependent callbacks (the reference's feFunctionArguments::t) are passed
as `coefficients`: (form, index, f) with f(t) returning the constant c0 of an
analytic coefficient or a per-element / per-quadrature-point table, written
with fe_hip_update_coeff before every Newton solve.  `essential(t)` returns
the nDofs - nInc essential values at time t (None: they keep their values);
`norms` are callables (S, t) -> float evaluated after every step and on the
initial condition, stored like getPostProcessingData (row 0: times).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import api as A
from .problems import NLSolverOptions, solve_newton

BDF2_START_BDF1 = "BDF1"                # BDF2Starter::BDF1 (the reference's default)
BDF2_START_EXACT = "InitialCondition"   # BDF2Starter::InitialCondition


def bdf_coefficients(order, dts):
    """computeBDFCoefficients (feSolutionContainer.cpp:82-99): dts[0] is the
    current step, dts[1] the previous one."""
    if order == 0:
        return [0.0]
    if order == 1:
        return [1.0 / dts[0], -1.0 / dts[0]]
    if order == 2:
        d0, d1 = dts[0], dts[1]
        return [1.0 / d0 + 1.0 / (d0 + d1), -1.0 / d0 - 1.0 / d1, d0 / d1 * 1.0 / (d0 + d1)]
    raise ValueError("BDF order must be 0, 1 or 2")


@dataclass
class StepReport:
    time: float
    newton_iterations: int
    linear_iterations: int
    asm_ms: float = 0.0    # device time of the element sweeps
    solve_ms: float = 0.0  # device time of the linear solves


@dataclass
class TimeIntegrator:
    """BDF1Integrator / BDF2Integrator (feTimeIntegration.cpp:198-551).

    S: the device linear system with its forms (one TRANSIENT_MASS per
    transient field); u0: the initial solution (nDofs values, t = t0);
    exact(t): the exact solution (nDofs values), needed only by the
    exact-solution BDF2 starter."""
    S: A.LinearSystemHIP
    order: int
    t0: float
    t1: float
    n_steps: int
    u0: np.ndarray
    nl: NLSolverOptions = field(default_factory=NLSolverOptions)
    coefficients: list = field(default_factory=list)
    essential: object = None
    norms: list = field(default_factory=list)
    exact: object = None
    start: str = BDF2_START_BDF1
    solve_kw: dict = field(default_factory=dict)
    fused_assembly: bool = True  # first Newton iteration: residual and matrix in one sweep

    def __post_init__(self):
        if self.order not in (1, 2):
            raise ValueError("order must be 1 (BDF1) or 2 (BDF2)")
        self.dt = (self.t1 - self.t0) / self.n_steps
        self.time = self.t0
        self.step = 0
        self.deltas = [0.0, 0.0]  # the integrator's _deltaT: current, previous
        self.post = [[] for _ in range(len(self.norms) + 1)]
        self.reports: list[StepReport] = []
        self.S.set_solution(self.u0)
        self.S.set_bdf([0.0])

    # -- pieces of makeSteps -------------------------------------------------
    def _at_time(self, t):
        """setCurrentTime + initializeEssentialBC + the callbacks' args.t."""
        for form, index, f in self.coefficients:
            self.S.update_coeff(form, index, f(t))
        if self.essential is not None:
            self.S.set_essential(self.essential(t))

    def _advance(self, dt):
        """updateTime + feSolutionContainer::rotate."""
        self.time += dt
        self.deltas = [dt, self.deltas[0]]
        self.S.rotate_solutions()

    def _newton(self):
        rep = solve_newton(self.S, self.nl, fused=self.fused_assembly, **self.solve_kw)
        if not rep.converged:
            raise A.FeError(f"Newton did not converge at t = {self.time:.6e}")
        self.reports.append(StepReport(self.time, rep.iterations, sum(r.iterations for r in rep.linear), rep.asm_ms,
                                       sum(r.ms_solve for r in rep.linear)))

    def _post_process(self):
        self.post[0].append(self.time)
        for i, nrm in enumerate(self.norms):
            self.post[i + 1].append(nrm(self.S, self.time))

    def _bdf_step(self, dt, order):
        self._advance(dt)
        self.S.set_bdf(bdf_coefficients(order, self.deltas))
        self._at_time(self.time)
        self._newton()

    def _start_bdf2(self):
        """The first BDF2 step (feTimeIntegration.cpp:383-437 / :484-500)."""
        if self.start == BDF2_START_EXACT:
            self._advance(self.dt)
            if self.exact is None:
                raise ValueError("the exact-solution starter needs exact(t)")
            u = self.exact(self.time)
            self.S.set_solution(u)
            self._at_time(self.time)
            return
        # BDF1 over a tenth of the step, then variable-step BDF2 to t0 + dt:
        # sol[2] = u(t0), sol[1] = u(t0 + dt / 10).
        r = 0.1
        d1 = r * self.dt
        d2 = self.dt - d1
        self._bdf_step(d1, 1)
        self._advance(d2)
        self.S.set_bdf(bdf_coefficients(2, [d2, d1]))
        self._at_time(self.time)
        self._newton()

    def make_steps(self, n=None):
        """makeSteps(nSteps): post-processing on the initial condition, then
        n steps of the scheme (BDF2: the starting step is the first)."""
        n = self.n_steps if n is None else n
        first = 0
        if self.step == 0:
            self._post_process()
            if self.order == 2:
                self._start_bdf2()
                self.step += 1
                self._post_process()
                first = 1
        for _ in range(first, n):
            self._bdf_step(self.dt, self.order)
            self.step += 1
            self._post_process()
        return self.post


def linf_nodal_error(num: A.Numbering, exact):
    """max over the DOFs of |u_h - u(x, t)| for a scalar exact(t, x, y, z)
    (the reference's LINF_ERROR, feNorm.cpp:1593-1616, samples quadrature
    points instead; both agree for solutions the space represents exactly in
    space)."""
    xyz = num.dof_xyz.reshape(-1, 3)

    def norm(S, t):
        u = S.get_solution()
        return float(np.max(np.abs(u - exact(t, xyz[:, 0], xyz[:, 1], xyz[:, 2]))))
    return norm
