"""Problem drivers over the GPU linear system.

`diffusion_system` wires the reference's scalar diffusion set-up
(createFiniteElementSpace x2, feMetaNumber, feSolution, Diffusion + Source
forms, createLinearSystem -- exe/example1.cpp:125-160) onto the device,
`heat_system` the transient one (TransientMass + Diffusion + Source), and
`solve_newton` is the Newton-Raphson loop of solveNewtonRaphson
(feNonLinearSolver.cpp:38-182) driving it.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import api as A


@dataclass
class NLSolverOptions:
    """feNLSolverOptions (feNonLinearSolver.h:12-40)."""
    tolResidual: float = 1e-10
    tolCorrection: float = 1e-10
    tolDivergence: float = 1e4
    maxIter: int = 20


@dataclass
class NewtonReport:
    iterations: int = 0
    residual_norms: list = field(default_factory=list)
    correction_norms: list = field(default_factory=list)
    linear: list = field(default_factory=list)
    asm_ms: float = 0.0  # device time of the element sweeps of all iterations
    converged: bool = False


def diffusion_system(mesh: A.Mesh, kind: int, degree: int, k: A.Fn, s: A.Fn, u_bc=None, coloring=None,
                     device=0, dk: A.Fn | None = None):
    """Scalar -div(k grad u) + S = 0 with essential boundary, on the GPU.
    u_bc(x, y, z, comp) gives the Dirichlet values (None: 0).  dk given: the
    conductivity depends on the solution, k(u) with derivative dk/du (e.g.
    two FN_POLY_U), through the NonlinearDiffusion form of the reference's
    nonlinear test (scalarFE.cpp:128-150)."""
    num = A.Numbering(mesh, [kind], [1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    if dk is None:
        S.add_form(A.DIFFUSION, [f], degree, [k])
    else:
        S.add_form(A.NONLINEAR_DIFFUSION, [f], degree, [k, dk])
    S.add_form(A.SOURCE, [f], degree, [s])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc]))
    return S, num


def advection_diffusion_system(mesh: A.Mesh, kind: int, degree: int, velocity, k: A.Fn, s: A.Fn, reaction=None,
                               u_bc=None, coloring=None, device=0):
    """Steady advection-diffusion(-reaction) with essential boundary: the form
    stack of the reference's scalarFE tests (scalarFE.cpp:241-246 and
    :345-353): Advection (velocity components), Diffusion (k), Mass
    (reaction coefficient, optional), Source."""
    num = A.Numbering(mesh, [kind], [1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    S.add_form(A.ADVECTION, [f], degree, list(velocity))
    S.add_form(A.DIFFUSION, [f], degree, [k])
    if reaction is not None:
        S.add_form(A.MASS, [f], degree, [reaction])
    S.add_form(A.SOURCE, [f], degree, [s])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc]))
    return S, num


def mixed_diffusion_system(mesh: A.Mesh, kind: int, degree: int, su: A.Fn, sv: A.Fn, u_bc=None, v_bc=None,
                           coloring=None, device=0):
    """Two coupled scalar fields u, v with essential boundaries,
    u + v^3 - lap v + su = 0 and u + v - lap u + sv = 0: the form stack of the
    reference's mixed diffusion test (scalarFE.cpp:455-475) -- Mass {u},
    MixedMassPower {u, v} (1, 3), MixedGradGrad {u, v} (-1), Source {u},
    Mass {v}, MixedMass {v, u}, MixedGradGrad {v, u} (-1), Source {v}."""
    num = A.Numbering(mesh, [kind, kind], [1, 1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    fu = S.add_field(kind, num.adr[0])
    fv = S.add_field(kind, num.adr[1])
    S.set_sizes(num.nInc, num.nDofs)
    one, minus_one = A.Fn(A.FN_CONST, 1.0), A.Fn(A.FN_CONST, -1.0)
    S.add_form(A.MASS, [fu], degree, [one])
    S.add_form(A.MIXED_MASS_POWER, [fu, fv], degree, [one, A.Fn(A.FN_CONST, 3.0)])
    S.add_form(A.MIXED_GRADGRAD, [fu, fv], degree, [minus_one])
    S.add_form(A.SOURCE, [fu], degree, [su])
    S.add_form(A.MASS, [fv], degree, [one])
    S.add_form(A.MIXED_MASS, [fv, fu], degree, [one])
    S.add_form(A.MIXED_GRADGRAD, [fv, fu], degree, [minus_one])
    S.add_form(A.SOURCE, [fv], degree, [sv])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc, v_bc]))
    return S, num


def heat_system(mesh: A.Mesh, kind: int, degree: int, m: A.Fn, k: A.Fn, s: A.Fn, essential=True, coloring=None,
                device=0, periodic=None, react=None):
    """Transient scalar m u_t - div(k grad u) + S = 0: the form stack of the
    reference's unsteady scalar tests (TransientMass, Diffusion, Source --
    scalarFEunsteady.cpp:270-300), on the GPU.  essential: Dirichlet boundary
    (True), none (False) or the names of its boundary groups; periodic:
    (master, slave) DOF pairs (Numbering.periodic_dofs) entering the
    pattern.  k = None leaves out the diffusion form: with no
    essential boundary that is the ODE m u_t + S = 0 of
    scalarFEunsteady.cpp:54-80.  react = (c, p): a MassPower form c u^p (with
    m u_t alone, the nonlinear ODE of scalarFEunsteady.cpp:134-165).  Returns (S, num, forms) with forms =
    {"mass", "diffusion", "source"} ids for fe_hip_update_coeff."""
    ess = list(essential) if isinstance(essential, (list, tuple)) else (1 if essential else 0)
    num = A.Numbering(mesh, [kind], [ess])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    forms = {"mass": S.add_form(A.TRANSIENT_MASS, [f], degree, [m])}
    if react is not None:
        forms["react"] = S.add_form(A.MASS, [f], degree, [react[0], A.Fn(A.FN_CONST, float(react[1]))])
    if k is not None:
        forms["diffusion"] = S.add_form(A.DIFFUSION, [f], degree, [k])
    if s is not None:
        forms["source"] = S.add_form(A.SOURCE, [f], degree, [s])
    if periodic is not None:
        S.set_periodic_dofs(*periodic)
    S.build_csr()
    return S, num, forms


def solve_newton(S: A.LinearSystemHIP, opts: NLSolverOptions = NLSolverOptions(), progress=None,
                 fused=False, **solve_kw) -> NewtonReport:
    """solveNewtonRaphson (feNonLinearSolver.cpp:64-142): residual, stopping
    test, matrix, solve, u += du, until ||NL(u)||_inf <= tolResidual or
    ||du||_inf <= tolCorrection.  progress(iteration, SolveResult) is called
    after every linear solve (the reference prints the same per-step norms).
    fused: the first iteration, which always needs the matrix, assembles
    residual and matrix in one element sweep per form (same values)."""
    rep = NewtonReport()
    it = 0
    prec = solve_kw.get("precision", A.PREC_FP64)
    while True:
        S.setToZero()
        both = fused and it == 0
        if both:
            S.assemble(3, prec)
        else:
            S.assembleResiduals()
        nres = S.getRHSMaxNorm()
        rep.asm_ms += S.kernel_stats(1)[0]  # read after the norm: no extra wait
        if it > 0 and nres <= opts.tolResidual:
            rep.residual_norms.append(nres)
            rep.converged = True
            break
        if not both:
            S.assembleMatrices(prec)
        S.constrainEssentialComponents()  # no-ops without constraints (feNonLinearSolver.cpp:94-95)
        S.applyPeriodicity()
        r = S.solve(**solve_kw)
        if not both:
            rep.asm_ms += S.kernel_stats(1)[0]
        rep.linear.append(r)
        if not r.converged and r.reason < 0 and r.reason != -3:
            raise A.FeError(f"linear solve failed at Newton iteration {it + 1}: reason {r.reason}")
        if r.norm_rhs_inf > opts.tolDivergence:
            raise A.FeError(f"residual norm {r.norm_rhs_inf:.4e} above the divergence tolerance")
        S.correctSolution()
        it += 1
        if progress is not None:
            progress(it, r)
        rep.residual_norms.append(r.norm_rhs_inf)
        rep.correction_norms.append(r.norm_dx_inf)
        if r.norm_rhs_inf <= opts.tolResidual or r.norm_dx_inf <= opts.tolCorrection or it > opts.maxIter:
            rep.converged = r.norm_rhs_inf <= opts.tolResidual or r.norm_dx_inf <= opts.tolCorrection
            break
    rep.iterations = it
    return rep


def taylor_hood_system(mesh: A.Mesh, degree: int, mu: float = 1.0, convection=None, source=None, u_bc=None,
                       p_bc=None, coloring=None, device=0, laplacian=False):
    """Taylor-Hood P2-P1 Stokes or Navier-Stokes with Dirichlet velocity and
    pressure on the boundary: the form stacks of the reference's MMS tests
    (stokes.cpp:76-85, divergence formulation: MixedDivergence {p, u} (1),
    VectorSource, DivergenceNewtonianStress {u, p} (1, mu); navier_stokes.cpp
    adds VectorConvectiveAcceleration first with coefficient `convection`).
    source: the two Fn components of the vector source (None: zero);
    u_bc(x, y, z, comp), p_bc(x, y, z, comp): the essential values.
    laplacian=True: the Laplacian formulation (stokes.cpp:80-84) -- VectorDiffusion
    (-1, mu) and MixedGradient {u, p} (-1) instead of the stress form."""
    num = A.Numbering(mesh, [A.TRI_VP2, A.TRI_P1], [1, 1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    fu = S.add_field(A.TRI_VP2, num.adr[0])
    fp = S.add_field(A.TRI_P1, num.adr[1])
    S.set_sizes(num.nInc, num.nDofs)
    if convection is not None:
        S.add_form(A.VECTOR_CONVECTIVE_ACCELERATION, [fu], degree, [A.Fn(A.FN_CONST, float(convection))])
    S.add_form(A.MIXED_DIVERGENCE, [fp, fu], degree, [A.Fn(A.FN_CONST, 1.0)])
    S.add_form(A.VECTOR_SOURCE, [fu], degree, source or [A.Fn(A.FN_CONST, 0.0), A.Fn(A.FN_CONST, 0.0)])
    if laplacian:
        S.add_form(A.VECTOR_DIFFUSION, [fu], degree, [A.Fn(A.FN_CONST, -1.0), A.Fn(A.FN_CONST, float(mu))])
        S.add_form(A.MIXED_GRADIENT, [fu, fp], degree, [A.Fn(A.FN_CONST, -1.0)])
    else:
        S.add_form(A.DIV_NEWTONIAN_STRESS, [fu, fp], degree, [A.Fn(A.FN_CONST, 1.0), A.Fn(A.FN_CONST, float(mu))])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc, p_bc]))
    return S, num


def weak_gradient_system(mesh: A.Mesh, degree: int, source: A.Fn, u_bc=None, coloring=None, device=0):
    """The weak gradient u = grad f of a scalar f for the vector P2 field with
    strong Dirichlet values (vectorFE.cpp:19-90): VectorMass (1) and
    GradSource (f)."""
    num = A.Numbering(mesh, [A.TRI_VP2], [1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    fu = S.add_field(A.TRI_VP2, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    S.add_form(A.VECTOR_MASS, [fu], degree, [A.Fn(A.FN_CONST, 1.0)])
    S.add_form(A.GRAD_SOURCE, [fu], degree, [source])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc]))
    return S, num


def vector_laplacian_system(mesh: A.Mesh, degree: int, source, u_bc=None, coloring=None, device=0):
    """-div(grad u) = f for the vector P2 field with strong Dirichlet values
    (vectorFE.cpp:246-259): VectorDiffusion (1, 1) and VectorSource."""
    num = A.Numbering(mesh, [A.TRI_VP2], [1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    fu = S.add_field(A.TRI_VP2, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    S.add_form(A.VECTOR_DIFFUSION, [fu], degree, [A.Fn(A.FN_CONST, 1.0), A.Fn(A.FN_CONST, 1.0)])
    S.add_form(A.VECTOR_SOURCE, [fu], degree, list(source))
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc]))
    return S, num


def kovasznay(Re):
    """Kovasznay flow: lambda = Re/2 - sqrt(Re^2/4 + 4 pi^2), u = 1 - e^{lx}
    cos 2 pi y, v = l/(2 pi) e^{lx} sin 2 pi y, p = (1 - e^{2 l x}) / 2."""
    lam = Re / 2.0 - np.sqrt(Re * Re / 4.0 + 4.0 * np.pi ** 2)

    def u(x, y, z, comp):
        ex = np.exp(lam * x)
        return np.where(comp == 0, 1.0 - ex * np.cos(2 * np.pi * y), lam / (2 * np.pi) * ex * np.sin(2 * np.pi * y))

    def p(x, y, z, comp):
        return 0.5 * (1.0 - np.exp(2.0 * lam * x))

    return lam, u, p


def navier_stokes_system(mesh: A.Mesh, Re: float, degree: int = 5, coloring=None, device=0, source=None,
                         mass=None, essential=True, periodic=None):
    """Steady incompressible Navier-Stokes, Taylor-Hood P2-P1, divergence
    formulation (the reference's navier_stokes.cpp:66-99 form stack):
    convective acceleration (coeff -1), mixed divergence {p, u} (1), vector
    source, divergence of the Newtonian stress {u, p} (1, mu = 1/Re).  u and p
    are essential on the boundary with the Kovasznay values; the interior
    starts from zero.  mass (an Fn): the unsteady equations, with a
    TransientVectorMass form first (vectorFEunsteady.cpp:561).  The momentum
    rows of this stack carry -(u . grad u - div sigma), so a well-posed
    unsteady problem takes the mass coefficient -1 (the reference's unsteady
    stack flips the convection and stress coefficients instead,
    vectorFEunsteady.cpp:560-572).  essential=False: no Dirichlet boundary
    (constraints through set_constrained_dofs instead); periodic: (master,
    slave) DOF pairs entering the pattern."""
    num = A.Numbering(mesh, [A.TRI_VP2, A.TRI_P1], [1, 1] if essential else [0, 0])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    fu = S.add_field(A.TRI_VP2, num.adr[0])
    fp = S.add_field(A.TRI_P1, num.adr[1])
    S.set_sizes(num.nInc, num.nDofs)
    if mass is not None:
        S.add_form(A.TRANSIENT_VECTOR_MASS, [fu], degree, [mass])
    S.add_form(A.VECTOR_CONVECTIVE_ACCELERATION, [fu], degree, [A.Fn(A.FN_CONST, -1.0)])
    S.add_form(A.MIXED_DIVERGENCE, [fp, fu], degree, [A.Fn(A.FN_CONST, 1.0)])
    src = source or [A.Fn(A.FN_CONST, 0.0), A.Fn(A.FN_CONST, 0.0)]
    S.add_form(A.VECTOR_SOURCE, [fu], degree, src)
    S.add_form(A.DIV_NEWTONIAN_STRESS, [fu, fp], degree, [A.Fn(A.FN_CONST, 1.0), A.Fn(A.FN_CONST, 1.0 / Re)])
    if periodic is not None:
        S.set_periodic_dofs(*periodic)
    S.build_csr()
    _, ufn, pfn = kovasznay(Re)
    S.set_solution(A.essential_values(num, [ufn, pfn]))
    return S, num
