"""Problem drivers over the GPU linear system.

`diffusion_system` wires the reference's scalar diffusion set-up
(createFiniteElementSpace x2, feMetaNumber, feSolution, Diffusion + Source
forms, createLinearSystem -- exe/example1.cpp:125-160) onto the device, and
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
    converged: bool = False


def diffusion_system(mesh: A.Mesh, kind: int, degree: int, k: A.Fn, s: A.Fn, u_bc=None, coloring=None,
                     device=0):
    """Scalar -div(k grad u) + S = 0 with essential boundary, on the GPU.
    u_bc(x, y, z, comp) gives the Dirichlet values (None: 0)."""
    num = A.Numbering(mesh, [kind], [1])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(mesh, coloring)
    f = S.add_field(kind, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    S.add_form(A.DIFFUSION, [f], degree, [k])
    S.add_form(A.SOURCE, [f], degree, [s])
    S.build_csr()
    S.set_solution(A.essential_values(num, [u_bc]))
    return S, num


def solve_newton(S: A.LinearSystemHIP, opts: NLSolverOptions = NLSolverOptions(), **solve_kw) -> NewtonReport:
    """solveNewtonRaphson (feNonLinearSolver.cpp:64-142): residual, stopping
    test, matrix, solve, u += du, until ||NL(u)||_inf <= tolResidual or
    ||du||_inf <= tolCorrection."""
    rep = NewtonReport()
    it = 0
    while True:
        S.setToZero()
        S.assembleResiduals()
        nres = S.getRHSMaxNorm()
        if it > 0 and nres <= opts.tolResidual:
            rep.residual_norms.append(nres)
            rep.converged = True
            break
        S.assembleMatrices(solve_kw.get("precision", A.PREC_FP64))
        r = S.solve(**solve_kw)
        rep.linear.append(r)
        if not r.converged and r.reason < 0 and r.reason != -3:
            raise A.FeError(f"linear solve failed at Newton iteration {it + 1}: reason {r.reason}")
        if r.norm_rhs_inf > opts.tolDivergence:
            raise A.FeError(f"residual norm {r.norm_rhs_inf:.4e} above the divergence tolerance")
        S.correctSolution()
        it += 1
        rep.residual_norms.append(r.norm_rhs_inf)
        rep.correction_norms.append(r.norm_dx_inf)
        if r.norm_rhs_inf <= opts.tolResidual or r.norm_dx_inf <= opts.tolCorrection or it > opts.maxIter:
            rep.converged = r.norm_rhs_inf <= opts.tolResidual or r.norm_dx_inf <= opts.tolCorrection
            break
    rep.iterations = it
    return rep
