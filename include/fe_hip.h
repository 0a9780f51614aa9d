/*
 * fe_hip.h -- C ABI of libfe_hip.so: the MI355X (gfx950) assemble-and-solve
 * path of feNG.  One context = one GPU-resident linear system: mesh, element
 * -> DOF maps, device-built CSR, the (bi-)linear forms, the solution vector
 * and the Krylov workspace.
 *
 * The entry points map one-to-one onto the reference's feLinearSystem
 * virtuals (feLinearSystem.h:85-168) and its createLinearSystem factory
 * (feLinearSystem.h:33-37); the element kernels replace the feSysElm
 * computeAe/computeBe virtuals (feSysElm.h:187-251) behind feBilinearForm.
 *
 * Conventions: every function returns an int feStatus (feMessage.h:44-62,
 * 0 = OK, negative = error); no exception crosses the ABI; host pointers are
 * borrowed for the duration of the call; device work is ordered on the
 * context's stream and a call returns after its work has completed (unless
 * stated otherwise).  One host thread per context.
 */
#ifndef FE_HIP_H
#define FE_HIP_H

#include <stdint.h>

#include "fe_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_hip_ctx fe_hip_ctx;

/* Weak-form ids: the reference's elementSystemType values (feSysElm.h:13-77)
 * for the terms the device kernels implement. */
#define FE_SYS_SOURCE 0
#define FE_SYS_VECTOR_SOURCE 2
#define FE_SYS_GRAD_SOURCE 4   /* vector field: Be += S div(phi_i) */
#define FE_SYS_MASS 6 /* coefficients (c): feSysElm_Mass; (c, FE_FN_CONST p): feSysElm_MassPower c u^p */
/* Mixed scalar forms, fields {test field, trial field} of one scalar kind: */
#define FE_SYS_MIXED_MASS 8       /* (c): Be -= c u phi_i */
#define FE_SYS_MIXED_MASS_POWER 9 /* (c, FE_FN_CONST p): Be -= c u^p phi_i */
#define FE_SYS_VECTOR_MASS 10
#define FE_SYS_TRANSIENT_MASS 13
#define FE_SYS_TRANSIENT_VECTOR_MASS 15
#define FE_SYS_DIFFUSION 16
#define FE_SYS_ADVECTION 20 /* coefficients: the dim velocity components */
#define FE_SYS_VECTOR_DIFFUSION 18
#define FE_SYS_NONLINEAR_DIFFUSION 19 /* coefficients (k(u), dk/du), e.g. two FE_FN_POLY_U */
#define FE_SYS_VECTOR_CONVECTIVE_ACCELERATION 22
#define FE_SYS_DIV_NEWTONIAN_STRESS 25
#define FE_SYS_MIXED_GRADIENT 26
#define FE_SYS_MIXED_GRADGRAD 28 /* {test field, trial field}, (c): Be -= c grad(u).grad(phi_i) */
#define FE_SYS_MIXED_DIVERGENCE 31

/* Coefficient kinds: the feFunction callbacks (feFunction.h:75-110) the
 * device evaluates at quadrature points, or caller tables. */
#define FE_FN_CONST 0   /* c0 */
#define FE_FN_SINSIN 1  /* c0 sin(pi x) sin(pi y) */
#define FE_FN_BLOCKS 2  /* exp(0.5(((7 bi + 13 bj) mod 11) - 5)) on c0 x c0 blocks */
#define FE_FN_SMOOTH3 3 /* exp(c0 sin(2 pi x) sin(2 pi y) sin(2 pi z)) */
#define FE_FN_ELEM 4    /* data[e] (count = E) */
#define FE_FN_QP 5      /* data[e nq + q] (count = E nq) */
#define FE_FN_KOV_U 6   /* Kovasznay velocity component c1 (0 or 1), Re = c0 */
#define FE_FN_KOV_P 7   /* Kovasznay pressure, Re = c0 */
#define FE_FN_POLY_U 8  /* sum_i data[i] u^i, u the solution at the point (count <= 8; NONLINEAR_DIFFUSION only) */

typedef struct {
  int kind;
  double c0, c1;
  const double *data; /* host table for FE_FN_ELEM / FE_FN_QP / FE_FN_POLY_U, copied at add time */
  int64_t count;
} fe_hip_fn;

/* Krylov method / preconditioner / precision (the KSP and PC types the
 * reference selects through PETSc options, feLinearSystemPETSc.cpp:900-1031). */
#define FE_KSP_CG 0
#define FE_KSP_GMRES 1
/* Direct solve: LU with partial pivoting of the dense copy of the system on
 * the device (rocSOLVER getrf / getrs), up to 46000 unknowns; the role of
 * the reference's direct back ends (MKL Pardiso,
 * feLinearSystemMklPardiso.cpp:1154-1222; PETSc -pc_type lu with MUMPS,
 * tests/withLinearSolver/navier_stokes.cpp:101-108).  pc, precision and the
 * iteration options are ignored; the FP64 true residual is checked against
 * rtol / atol; iterations = 1. */
#define FE_KSP_DIRECT 2

#define FE_PC_NONE 0
#define FE_PC_JACOBI 1
#define FE_PC_IC0 2
#define FE_PC_ILU0 3
#define FE_PC_BJ_ILU0 4
#define FE_PC_LEARNED 5
#define FE_PC_AMG 6 /* aggregation multigrid V-cycle (PETSc -pc_type gamg role) */
/* GMRES on saddle-point systems (a field whose diagonal block is zero, e.g.
 * the pressure of Taylor-Hood Navier-Stokes): block upper-triangular Schur
 * preconditioner, the role of PETSc's -pc_type fieldsplit
 * -pc_fieldsplit_type schur -pc_fieldsplit_schur_fact_type upper
 * -pc_fieldsplit_schur_precondition selfp; the other block by one
 * aggregation-multigrid V-cycle on its same-component entries, the Schur
 * complement by the diagonal of A_pu diag(A_uu)^-1 A_up */
#define FE_PC_FIELDSPLIT 7

#define FE_PREC_FP64 0    /* FP64 Krylov */
#define FE_PREC_FP32_IR 1 /* FP32 inner Krylov + FP64 iterative refinement */

typedef struct {
  int ksp;           /* FE_KSP_* */
  int pc;            /* FE_PC_* */
  int precision;     /* FE_PREC_* */
  int restart;       /* GMRES restart length */
  int max_iter;      /* total inner iterations (reference default 1e4) */
  double rtol;       /* ||b - A x||_2 / ||b||_2 target, checked in FP64 */
  double atol;       /* absolute ||b - A x||_2 target (reference default 1e-14) */
  double dtol;       /* divergence tolerance (reference default 1e6) */
  double inner_rtol; /* FP32_IR: relative reduction per inner solve */
  int check_every;   /* iterations between host convergence checks */
} fe_hip_solver_opts;

typedef struct {
  int iterations;       /* total Krylov iterations */
  int outer_iterations; /* refinement steps (FP32_IR) */
  int converged;        /* 1 when the FP64 true residual met the target */
  int reason;           /* >0 converged, <0 diverged (PETSc KSPConvergedReason style) */
  double relres_fp64;   /* ||b - A x||_2 / ||b||_2 in FP64 */
  double norm_dx_inf;   /* ||du||_inf */
  double norm_rhs_inf;  /* ||b||_inf */
  double norm_axb_inf;  /* ||A du - b||_inf */
  double ms_solve;      /* device time of the solve (hipEvents) */
  double ms_setup;      /* preconditioner setup included in ms_solve */
} fe_hip_solve_stats;

/* ---- context ------------------------------------------------------------ */

/* createLinearSystem (feLinearSystem.cpp:7-67) counterpart: a context on HIP
 * device `device`.  Fails with FE_STATUS_ERROR when no GPU is present. */
int fe_hip_ctx_create(int device, fe_hip_ctx **out);
int fe_hip_ctx_destroy(fe_hip_ctx *ctx);
/* Last error message of this context (static storage owned by the context). */
const char *fe_hip_last_error(const fe_hip_ctx *ctx);

/* Mesh geometry (feMesh getCoord, feMesh.cpp:119-134): dim 2 triangles or
 * dim 3 tets, straight (P1) geometry. */
int fe_hip_upload_mesh(fe_hip_ctx *ctx, int dim, int V, const double *xyz, int E, int nvpe,
                       const int *conn);
/* Element coloring (feCncGeo.cpp:735-794 semantics): colorOf[E]; elements of
 * one color share no vertex, so each color is one atomic-free scatter sweep. */
int fe_hip_set_coloring(fe_hip_ctx *ctx, const int *colorOf);
/* One field (an feSpace): kind FE_TRI_* / FE_TET_P1, element -> DOF map
 * adr[E nloc] from feMetaNumber (initializeAddressingVector order). Returns
 * the field index in *field. */
int fe_hip_add_field(fe_hip_ctx *ctx, int kind, const int *adr, int *field);
/* System size n = nInc and total DOFs (feMetaNumber::getNbUnknowns /
 * getNbDOFs, feNumber.h:202-203). */
int fe_hip_set_sizes(fe_hip_ctx *ctx, int nInc, int nDofs);

/* One (bi-)linear form (createBilinearForm, feBilinearForm.cpp:6-73):
 * sysType FE_SYS_*, the fields it acts on (same order as the reference's
 * space vector), quadrature degree, up to 4 coefficient callbacks in the
 * order the reference's feSysElm constructor takes them. Returns the form id. */
int fe_hip_add_form(fe_hip_ctx *ctx, int sysType, int nfields, const int *fields,
                    int quad_degree, int ncoeff, const fe_hip_fn *coeffs, int *form);

/* feEZCompressedRowStorage (feCompressedRowStorage.cpp:15-133) built on the
 * device from the element -> DOF maps of the matrix forms: ia[n+1], ja[nnz],
 * rows sorted, diagonal always present, only unknown rows and columns. */
int fe_hip_build_csr(fe_hip_ctx *ctx, int64_t *nnz);
int fe_hip_get_csr(fe_hip_ctx *ctx, int *ia, int *ja);

/* ---- solution, assembly ------------------------------------------------- */

/* feSolution values (nDofs doubles, essential DOFs holding their BC values). */
int fe_hip_set_solution(fe_hip_ctx *ctx, const double *u);
int fe_hip_get_solution(fe_hip_ctx *ctx, double *u);

/* setToZero / setMatrixToZero / setResidualToZero (feLinearSystem.h:112-114). */
int fe_hip_set_to_zero(fe_hip_ctx *ctx, int matrix, int residual);

/* assembleMatrices / assembleResiduals / assemble (feLinearSystemPETSc.cpp:
 * 543-805): every form's element kernel on every element, adding into the
 * CSR values and the RHS in the summation order of the reference's colored
 * element loop.  Fused scalar sweeps (DIFFUSION, TRANSIENT_MASS, SOURCE) of
 * P1/P2 triangles and P1 tets run as one row-gather launch (each CSR row
 * sums its incident elements' rows in color order; tuning key asm_kernel =
 * 2, the default); the other forms as one colored, atomic-free sweep per
 * color.  what: 1 = matrix, 2 = residual, 3 = both (fused: Ae and Be
 * together).  precision: FE_PREC_FP64 computes in FP64; FE_PREC_FP32_IR
 * additionally keeps an FP32 copy of the values for the inner solve (the
 * FP32 rounding of the FP64 values, or with asm_fp32_math = 1 element rows
 * computed and summed in FP32). */
int fe_hip_assemble(fe_hip_ctx *ctx, int what, int precision);

/* getRHSMaxNorm (feLinearSystem.h:104). */
int fe_hip_rhs_max_norm(fe_hip_ctx *ctx, double *norm);

/* Copy back the assembled system: val[nnz] (may be NULL), rhs[n] (may be NULL). */
int fe_hip_get_system(fe_hip_ctx *ctx, double *val, double *rhs);
/* Copy back the FP32 values val32[nnz] that an FE_PREC_FP32_IR assembly
 * keeps for the inner solve (no reference counterpart: the reference
 * assembles in FP64 only).  FE_STATUS_ERROR before any FP32 assembly. */
int fe_hip_get_system32(fe_hip_ctx *ctx, float *val32);
/* Overwrite the system values / RHS from the host (used for externally
 * assembled or perturbed systems). */
int fe_hip_set_system(fe_hip_ctx *ctx, const double *val, const double *rhs);

/* ---- solve -------------------------------------------------------------- */

void fe_hip_default_opts(fe_hip_solver_opts *opts);

/* solve (feLinearSystemPETSc.cpp:900-1031): du = A^-1 rhs with the Krylov
 * method and preconditioner of opts, x0 = 0, stopped on the FP64 true
 * residual.  du (n doubles, may be NULL) receives the correction. */
int fe_hip_solve(fe_hip_ctx *ctx, const fe_hip_solver_opts *opts, double *du,
                 fe_hip_solve_stats *stats);

/* correctSolution (feLinearSystemPETSc.cpp:1033-1060): u[i] += du[i], i < n. */
int fe_hip_correct_solution(fe_hip_ctx *ctx);

/* ---- constraints on the assembled system ------------------------------------
 * The linear-system hooks Newton calls after the matrix assembly
 * (feNonLinearSolver.cpp:94-95).
 * fe_hip_set_constrained_dofs: the DOFs of the essential components of vector
 * spaces (feSpace::setEssentialComponent; the rows the reference collects in
 * feLinearSystemMklPardiso.cpp:1000-1037); DOFs >= nInc are skipped as there.
 * fe_hip_constrain_essential_components: constrainEssentialComponents
 * (feLinearSystemPETSc.cpp:816-853 MatZeroRowsColumns with zero values,
 * feLinearSystemMklPardiso.cpp:994-1097): columns and rows of those DOFs
 * zeroed, 1 on their diagonal, 0 in their RHS (the FP32 value copy too).
 * fe_hip_set_periodic_dofs: the numbering's periodic (master, slave) pairs
 * (feMetaNumber::PeriodicDOF, one slave per master), before
 * fe_hip_build_csr: each slave row of two unknowns gains the master column
 * (feCompressedRowStorage.cpp:96-108).
 * fe_hip_apply_periodicity: applyPeriodicity
 * (feLinearSystemMklPardiso.cpp:1099-1127): slave row = 1 at the slave,
 * -1 at the master, 0 elsewhere, RHS 0.  Both apply hooks are no-ops
 * without constraints. */
int fe_hip_set_constrained_dofs(fe_hip_ctx *ctx, int count, const int *dofs);
int fe_hip_constrain_essential_components(fe_hip_ctx *ctx);
int fe_hip_set_periodic_dofs(fe_hip_ctx *ctx, int count, const int *master, const int *slave);
int fe_hip_apply_periodicity(fe_hip_ctx *ctx);

/* ILU(0)/IC(0) factor diagnostics: number of row levels of the factorization
 * schedule and zero pivots shifted in the last factorization. */
int fe_hip_pc_info(fe_hip_ctx *ctx, int *levels, int *shifts);
/* Copy back the last ILU(0) factor (L strict lower, unit diagonal implied,
 * and U, stored on the CSR pattern of A: nnz doubles). */
int fe_hip_get_factor(fe_hip_ctx *ctx, double *lu);

/* ---- measurement -------------------------------------------------------- */

/* Average device time (ms, hipEvents on the context stream) and launch count
 * of the last `fe_hip_solve`'s dominant kernel (which = 0: the iteration
 * kernel of the Jacobi, IC(0) and learned paths, the level-0 restriction
 * `pcgDown0` of the nested multigrid PCG), of the last assembly (which = 1),
 * or of one whole nested multigrid PCG iteration with all its launches
 * (which = 2).  Kernel and iteration times are taken with profiling on. */
int fe_hip_kernel_stats(fe_hip_ctx *ctx, int which, double *avg_ms, int64_t *launches);
/* Enable per-launch event timing of the dominant kernel (adds two events per
 * launch). */
int fe_hip_set_profiling(fe_hip_ctx *ctx, int on);

/* Launch shapes of the Krylov iteration kernels (no reference counterpart:
 * the reference delegates the solve to PETSc).  Keys: "spmv_lanes" (lanes
 * per CSR row: 4, 8, 16, 32), "spmv_unroll" (row groups in flight per lane
 * group: 1, 2), "iter_block" (256, 512, 1024), "iter_grid" (blocks, a
 * multiple of 8 up to 4096); FE_PC_AMG: "amg_omega_x100" (l1-Jacobi factor
 * times 100, default 180), "amg_oc_x100" (coarse over-correction times 100,
 * default 180), "amg_block_rows" (levels up to this many rows run in one
 * workgroup, default 1024). */
int fe_hip_set_tuning(fe_hip_ctx *ctx, const char *key, int value);
int fe_hip_get_tuning(fe_hip_ctx *ctx, const char *key, int *value);

/* Replace the values of coefficient `index` of form `form` (an FE_FN_ELEM or
 * FE_FN_QP table, same length as given to fe_hip_add_form).  Replaces the
 * reference's per-instance feFunction objects (feFunction.h) when many
 * coefficient fields are assembled on one mesh.  For the analytic kinds
 * count = 1 replaces c0: a callback of time only (feFunctionArguments::t,
 * feFunction.h) is set this way before every time step. */
int fe_hip_update_coeff(fe_hip_ctx *ctx, int form, int index, const double *data, int64_t count);

/* ---- time integration (BDF) ---------------------------------------------
 * The solution history of feSolutionContainer / BDFContainer
 * (feSolutionContainer.cpp:82-99,291-305,338-348) kept on the device: sol[0]
 * is the current solution (fe_hip_set_solution), sol[1..4] are past
 * solutions.  FE_SYS_TRANSIENT_MASS (feSysElm.cpp:426-458) assembles
 * Ae += phi_i m c0 phi_j and Be -= phi_i m u_dot with
 * u_dot = sum_{j < nsol} cn[j] sol[j] and c0 = cn[0], gathered per element
 * from the history vectors; FE_SYS_TRANSIENT_VECTOR_MASS
 * (feVectorSysElm.cpp:391-425) the same with phi_i . phi_j on the vector P2
 * field, fused into the Taylor-Hood element sweep.
 *
 * fe_hip_set_bdf: setC0 + computeBDFCoefficients; nsol in [1, 5] (the
 * reference's order 0 is nsol = 1, cn = {0}).
 * fe_hip_rotate_solutions: feSolutionContainer::rotate: sol[j] = sol[j - 1]
 * for j = 4..1; sol[0] keeps its values (Newton's initial guess).
 * fe_hip_set_past_solution: sol[j] (1 <= j <= 4) from nDofs host values
 * (initial conditions, the exact-solution BDF2 starter, restarts);
 * fe_hip_get_past_solution copies it back.
 * fe_hip_set_essential: the essential part of sol[0] (nDofs - nInc values,
 * feSolution::initializeEssentialBC at the new time). */
int fe_hip_set_bdf(fe_hip_ctx *ctx, int nsol, const double *cn);
int fe_hip_rotate_solutions(fe_hip_ctx *ctx);
int fe_hip_set_past_solution(fe_hip_ctx *ctx, int j, const double *u);
int fe_hip_get_past_solution(fe_hip_ctx *ctx, int j, double *u);
int fe_hip_set_essential(fe_hip_ctx *ctx, const double *values);

/* ---- norms ---------------------------------------------------------------
 * feNorm::compute (feNorm.cpp:124-246) on the device: norm `type` (the
 * reference's normType values, feNorm.h:10-85) of field `field` of the
 * current solution with the quadrature rule of degree quad_degree.  exact
 * holds the exact-solution callbacks the type needs (the reference's scalar
 * or vector feFunction; quadrature-point tables are E nq of this degree):
 *   L1/L2/LINF/SEMI_H1/H1/AREA, VECTOR_L2: none;
 *   L1_ERROR, L2_ERROR, LINF_ERROR: u; LINF_UDOT_ERROR: du/dt (the time
 *   derivative of the solution from the BDF history, fe_hip_set_bdf);
 *   VECTOR_L2_ERROR, VECTOR_LINF_ERROR, VECTOR_LINF_UDOT_ERROR: the two
 *   components of the vector P2 field;
 *   SEMI_H1_ERROR: the dim components of grad u; H1_ERROR: u, then grad u;
 *   VECTOR_SEMI_H1_ERROR (2D vector P2): the four components of grad u in
 *   the reference's order d u_n / d x_m at m 2 + n;
 *   INTEGRAL, INTEGRAL_DT: none; INTEGRAL_F: f; INTEGRAL_SOL_DEPENDENT_F:
 *   f(x, u), e.g. an FE_FN_POLY_U of the field's value.
 * The value is bitwise repeatable (fixed reduction order); integrals are
 * signed. */
#define FE_NORM_L1 1
#define FE_NORM_L1_ERROR 2
#define FE_NORM_L2 3
#define FE_NORM_VECTOR_L2 4
#define FE_NORM_L2_ERROR 5
#define FE_NORM_VECTOR_L2_ERROR 7
#define FE_NORM_LINF 8
#define FE_NORM_LINF_ERROR 9
#define FE_NORM_LINF_UDOT_ERROR 10
#define FE_NORM_VECTOR_LINF_ERROR 11
#define FE_NORM_VECTOR_LINF_UDOT_ERROR 12
#define FE_NORM_SEMI_H1 13
#define FE_NORM_SEMI_H1_ERROR 14
#define FE_NORM_VECTOR_SEMI_H1_ERROR 16
#define FE_NORM_H1 17
#define FE_NORM_H1_ERROR 18
#define FE_NORM_AREA 19
#define FE_NORM_INTEGRAL 20                 /* integral of u_h */
#define FE_NORM_INTEGRAL_SOL_DEPENDENT_F 21 /* integral of f(x, u_h) */
#define FE_NORM_INTEGRAL_DT 23              /* integral of du_h/dt from the BDF history */
#define FE_NORM_INTEGRAL_F 25               /* integral of f(x) */
int fe_hip_norm(fe_hip_ctx *ctx, int type, int field, int quad_degree, int nexact, const fe_hip_fn *exact,
                double *value);

/* Allocation-free fe_hip_norm.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight time-stepping or serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and value as fe_hip_norm, but it calls no
 * hipMalloc / hipFree: the reference tables, the block partials, the result
 * and the exact-solution tables (FE_FN_ELEM / FE_FN_QP / FE_FN_POLY_U data)
 * live in a static device scratch of the library, one copy per GPU, about
 * 8.1 MiB, held until the process ends.  Exact-solution tables larger than
 * 1,048,576 doubles in total are rejected with FE_STATUS_ERROR and a
 * message (e.g. an FE_FN_QP table of E nq > 2^20 values).  Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU (two host
 * threads, or two contexts on different streams of the same device) race and
 * must be serialized by the caller.  Nothing of the scratch needs to start
 * at zero: every value a call reads was written earlier in the same call. */
int fe_hip_norm_noalloc(fe_hip_ctx *ctx, int type, int field, int quad_degree, int nexact,
                        const fe_hip_fn *exact, double *value);

/* ---- aggregation multigrid (FE_PC_AMG) ----------------------------------
 * The algebraic multigrid the reference reaches through PETSc's runtime
 * options (-pc_type gamg, feLinearSystemPETSc.cpp:900-1031).  The hierarchy
 * of the CSR pattern is built on the first AMG solve; the numeric set-up runs
 * in every solve.  fe_hip_amg_levels reports the rows and nonzeros per level
 * (up to cap levels, *levels = 0 before the first AMG solve);
 * fe_hip_amg_apply applies one V-cycle z = M^-1 r (host arrays of n values,
 * FP64), setting the hierarchy up from the current values first. */
int fe_hip_amg_levels(fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap);
int fe_hip_amg_apply(fe_hip_ctx *ctx, const double *r, double *z);

/* ---- batched independent problems ---------------------------------------
 * No reference counterpart (the reference solves one system per
 * feLinearSystem).  The assembled system must be block diagonal: problem p
 * owns the unknown rows [rows[p], rows[p+1]) (rows[0] = 0, rows[nprob] = n)
 * and couples to no other problem, as when the meshes of many problems are
 * uploaded as one disjoint mesh with per-problem contiguous numbering.  Every
 * problem gets its own Jacobi-PCG (own scalars, own stop on
 * ||b_p - A_p x_p||_2 <= max(rtol ||b_p||_2, atol), FP64), all problems
 * advancing in the same launches; FE_PREC_FP32_IR runs the inner solves on
 * the FP32 values with FP64 refinement.  With pc = FE_PC_AMG every problem
 * runs its own pipelined PCG (own scalars and stop, FP64 vectors) with the
 * union AMG V-cycle as preconditioner: the hierarchy of a block-diagonal
 * matrix is block diagonal, so each problem sees its own V-cycle (tuning key
 * "batch_amg_global" = 1 instead runs one CG over the union system to
 * ||r||_2 <= min_p max(rtol ||b_p||_2, atol)).
 * The correction goes to the device du (fe_hip_correct_solution applies it).
 * ms_solve: device time. */
typedef struct {
  int iterations;
  int converged;
  double relres_fp64;
} fe_hip_batch_stats;
int fe_hip_solve_batched(fe_hip_ctx *ctx, int nprob, const int *rows, const fe_hip_solver_opts *o,
                         fe_hip_batch_stats *stats, double *ms_solve);
/* Times the batched iteration for every launch shape (lanes per row 2/4/8,
 * block 256/512/1024) and keeps the fastest in the context's tuning. */
int fe_hip_autotune_batched(fe_hip_ctx *ctx, int nprob, const int *rows, int precision, int *lanes, int *block,
                            double *best_ms);
/* Times one assembly of the current forms (zero + sweeps, best of three) for
 * every kernel variant -- the row gather in two slice orders, the colored
 * sweep with one thread per element or one lane per local row, the colored
 * MFMA sweep with 16 elements per wavefront -- and keeps the fastest in the
 * context's tuning (asm_kernel, scalar_lanes, gather_xcd); the variants that
 * do not apply to a form fall back to the sweep that does.  The system holds
 * an assembly afterwards.  No reference counterpart (the reference has one
 * element loop). */
int fe_hip_autotune_assembly(fe_hip_ctx *ctx, int precision, int *asm_kernel, int *scalar_lanes, int *gather_xcd,
                             double *best_ms);

/* ---- learned IC(0)-pattern correction (FE_PC_LEARNED) --------------------
 * No reference counterpart (the reference's PCG uses PETSc's IC(0),
 * feLinearSystemPETSc.cpp:900-1031).  A 3-layer, 32-wide message-passing
 * network on the graph of A corrects the split IC(0) factor
 * M^-1 = G^T D^-1 G (G = I - L_strict); with FE_PC_LEARNED, fe_hip_solve
 * factors IC(0) (three parallel sweeps), runs the network and applies the corrected factor.
 * Parameters: fe_hip_mpnn_param_count() floats.  Training minimises the
 * logarithm of Kaporin's condition number of the preconditioned matrix (a
 * bound on the CG iteration count), exactly, on the assembled matrix. */
typedef struct {
  double lr, beta1, beta2, eps; /* Adam */
} fe_hip_mpnn_opts;

int fe_hip_mpnn_param_count(void);
void fe_hip_mpnn_default_opts(fe_hip_mpnn_opts *o);
int fe_hip_mpnn_init(fe_hip_ctx *ctx, uint64_t seed);
int fe_hip_mpnn_set_params(fe_hip_ctx *ctx, const float *theta, int count);
int fe_hip_mpnn_get_params(fe_hip_ctx *ctx, float *theta, int count);
/* Factor IC(0) of the assembled matrix, run the network, the loss and its
 * gradient (grad may be NULL). */
int fe_hip_mpnn_loss_grad(fe_hip_ctx *ctx, double *loss, float *grad);
/* fe_hip_mpnn_loss_grad followed by one Adam step. */
int fe_hip_mpnn_train_step(fe_hip_ctx *ctx, const fe_hip_mpnn_opts *o, double *loss);
/* Intermediates of the last network run, as floats: "X" (n x 16), "h0".."h3",
 * "m1".."m3" (n x 32), "p", "q" (n x 8), "heads" (n x 5: tau, s, r, sigma,
 * rho), "B", "d" (n), "g", "w0", "w" (nnz; normalised strict-lower factor
 * entries, the IC(0) row products and the corrected ones). */
int fe_hip_mpnn_get_tensor(fe_hip_ctx *ctx, const char *name, float *out, int64_t count);

#ifdef __cplusplus
}
#endif

#endif
