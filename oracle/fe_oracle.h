/* CPU oracle for the feNG assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline); the product path never calls it.
 *
 * Each function restates one step of the reference algorithm and cites the
 * reference file:line it follows.  All arrays are caller-owned, 0-based,
 * int32 indices and float64 values unless stated otherwise.
 */
#ifndef FE_ORACLE_H
#define FE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element families (local DOF layouts of feSpace_2D.cpp / feSpace_3D.cpp). */
enum { OR_TRI_P1 = 1, OR_TRI_P2 = 2, OR_TRI_VP2 = 3, OR_TET_P1 = 4, OR_TRI_P3 = 5, OR_TRI_P4 = 6 };

/* Coefficient / source functions evaluated at quadrature points. */
enum {
  OR_FN_CONST = 0,     /* c0 */
  OR_FN_SINSIN = 1,    /* c0 * sin(pi x) sin(pi y) */
  OR_FN_BLOCKS = 2,    /* exp(0.5*(((7 bi + 13 bj) % 11) - 5)), B = c0 blocks */
  OR_FN_SMOOTH3 = 3,   /* exp(c0 sin(2 pi x) sin(2 pi y) sin(2 pi z)) */
  OR_FN_ELEM = 4,      /* data[e] */
  OR_FN_QP = 5,        /* data[e*nq + q] */
  OR_FN_KOV_U = 6,     /* Kovasznay velocity component c1 at Re = c0 */
  OR_FN_KOV_P = 7      /* Kovasznay pressure at Re = c0 */
};

typedef struct {
  int kind;
  double c0, c1;
  const double *data;
} or_fn;

/* feMeshRead.cpp:1412-1456 -- signed 1-based edge tags by first encounter. */
int or_tri_edges(int E, const int *conn, int *edges);

/* feCncGeo.cpp:735-794 -- greedy independent-set coloring (algorithm 1). */
int or_color_elements(int E, int nvpe, const int *conn, int V, int *colorOf);

/* feNumber.cpp:370-483, 677-843 -- DOF numbering.
 * fields: kinds[f] in {OR_TRI_P1, OR_TRI_P2, OR_TRI_VP2, OR_TET_P1};
 * essential[f] != 0 marks the field's boundary entities as essential.
 * Boundary: nB boundary elements with bnv vertices each (bconn), and for
 * triangles bedge[nB] = edge index (|tag|-1) of each boundary segment, in the
 * order the boundary elements appear in the mesh file.
 * adr_out[f] receives E * nloc(kind) global DOF ids. */
int or_number(int V, int E, int nvpe, const int *conn, int nEdges, const int *edges,
              int nB, int bnv, const int *bconn, const int *bedge,
              int nfields, const int *kinds, const int *essential,
              int **adr_out, int *nInc, int *nDofs);

/* feCompressedRowStorage.cpp:15-133 -- CSR pattern.  forms: nforms matrix forms,
 * each with E elements, nI/nJ local sizes and adrI/adrJ (E*nI, E*nJ).
 * Call with ja == NULL to get nnz only. */
int64_t or_csr(int n, int nforms, const int *E, const int *nI, const int *nJ,
               const int *const *adrI, const int *const *adrJ, int *ia, int *ja);

/* Quadrature (feQuadratureTri.cpp:795-978 up to degree 12, feQuadratureTet.cpp:96-110). */
int or_quad_tri(int degree, double *r, double *s, double *w);
int or_quad_tet(int degree, double *r, double *s, double *t, double *w);

/* Scalar diffusion + source assembly (feSysElm.cpp:16-27, 520-589), scatter per
 * feLinearSystemMklPardiso.cpp:501-749.  kind: OR_TRI_P1/P2 or OR_TET_P1.
 * u: solution with essential values (nDofs).  Adds into val/rhs. */
int or_assemble_diffusion(int kind, int V, const double *xyz, int E, const int *conn,
                          const int *adr, int n, const int *ia, const int *ja,
                          int quad_degree, or_fn coeff, or_fn source, const double *u,
                          const int *elem_order, double *val, double *rhs);

/* The same assembly over the element colors with `threads` OpenMP threads
 * (colorOf from or_color_elements, ncolors colors). */
int or_assemble_diffusion_omp(int kind, int V, const double *xyz, int E, const int *conn,
                              const int *adr, int n, const int *ia, const int *ja,
                              int quad_degree, or_fn coeff, or_fn source, const double *u,
                              const int *colorOf, int ncolors, int threads, double *val, double *rhs);

/* Element matrix of one scalar diffusion element (for unit checks). */
int or_element_diffusion(int kind, const double *coords, int quad_degree, or_fn coeff,
                         int elem, double *Ae);

/* Transient mass (feSysElm.cpp:426-458): Ae += phi_i m c0 phi_j, Be -=
 * phi_i m udot_h with udot the nodal time derivative (nDofs).  val may be
 * NULL (residual only).  Adds into val/rhs. */
int or_assemble_transient_mass(int kind, int V, const double *xyz, int E, const int *conn,
                               const int *adr, int n, const int *ia, const int *ja, int quad_degree,
                               or_fn coeff, double c0, const double *udot, double *val, double *rhs);
/* TransientVectorMass (feVectorSysElm.cpp:391-425) on the vector P2 field
 * (adrU: E*12), constant coefficient m; val may be NULL.  Adds into val/rhs. */
int or_assemble_vector_transient_mass(int V, const double *xyz, int E, const int *conn, const int *adrU, int n,
                                      const int *ia, const int *ja, int quad_degree, double m, double c0,
                                      const double *udot, double *val, double *rhs);
/* Mass (p = 1) / mass power c u^p (feSysElm.cpp:186-225, :260-278) with
 * the derivative matrix c p u^(p-1) phi_i phi_j; val may be NULL. */
int or_assemble_mass_power(int kind, int V, const double *xyz, int E, const int *conn, const int *adr, int n,
                           const int *ia, const int *ja, int quad_degree, or_fn coeff, double p, const double *u,
                           double *val, double *rhs);
/* Advection in conservative form (feSysElm.cpp:766-835), velocity (cx, cy, cz). */
int or_assemble_advection(int kind, int V, const double *xyz, int E, const int *conn, const int *adr, int n,
                          const int *ia, const int *ja, int quad_degree, or_fn cx, or_fn cy, or_fn cz,
                          const double *u, double *val, double *rhs);
/* Mixed scalar forms with test functions of one field (adrI) and the
 * unknown u of another field (adrJ) of the same kind, constant coefficients:
 * mixed mass cm (feSysElm.cpp:295-337), mixed mass power cp u^p (:364-403),
 * mixed grad-grad cg (:899-976):
 *   Be_i -= (cm u + cp u^p) phi_i + cg grad(u).grad(phi_i),
 *   Ae_ij += (cm + cp p u^(p-1)) phi_i phi_j + cg grad(phi_i).grad(phi_j).
 * Rows adrI >= n are dropped; val may be NULL. */
int or_assemble_mixed(int kind, const double *xyz, int E, const int *conn, const int *adrI, const int *adrJ, int n,
                      const int *ia, const int *ja, int quad_degree, double cm, double cp, double p, double cg,
                      const double *u, double *val, double *rhs);
/* Shape functions of a kind at a reference point and their r, s (t)
 * derivatives; returns the number of functions (dt may be NULL in 2D). */
int or_basis(int kind, double r, double s, double t, double *L, double *dr, double *ds, double *dt);
/* Nonlinear diffusion (feSysElm.cpp:681-760): conductivity k(u) and its
 * derivative dk/du given as polynomials in the solution's value at the
 * quadrature point (coefficients of u^0, u^1, ...):
 *   Ae_ij += k grad(phi_i).grad(phi_j) + dk phi_j grad(u).grad(phi_i),
 *   Be_i  -= k grad(phi_i).grad(u).  val may be NULL. */
int or_assemble_nonlinear_diffusion(int kind, int V, const double *xyz, int E, const int *conn, const int *adr,
                                    int n, const int *ia, const int *ja, int quad_degree, int nk, const double *k,
                                    int ndk, const double *dk, const double *u, double *val, double *rhs);
/* L2 error (triangles) against exact values at the quadrature points ue[E nq]. */
double or_l2_error_qp(int kind, const double *xyz, int E, const int *conn, const int *adr, const double *u,
                      int quad_degree, const double *ue);
/* Element mass matrix (coefficient m, c0 = 1) of one element (for unit checks). */
int or_element_mass(int kind, const double *coords, int quad_degree, or_fn coeff, int elem, double *Ae);

/* Taylor-Hood P2-P1 Navier-Stokes forms (feVectorSysElm.cpp).  adrU: E*12,
 * adrP: E*3.  Form stack: convective (c=-1), mixed divergence (c=1), zero
 * vector source, divergence of the Newtonian stress (c=1, mu). */
int or_assemble_ns(int V, const double *xyz, int E, const int *conn, const int *adrU,
                   const int *adrP, int n, const int *ia, const int *ja, int quad_degree,
                   double mu, const double *u, const int *elem_order, double *val,
                   double *rhs);

/* Krylov solvers (semantics of feLinearSystemPETSc.cpp:900-1031 with the
 * KSP/PC the configs select). Return iterations (<0 on breakdown). */
int or_spmv(int n, const int *ia, const int *ja, const double *val, const double *x,
            double *y);
int or_cg_jacobi(int n, const int *ia, const int *ja, const double *val, const double *b,
                 double *x, double rtol, int maxit, double *relres);
int or_cg_jacobi_omp(int n, const int *ia, const int *ja, const double *val, const double *b,
                     double *x, double rtol, int maxit, int threads, double *relres);
int or_ic0(int n, const int *ia, const int *ja, const double *val, double *lval);
int or_pcg_ic0(int n, const int *ia, const int *ja, const double *val, const double *lval,
               const double *b, double *x, double rtol, int maxit, double *relres);
int or_ilu0(int n, const int *ia, const int *ja, const double *val, double *luval);
int or_gmres_ilu0(int n, const int *ia, const int *ja, const double *val,
                  const double *luval, const double *b, double *x, int restart,
                  double rtol, int maxit, double *relres);

/* feNorm.cpp:323-398 -- L2 error of a scalar P1/P2 field vs the manufactured
 * u = sin(pi x) sin(pi y) with the given quadrature degree. */
double or_l2_error_sinsin(int kind, const double *xyz, int E, const int *conn,
                          const int *adr, const double *u, int quad_degree);

#ifdef __cplusplus
}
#endif

#endif
