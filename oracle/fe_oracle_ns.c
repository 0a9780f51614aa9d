/* Taylor-Hood P2-P1 Navier-Stokes forms of the CPU oracle -- TEST
 * INFRASTRUCTURE ONLY (see fe_oracle.c).
 *
 * Restates, per element and quadrature point, the reference terms of the
 * navier_stokes tests' form stack (tests/withLinearSolver/navier_stokes.cpp:
 * 83-99, divergence formulation):
 *   VectorConvectiveAcceleration (coeff -1)  feVectorSysElm.cpp:1153-1246
 *   MixedDivergence {p, u} (coeff 1)         feVectorSysElm.cpp:669-753
 *   VectorSource (S = 0)                     feVectorSysElm.cpp:101-132
 *   DivergenceNewtonianStress {u, p} (1, mu) feVectorSysElm.cpp:1434-1536
 * with the tensor helpers of feVectorSpace.cpp:33-57, 202-292, 353-501 and
 * the gradient layout grad[m*2+n] = d(phi)_n / dx_m (feSpace.cpp:748-759).
 * Vector P2 function f = 2 a + c is psi_a e_c (vectorLagrangeLayout,
 * feSpace_2D.cpp:40-54).
 *
 * Parity pin: tests/test_oracle_golden.py compares the assembled matrix and
 * residual with the reference's own values (<= 1e-12) on jittered meshes.
 * or_assemble_vector_transient_mass: its matrix is the scalar P2 mass
 * matrix (or_element_mass) on each component, uncoupled
 * (tests/test_oracle_time.py).
 */
#include "fe_oracle.h"

#include <math.h>
#include <string.h>

int or_quad_tri(int degree, double *r, double *s, double *w);

static int find_pos(const int *ia, const int *ja, int row, int col)
{
  for(int k = ia[row]; k < ia[row + 1]; ++k)
    if(ja[k] == col) return k;
  return -1;
}

static void p2_tables(double r, double s, double *L, double *dr, double *ds)
{
  L[0] = (1. - r - s) * (1. - 2. * r - 2. * s);
  L[1] = r * (2. * r - 1.);
  L[2] = s * (2. * s - 1.);
  L[3] = 4. * r * (1. - r - s);
  L[4] = 4. * r * s;
  L[5] = 4. * s * (1. - r - s);
  dr[0] = 4. * (r + s) - 3.; dr[1] = 4. * r - 1.; dr[2] = 0.;
  dr[3] = 4. * (1. - 2. * r - s); dr[4] = 4. * s; dr[5] = -4. * s;
  ds[0] = 4. * (r + s) - 3.; ds[1] = 0.; ds[2] = 4. * s - 1.;
  ds[3] = -4. * r; ds[4] = 4. * r; ds[5] = 4. * (1. - r - 2. * s);
}

/* Scatter one local block (Pardiso semantics). */
static int scatter(const int *ia, const int *ja, int n, const int *I, int nI, const int *Jc, int nJ,
                   const double *A, double *val)
{
  for(int i = 0; i < nI; ++i) {
    if(I[i] >= n) continue;
    for(int j = 0; j < nJ; ++j) {
      if(Jc[j] >= n) continue;
      int p = find_pos(ia, ja, I[i], Jc[j]);
      if(p < 0) return -1;
      val[p] += A[i * nJ + j];
    }
  }
  return 0;
}

int or_assemble_ns(int V, const double *xyz, int E, const int *conn, const int *adrU,
                   const int *adrP, int n, const int *ia, const int *ja, int quad_degree,
                   double mu, const double *u, const int *elem_order, double *val,
                   double *rhs)
{
  (void)V;
  double qr[36], qs[36], qw[36];
  const int nq = or_quad_tri(quad_degree, qr, qs, qw);
  if(nq <= 0) return -1;
  /* form order of the reference stack: conv, div, source, stress; each form
   * is assembled over all elements before the next (the reference loops
   * forms outermost, feLinearSystemMklPardiso.cpp:546-663). */
  for(int form = 0; form < 4; ++form) {
    for(int ee = 0; ee < E; ++ee) {
      const int e = elem_order ? elem_order[ee] : ee;
      double c[3][3];
      for(int j = 0; j < 3; ++j)
        for(int d = 0; d < 3; ++d) c[j][d] = xyz[3 * conn[3 * e + j] + d];
      const double dxdr0 = -c[0][0] + c[1][0], dxdr1 = -c[0][1] + c[1][1];
      const double dxds0 = -c[0][0] + c[2][0], dxds1 = -c[0][1] + c[2][1];
      const double J = dxdr0 * dxds1 - dxdr1 * dxds0;
      const double drdx0 = dxds1 / J, drdx1 = -dxdr1 / J, drdy0 = -dxds0 / J, drdy1 = dxdr0 / J;
      const int *aU = adrU + (size_t)e * 12, *aP = adrP + (size_t)e * 3;
      double su[12], sp[3];
      for(int i = 0; i < 12; ++i) su[i] = u[aU[i]];
      for(int i = 0; i < 3; ++i) sp[i] = u[aP[i]];
      double Aconv[144], Adiv[36], Astr[180], Bu[12], Bp[3];
      memset(Aconv, 0, sizeof(Aconv));
      memset(Adiv, 0, sizeof(Adiv));
      memset(Astr, 0, sizeof(Astr));
      memset(Bu, 0, sizeof(Bu));
      memset(Bp, 0, sizeof(Bp));
      for(int k = 0; k < nq; ++k) {
        double L[6], dr[6], ds[6];
        p2_tables(qr[k], qs[k], L, dr, ds);
        const double Lp[3] = {1. - qr[k] - qs[k], qr[k], qs[k]};
        const double jw = J * qw[k];
        /* gradPhi[f][m*2+n] of vector function f = 2a + c */
        double G[12][4];
        for(int a = 0; a < 6; ++a) {
          const double gx = dr[a] * drdx0 + ds[a] * drdx1;
          const double gy = dr[a] * drdy0 + ds[a] * drdy1;
          /* component x: phi = (psi, 0); component y: phi = (0, psi) */
          G[2 * a][0] = gx; G[2 * a][1] = 0.; G[2 * a][2] = gy; G[2 * a][3] = 0.;
          G[2 * a + 1][0] = 0.; G[2 * a + 1][1] = gx; G[2 * a + 1][2] = 0.; G[2 * a + 1][3] = gy;
        }
        /* phi[f][n] */
        double Ph[12][2];
        for(int a = 0; a < 6; ++a) {
          Ph[2 * a][0] = L[a]; Ph[2 * a][1] = 0.;
          Ph[2 * a + 1][0] = 0.; Ph[2 * a + 1][1] = L[a];
        }
        /* u and grad u (feSpace.cpp:1352-1405 sums over the 12 functions) */
        double uq[2] = {0., 0.}, gu[4] = {0., 0., 0., 0.};
        for(int f = 0; f < 12; ++f) {
          uq[0] += su[f] * Ph[f][0];
          uq[1] += su[f] * Ph[f][1];
          gu[0] += su[f] * G[f][0];
          gu[1] += su[f] * G[f][1];
          gu[2] += su[f] * G[f][2];
          gu[3] += su[f] * G[f][3];
        }
        double div[12];
        for(int f = 0; f < 12; ++f) div[f] = G[f][0] + G[f][3];
        if(form == 0) {
          const double coeff = -1.;
          for(int i = 0; i < 12; ++i)
            for(int j = 0; j < 12; ++j) {
              const double a1 = uq[0] * (G[j][0] * Ph[i][0] + G[j][1] * Ph[i][1]) +
                                uq[1] * (G[j][2] * Ph[i][0] + G[j][3] * Ph[i][1]);
              const double a2 = Ph[j][0] * (gu[0] * Ph[i][0] + gu[1] * Ph[i][1]) +
                                Ph[j][1] * (gu[2] * Ph[i][0] + gu[3] * Ph[i][1]);
              Aconv[i * 12 + j] += coeff * (a1 + a2) * jw;
            }
          double ugu[2] = {0., 0.};
          for(int m = 0; m < 2; ++m)
            for(int nn = 0; nn < 2; ++nn) ugu[m] += uq[nn] * gu[nn * 2 + m];
          for(int i = 0; i < 12; ++i) Bu[i] -= coeff * (Ph[i][0] * ugu[0] + Ph[i][1] * ugu[1]) * jw;
        } else if(form == 1) {
          const double coeff = 1.;
          for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 12; ++j) Adiv[i * 12 + j] += coeff * Lp[i] * div[j] * jw;
          double divu = gu[0] + gu[3];
          for(int i = 0; i < 3; ++i) Bp[i] -= coeff * divu * Lp[i] * jw;
        } else if(form == 2) {
          /* zero vector source: Be -= (0 . phi) J w */
          for(int i = 0; i < 12; ++i) Bu[i] -= (Ph[i][0] * 0. + Ph[i][1] * 0.) * jw;
        } else {
          const double coeff = 1.;
          double pq = 0.;
          for(int j = 0; j < 3; ++j) pq += sp[j] * Lp[j];
          double sym[4];
          for(int m = 0; m < 2; ++m)
            for(int nn = 0; nn < 2; ++nn) sym[m * 2 + nn] = gu[m * 2 + nn] + gu[nn * 2 + m];
          for(int i = 0; i < 12; ++i) {
            int Jc = 0;
            for(int j = 0; j < 12; ++j) {
              const double dc = G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2] + G[i][3] * G[j][3];
              const double dt = G[i][0] * G[j][0] + G[i][1] * G[j][2] + G[i][2] * G[j][1] + G[i][3] * G[j][3];
              Astr[i * 15 + Jc++] += -coeff * mu * (dc + dt) * jw;
            }
            for(int j = 0; j < 3; ++j) Astr[i * 15 + Jc++] += coeff * Lp[j] * div[i] * jw;
            const double dcs = G[i][0] * sym[0] + G[i][1] * sym[1] + G[i][2] * sym[2] + G[i][3] * sym[3];
            Bu[i] -= coeff * (pq * div[i] - mu * dcs) * jw;
          }
        }
      }
      int JuP[15];
      for(int j = 0; j < 12; ++j) JuP[j] = aU[j];
      for(int j = 0; j < 3; ++j) JuP[12 + j] = aP[j];
      if(form == 0) {
        if(scatter(ia, ja, n, aU, 12, aU, 12, Aconv, val)) return -2;
        for(int i = 0; i < 12; ++i)
          if(aU[i] < n) rhs[aU[i]] += Bu[i];
      } else if(form == 1) {
        if(scatter(ia, ja, n, aP, 3, aU, 12, Adiv, val)) return -2;
        for(int i = 0; i < 3; ++i)
          if(aP[i] < n) rhs[aP[i]] += Bp[i];
      } else if(form == 2) {
        for(int i = 0; i < 12; ++i)
          if(aU[i] < n) rhs[aU[i]] += Bu[i];
      } else {
        if(scatter(ia, ja, n, aU, 12, JuP, 15, Astr, val)) return -2;
        for(int i = 0; i < 12; ++i)
          if(aU[i] < n) rhs[aU[i]] += Bu[i];
      }
    }
  }
  return 0;
}

/* TransientVectorMass (feVectorSysElm.cpp:391-425) on the vector P2 field:
 * Ae_ij += m c0 (phi_i . phi_j) J w (dotProductShapeShape), Be_i -= m
 * (udot_h . phi_i) J w with udot_h interpolated from the nodal time
 * derivative (dotProductShapeOther).  Constant coefficient m. */
int or_assemble_vector_transient_mass(int V, const double *xyz, int E, const int *conn, const int *adrU, int n,
                                      const int *ia, const int *ja, int quad_degree, double m, double c0,
                                      const double *udot, double *val, double *rhs)
{
  (void)V;
  double qr[36], qs[36], qw[36];
  const int nq = or_quad_tri(quad_degree, qr, qs, qw);
  if(nq <= 0) return -1;
  for(int e = 0; e < E; ++e) {
    double c[3][3];
    for(int j = 0; j < 3; ++j)
      for(int d = 0; d < 3; ++d) c[j][d] = xyz[3 * conn[3 * e + j] + d];
    const double J = (-c[0][0] + c[1][0]) * (-c[0][1] + c[2][1]) - (-c[0][1] + c[1][1]) * (-c[0][0] + c[2][0]);
    const int *aU = adrU + (size_t)e * 12;
    double ud[12], A[144], B[12];
    for(int i = 0; i < 12; ++i) ud[i] = udot[aU[i]];
    memset(A, 0, sizeof(A));
    memset(B, 0, sizeof(B));
    for(int k = 0; k < nq; ++k) {
      double L[6], dr[6], ds[6];
      p2_tables(qr[k], qs[k], L, dr, ds);
      double Ph[12][2];
      for(int a = 0; a < 6; ++a) {
        Ph[2 * a][0] = L[a]; Ph[2 * a][1] = 0.;
        Ph[2 * a + 1][0] = 0.; Ph[2 * a + 1][1] = L[a];
      }
      double uq[2] = {0., 0.};
      for(int f = 0; f < 12; ++f) {
        uq[0] += ud[f] * Ph[f][0];
        uq[1] += ud[f] * Ph[f][1];
      }
      const double jw = J * qw[k];
      for(int i = 0; i < 12; ++i) {
        for(int j = 0; j < 12; ++j) A[i * 12 + j] += m * c0 * (Ph[i][0] * Ph[j][0] + Ph[i][1] * Ph[j][1]) * jw;
        B[i] -= m * (uq[0] * Ph[i][0] + uq[1] * Ph[i][1]) * jw;
      }
    }
    for(int i = 0; i < 12; ++i)
      if(aU[i] < n) rhs[aU[i]] += B[i];
    if(val && scatter(ia, ja, n, aU, 12, aU, 12, A, val)) return -1;
  }
  return 0;
}
