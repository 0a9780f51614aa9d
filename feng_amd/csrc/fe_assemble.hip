// Element kernels and the colored, atomic-free scatter (the device
// counterpart of feLinearSystemPETSc::assembleMatrices/assembleResiduals,
// feLinearSystemPETSc.cpp:543-805, with the Pardiso CSR scatter semantics of
// feLinearSystemMklPardiso.cpp:501-749).
//
// Layout: elements are processed in color-sorted order; one launch per
// color, one thread per element.  Within a color no two elements share a
// vertex, hence no two threads add into the same CSR value or RHS entry, so
// the scatter uses plain read-modify-write (deterministic, no atomics).  The
// reference-element tables are staged in LDS; geometry is recomputed from the
// vertex coordinates in registers.
#include "fe_hip_internal.h"
#include "fe_element.h"

#include <cmath>

namespace {

struct ScalarAsmArgs {
  int begin, count, E, n, what;
  const int *perm, *conn, *adr, *pos;
  const double *xyz, *tab, *u;
  FeFn coeff, src, mass, react, adv[3];
  FeFn dcoeff;  // NONLINEAR_DIFFUSION: dk/du (coeff is then k(u))
  int hasDiff, hasSrc, hasMass, hasReact, hasAdv, nlDiff;
  double reactPow;  // MASS form: c u^p (p = 1: the plain mass)
  // time derivative u_dot = cn[0] u + sum_{j = 1}^{nsol - 1} cn[j] past[j - 1]
  int nsol;
  double cn[fe_hip_ctx::kMaxPast + 1];
  const double *past[fe_hip_ctx::kMaxPast];
  double *val, *rhs;
};

// Scalar diffusion (feSysElm.cpp:530-583) or nonlinear diffusion
// (feSysElm.cpp:681-760), advection (feSysElm.cpp:766-835),
// transient mass (feSysElm.cpp:426-458), mass / mass power
// (feSysElm.cpp:186-225, :260-278) and source (feSysElm.cpp:16-27) of one
// field and quadrature rule, fused, on one element per thread.
template <int NF, int DIM, bool RARE = true>
__global__ __launch_bounds__(256) void asmScalarFP64(ScalarAsmArgs a)
{
  __shared__ FeRefTablesFor<NF> T;
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(T) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= a.count) return;
  const int ee = a.begin + t;
  const int e = a.perm[ee];
  constexpr int NV = DIM + 1;
  double c[NV][3];
#pragma unroll
  for(int j = 0; j < NV; ++j) {
    const int v = a.conn[(size_t)e * NV + j];
#pragma unroll
    for(int d = 0; d < 3; ++d) c[j][d] = a.xyz[3 * (size_t)v + d];
  }
  double J, drdx[3], drdy[3], drdz[3];
  feGeometry<DIM>(c, J, drdx, drdy, drdz);
  int ad[NF];
  double ul[NF];
#pragma unroll
  for(int i = 0; i < NF; ++i) {
    ad[i] = a.adr[(size_t)ee * NF + i];
    ul[i] = a.u[ad[i]];
  }
  double Ae[NF][NF], Be[NF];
#pragma unroll
  for(int i = 0; i < NF; ++i) {
    Be[i] = 0.;
#pragma unroll
    for(int j = 0; j < NF; ++j) Ae[i][j] = 0.;
  }
  const bool doMat = (a.what & 1) && (a.hasDiff || a.hasMass || a.hasReact || a.hasAdv), doRes = (a.what & 2);
  const int nq = T.nq;
  if(a.hasDiff) {
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      double uq = 0.;
      if((RARE && a.nlDiff))
#pragma unroll
        for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ul[j];
      const double kc = (RARE && a.nlDiff) ? feEvalFnU(a.coeff, x, e, nq, k, uq) : feEvalFn(a.coeff, x, e, nq, k);
      double gp[NF][DIM];
#pragma unroll
      for(int i = 0; i < NF; ++i) {
        if(DIM == 2) {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1];
        } else {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1] + T.dLt[k][i] * drdx[2];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1] + T.dLt[k][i] * drdy[2];
          gp[i][2 % DIM] = T.dLr[k][i] * drdz[0] + T.dLs[k][i] * drdz[1] + T.dLt[k][i] * drdz[2];
        }
      }
      const double jw = J * T.w[k];
      double gu[DIM];
      if(doRes || (RARE && a.nlDiff)) {
#pragma unroll
        for(int d = 0; d < DIM; ++d) gu[d] = 0.;
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int d = 0; d < DIM; ++d) gu[d] += ul[i] * gp[i][d];
      }
      if(doMat) {
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int j = 0; j < NF; ++j)
#pragma unroll
            for(int d = 0; d < DIM; ++d) Ae[i][j] += kc * gp[i][d] * gp[j][d] * jw;
        if((RARE && a.nlDiff)) {
          // dk/du phi_j grad(u) . grad(phi_i): the derivative of k(u)
          const double dk = feEvalFnU(a.dcoeff, x, e, nq, k, uq);
#pragma unroll
          for(int i = 0; i < NF; ++i) {
            double gi = 0.;
#pragma unroll
            for(int d = 0; d < DIM; ++d) gi += gu[d] * gp[i][d];
#pragma unroll
            for(int j = 0; j < NF; ++j) Ae[i][j] += dk * T.L[k][j] * gi * jw;
          }
        }
      }
      if(doRes) {
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int d = 0; d < DIM; ++d) Be[i] -= kc * gp[i][d] * gu[d] * jw;
      }
    }
  }
  if(RARE && a.hasAdv) {
    // Ae_ij -= c . grad(phi_i) phi_j, Be_i += u c . grad(phi_i) (conservative
    // form, feSysElm.cpp:766-835)
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      double cv[DIM];
#pragma unroll
      for(int d = 0; d < DIM; ++d) cv[d] = feEvalFn(a.adv[d], x, e, nq, k);
      double gp[NF][DIM];
#pragma unroll
      for(int i = 0; i < NF; ++i) {
        if(DIM == 2) {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1];
        } else {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1] + T.dLt[k][i] * drdx[2];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1] + T.dLt[k][i] * drdy[2];
          gp[i][2 % DIM] = T.dLr[k][i] * drdz[0] + T.dLs[k][i] * drdz[1] + T.dLt[k][i] * drdz[2];
        }
      }
      const double jw = J * T.w[k];
      if(doMat) {
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int j = 0; j < NF; ++j)
#pragma unroll
            for(int d = 0; d < DIM; ++d) Ae[i][j] -= cv[d] * gp[i][d] * T.L[k][j] * jw;
      }
      if(doRes) {
        double uq = 0.;
#pragma unroll
        for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ul[j];
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int d = 0; d < DIM; ++d) Be[i] += cv[d] * uq * gp[i][d] * jw;
      }
    }
  }
  if(a.hasMass) {
    // nodal time derivative of the element's DOFs
    double ud[NF];
    if(doRes) {
#pragma unroll
      for(int i = 0; i < NF; ++i) {
        double v = a.cn[0] * ul[i];
        for(int j = 1; j < a.nsol; ++j) v += a.cn[j] * a.past[j - 1][ad[i]];
        ud[i] = v;
      }
    }
    const double c0 = a.cn[0];
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double m = feEvalFn(a.mass, x, e, nq, k);
      const double jw = J * T.w[k];
      if(doMat) {
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int j = 0; j < NF; ++j) Ae[i][j] += T.L[k][i] * m * c0 * T.L[k][j] * jw;
      }
      if(doRes) {
        double uq = 0.;
#pragma unroll
        for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ud[j];
#pragma unroll
        for(int i = 0; i < NF; ++i) Be[i] -= T.L[k][i] * m * uq * jw;
      }
    }
  }
  if(RARE && a.hasReact) {
    // Be -= c u^p phi_i; Ae += c p u^(p-1) phi_i phi_j (the derivative the
    // reference approximates by finite differences for p != 1)
    const double pw = a.reactPow;
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double cr = feEvalFn(a.react, x, e, nq, k);
      const double jw = J * T.w[k];
      double uq = 0.;
#pragma unroll
      for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ul[j];
      if(doMat) {
        const double d = pw == 1. ? cr : cr * pw * pow(uq, pw - 1.);
#pragma unroll
        for(int i = 0; i < NF; ++i)
#pragma unroll
          for(int j = 0; j < NF; ++j) Ae[i][j] += d * T.L[k][i] * T.L[k][j] * jw;
      }
      if(doRes) {
        const double up = pw == 1. ? uq : pow(uq, pw);
#pragma unroll
        for(int i = 0; i < NF; ++i) Be[i] -= cr * up * T.L[k][i] * jw;
      }
    }
  }
  if(doRes && a.hasSrc) {
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double S = feEvalFn(a.src, x, e, nq, k);
      const double jw = J * T.w[k];
#pragma unroll
      for(int i = 0; i < NF; ++i) Be[i] -= T.L[k][i] * S * jw;
    }
  }
  // Atomic-free scatter (exclusive ownership within the color).
#pragma unroll
  for(int i = 0; i < NF; ++i) {
    if(doRes && ad[i] < a.n) a.rhs[ad[i]] += Be[i];
    if(doMat) {
#pragma unroll
      for(int j = 0; j < NF; ++j) {
        const int p = a.pos[(size_t)(i * NF + j) * a.E + ee];
        if(p >= 0) a.val[p] += Ae[i][j];
      }
    }
  }
}


// Mixed scalar forms of a (test field, trial field) pair of one kind and
// quadrature rule, fused: mixed mass (feSysElm.cpp:295-337), mixed mass
// power (:364-403) and mixed grad-grad (:899-976), one element per thread.
struct MixedAsmArgs {
  int begin, count, E, n, what;
  const int *perm, *conn, *adrI, *adrJ, *pos;
  const double *xyz, *tab, *u;
  FeFn cMass, cPow, cGG;
  double pw;
  int hasMass, hasPow, hasGG;
  double *val, *rhs;
};

template <int NF, int DIM>
__global__ __launch_bounds__(256) void asmMixedScalarFP64(MixedAsmArgs a)
{
  __shared__ FeRefTablesFor<NF> T;
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(T) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= a.count) return;
  const int ee = a.begin + t;
  const int e = a.perm[ee];
  constexpr int NV = DIM + 1;
  double c[NV][3];
#pragma unroll
  for(int j = 0; j < NV; ++j) {
    const int v = a.conn[(size_t)e * NV + j];
#pragma unroll
    for(int d = 0; d < 3; ++d) c[j][d] = a.xyz[3 * (size_t)v + d];
  }
  double J, drdx[3], drdy[3], drdz[3];
  feGeometry<DIM>(c, J, drdx, drdy, drdz);
  double uj[NF];
#pragma unroll
  for(int j = 0; j < NF; ++j) uj[j] = a.u[a.adrJ[(size_t)ee * NF + j]];
  double Ae[NF][NF], Be[NF];
#pragma unroll
  for(int i = 0; i < NF; ++i) {
    Be[i] = 0.;
#pragma unroll
    for(int j = 0; j < NF; ++j) Ae[i][j] = 0.;
  }
  const bool doMat = a.what & 1, doRes = a.what & 2;
  for(int k = 0; k < T.nq; ++k) {
    double x[3] = {0., 0., 0.};
#pragma unroll
    for(int j = 0; j < NV; ++j)
#pragma unroll
      for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
    const double jw = J * T.w[k];
    double uq = 0.;
#pragma unroll
    for(int j = 0; j < NF; ++j) uq += T.L[k][j] * uj[j];
    // value terms: r = cM u + cP u^p, with derivative dr/du
    double r = 0., dr = 0.;
    if(a.hasMass) {
      const double cm = feEvalFn(a.cMass, x, e, T.nq, k);
      r += cm * uq;
      dr += cm;
    }
    if(a.hasPow) {
      const double cp = feEvalFn(a.cPow, x, e, T.nq, k);
      r += cp * pow(uq, a.pw);
      dr += cp * a.pw * pow(uq, a.pw - 1.);
    }
    double gp[NF][DIM], gu[DIM], cg = 0.;
    if(a.hasGG) {
      cg = feEvalFn(a.cGG, x, e, T.nq, k);
#pragma unroll
      for(int i = 0; i < NF; ++i) {
        if(DIM == 2) {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1];
        } else {
          gp[i][0] = T.dLr[k][i] * drdx[0] + T.dLs[k][i] * drdx[1] + T.dLt[k][i] * drdx[2];
          gp[i][1 % DIM] = T.dLr[k][i] * drdy[0] + T.dLs[k][i] * drdy[1] + T.dLt[k][i] * drdy[2];
          gp[i][2 % DIM] = T.dLr[k][i] * drdz[0] + T.dLs[k][i] * drdz[1] + T.dLt[k][i] * drdz[2];
        }
      }
#pragma unroll
      for(int d = 0; d < DIM; ++d) gu[d] = 0.;
#pragma unroll
      for(int j = 0; j < NF; ++j)
#pragma unroll
        for(int d = 0; d < DIM; ++d) gu[d] += uj[j] * gp[j][d];
    }
#pragma unroll
    for(int i = 0; i < NF; ++i) {
      if(doRes) {
        double b = r * T.L[k][i];
        if(a.hasGG)
#pragma unroll
          for(int d = 0; d < DIM; ++d) b += cg * gu[d] * gp[i][d];
        Be[i] -= b * jw;
      }
      if(doMat)
#pragma unroll
        for(int j = 0; j < NF; ++j) {
          double v = dr * T.L[k][i] * T.L[k][j];
          if(a.hasGG)
#pragma unroll
            for(int d = 0; d < DIM; ++d) v += cg * gp[i][d] * gp[j][d];
          Ae[i][j] += v * jw;
        }
    }
  }
#pragma unroll
  for(int i = 0; i < NF; ++i) {
    const int row = a.adrI[(size_t)ee * NF + i];
    if(doRes && row < a.n) a.rhs[row] += Be[i];
    if(doMat) {
#pragma unroll
      for(int j = 0; j < NF; ++j) {
        const int p = a.pos[(size_t)(i * NF + j) * a.E + ee];
        if(p >= 0) a.val[p] += Ae[i][j];
      }
    }
  }
}

// The same fused forms with one lane per local row: LPE lanes (4 for P1, 8
// for P2) share an element, lane i owning row i of Ae and Be.  The element's
// vertices, DOF values and time derivatives are loaded once by the lane
// that owns them and exchanged with shuffles; every lane rebuilds the
// reference-element gradients from the LDS tables and accumulates its own
// row in the scalar kernel's order.  A color of ~50k elements becomes
// ~3,100 wavefronts instead of ~780, so the dependent-load chain of the
// colored sweep is hidden by four times the occupancy.
template <int NF, int DIM>
__global__ __launch_bounds__(256) void asmScalarLanes(ScalarAsmArgs a)
{
  __shared__ FeRefTables T;
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(FeRefTables) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  constexpr int LPE = NF <= 4 ? 4 : 8;
  constexpr int NV = DIM + 1;
  const int lane = threadIdx.x & 63;
  const int i = lane % LPE, base = lane - i;
  const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) / LPE);
  const bool valid = t < a.count;
  const int ee = a.begin + (valid ? t : 0);
  const int e = a.perm[ee];
  double vc[3] = {0., 0., 0.};
  if(i < NV) {
    const int v = a.conn[(size_t)e * NV + i];
#pragma unroll
    for(int d = 0; d < 3; ++d) vc[d] = a.xyz[3 * (size_t)v + d];
  }
  double c[NV][3];
#pragma unroll
  for(int j = 0; j < NV; ++j)
#pragma unroll
    for(int d = 0; d < 3; ++d) c[j][d] = __shfl(vc[d], base + j, 64);
  double J, drdx[3], drdy[3], drdz[3];
  feGeometry<DIM>(c, J, drdx, drdy, drdz);
  const bool row = i < NF;
  const int ad = row ? a.adr[(size_t)ee * NF + i] : 0;
  const double uo = row ? a.u[ad] : 0.;
  double udo = 0.;
  const bool doMat = (a.what & 1) && (a.hasDiff || a.hasMass || a.hasReact || a.hasAdv), doRes = (a.what & 2);
  if(a.hasMass && doRes && row) {
    udo = a.cn[0] * uo;
    for(int j = 1; j < a.nsol; ++j) udo += a.cn[j] * a.past[j - 1][ad];
  }
  double ul[NF], ud[NF];
#pragma unroll
  for(int j = 0; j < NF; ++j) {
    ul[j] = __shfl(uo, base + j, 64);
    ud[j] = __shfl(udo, base + j, 64);
  }
  const int ir = row ? i : 0;  // LDS column of this lane's row
  double Ae[NF], Be = 0.;
#pragma unroll
  for(int j = 0; j < NF; ++j) Ae[j] = 0.;
  const int nq = T.nq;
  if(a.hasDiff) {
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double kc = feEvalFn(a.coeff, x, e, nq, k);
      double gp[NF][DIM], gi[DIM];
#pragma unroll
      for(int j = 0; j < NF; ++j) {
        if(DIM == 2) {
          gp[j][0] = T.dLr[k][j] * drdx[0] + T.dLs[k][j] * drdx[1];
          gp[j][1 % DIM] = T.dLr[k][j] * drdy[0] + T.dLs[k][j] * drdy[1];
        } else {
          gp[j][0] = T.dLr[k][j] * drdx[0] + T.dLs[k][j] * drdx[1] + T.dLt[k][j] * drdx[2];
          gp[j][1 % DIM] = T.dLr[k][j] * drdy[0] + T.dLs[k][j] * drdy[1] + T.dLt[k][j] * drdy[2];
          gp[j][2 % DIM] = T.dLr[k][j] * drdz[0] + T.dLs[k][j] * drdz[1] + T.dLt[k][j] * drdz[2];
        }
      }
      if(DIM == 2) {
        gi[0] = T.dLr[k][ir] * drdx[0] + T.dLs[k][ir] * drdx[1];
        gi[1 % DIM] = T.dLr[k][ir] * drdy[0] + T.dLs[k][ir] * drdy[1];
      } else {
        gi[0] = T.dLr[k][ir] * drdx[0] + T.dLs[k][ir] * drdx[1] + T.dLt[k][ir] * drdx[2];
        gi[1 % DIM] = T.dLr[k][ir] * drdy[0] + T.dLs[k][ir] * drdy[1] + T.dLt[k][ir] * drdy[2];
        gi[2 % DIM] = T.dLr[k][ir] * drdz[0] + T.dLs[k][ir] * drdz[1] + T.dLt[k][ir] * drdz[2];
      }
      const double jw = J * T.w[k];
      if(doMat) {
#pragma unroll
        for(int j = 0; j < NF; ++j)
#pragma unroll
          for(int d = 0; d < DIM; ++d) Ae[j] += kc * gi[d] * gp[j][d] * jw;
      }
      if(doRes) {
        double gu[DIM];
#pragma unroll
        for(int d = 0; d < DIM; ++d) gu[d] = 0.;
#pragma unroll
        for(int j = 0; j < NF; ++j)
#pragma unroll
          for(int d = 0; d < DIM; ++d) gu[d] += ul[j] * gp[j][d];
#pragma unroll
        for(int d = 0; d < DIM; ++d) Be -= kc * gi[d] * gu[d] * jw;
      }
    }
  }
  if(a.hasAdv) {
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      double cv[DIM], gi[DIM];
#pragma unroll
      for(int d = 0; d < DIM; ++d) cv[d] = feEvalFn(a.adv[d], x, e, nq, k);
      if(DIM == 2) {
        gi[0] = T.dLr[k][ir] * drdx[0] + T.dLs[k][ir] * drdx[1];
        gi[1 % DIM] = T.dLr[k][ir] * drdy[0] + T.dLs[k][ir] * drdy[1];
      } else {
        gi[0] = T.dLr[k][ir] * drdx[0] + T.dLs[k][ir] * drdx[1] + T.dLt[k][ir] * drdx[2];
        gi[1 % DIM] = T.dLr[k][ir] * drdy[0] + T.dLs[k][ir] * drdy[1] + T.dLt[k][ir] * drdy[2];
        gi[2 % DIM] = T.dLr[k][ir] * drdz[0] + T.dLs[k][ir] * drdz[1] + T.dLt[k][ir] * drdz[2];
      }
      const double jw = J * T.w[k];
      if(doMat) {
#pragma unroll
        for(int j = 0; j < NF; ++j)
#pragma unroll
          for(int d = 0; d < DIM; ++d) Ae[j] -= cv[d] * gi[d] * T.L[k][j] * jw;
      }
      if(doRes) {
        double uq = 0.;
#pragma unroll
        for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ul[j];
#pragma unroll
        for(int d = 0; d < DIM; ++d) Be += cv[d] * uq * gi[d] * jw;
      }
    }
  }
  if(a.hasMass) {
    const double c0 = a.cn[0];
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double m = feEvalFn(a.mass, x, e, nq, k);
      const double jw = J * T.w[k];
      const double Li = T.L[k][ir];
      if(doMat) {
#pragma unroll
        for(int j = 0; j < NF; ++j) Ae[j] += Li * m * c0 * T.L[k][j] * jw;
      }
      if(doRes) {
        double uq = 0.;
#pragma unroll
        for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ud[j];
        Be -= Li * m * uq * jw;
      }
    }
  }
  if(a.hasReact) {
    const double pw = a.reactPow;
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double cr = feEvalFn(a.react, x, e, nq, k);
      const double jw = J * T.w[k];
      const double Li = T.L[k][ir];
      double uq = 0.;
#pragma unroll
      for(int j = 0; j < NF; ++j) uq += T.L[k][j] * ul[j];
      if(doMat) {
        const double dd = pw == 1. ? cr : cr * pw * pow(uq, pw - 1.);
#pragma unroll
        for(int j = 0; j < NF; ++j) Ae[j] += dd * Li * T.L[k][j] * jw;
      }
      if(doRes) {
        const double up = pw == 1. ? uq : pow(uq, pw);
        Be -= cr * up * Li * jw;
      }
    }
  }
  if(doRes && a.hasSrc) {
    for(int k = 0; k < nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double S = feEvalFn(a.src, x, e, nq, k);
      const double jw = J * T.w[k];
      Be -= T.L[k][ir] * S * jw;
    }
  }
  if(!valid || !row) return;
  if(doRes && ad < a.n) a.rhs[ad] += Be;
  if(doMat) {
#pragma unroll
    for(int j = 0; j < NF; ++j) {
      const int p = a.pos[(size_t)(i * NF + j) * a.E + ee];
      if(p >= 0) a.val[p] += Ae[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Taylor-Hood P2-P1 Navier-Stokes form stack fused in one kernel (divergence
// formulation of the reference's navier_stokes tests, navier_stokes.cpp:83-99):
//   VectorConvectiveAcceleration  feVectorSysElm.cpp:1153-1246
//   MixedDivergence {p, u}        feVectorSysElm.cpp:669-753
//   VectorSource                  feVectorSysElm.cpp:101-132
//   DivergenceNewtonianStress     feVectorSysElm.cpp:1434-1536
//   TransientVectorMass           feVectorSysElm.cpp:391-425
//   VectorDiffusion               feVectorSysElm.cpp:450-503
//   MixedGradient {u, p}          feVectorSysElm.cpp:529-576
//   VectorMass, GradSource        feVectorSysElm.cpp:196-237, :151-172
// 16 lanes per element: lane i < 12 owns velocity row i = 2a + c (function
// psi_a e_c), lanes 12..14 own the pressure rows, lane 15 only loads.  The
// element's local solution sits in LDS; every lane rebuilds the P2 gradients
// at each quadrature point from the LDS tables and accumulates its own row of
// every local matrix in registers, then scatters it through the position
// maps (exclusive ownership within the color: no atomics).
struct NsAsmArgs {
  int begin, count, E, n, what;
  const int *perm, *conn;
  const int *adrU, *adrP;  // color-sorted element maps (12, 3 per element); adrP null: no pressure
  const int *posUU, *posDiv, *posStr, *posGrad;  // posUU: the velocity-velocity group (convection, mass, diffusion)
  const double *xyz, *tab, *u;
  int hasConv, hasDiv, hasStr, hasSrc, hasMass, hasVDiff, hasGrad, hasVMass, hasGSrc;
  FeFn cConv, cDiv, cStr, mu, src0, src1, cMass, cVDiff, dVDiff, cGrad, cVMass, gSrc;
  // time derivative u_dot = cn[0] u + sum_{j = 1}^{nsol - 1} cn[j] past[j - 1]
  int nsol;
  double cn[fe_hip_ctx::kMaxPast + 1];
  const double *past[fe_hip_ctx::kMaxPast];
  double *val, *rhs;
};

constexpr int kNsElemPerBlock = 16;

__global__ __launch_bounds__(256) void asmNavierStokesFP64(NsAsmArgs a)
{
  __shared__ FeRefTables T;
  __shared__ double su[kNsElemPerBlock][16];  // 12 velocity + 3 pressure values
  __shared__ double sud[kNsElemPerBlock][12];  // velocity time derivative (TransientVectorMass)
  __shared__ double geo[kNsElemPerBlock][8];  // J, drdx0, drdx1, drdy0, drdy1, x0, y0 (pos origin)
  __shared__ double cxy[kNsElemPerBlock][6];
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(FeRefTables) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  const int le = threadIdx.x >> 4, lane = threadIdx.x & 15;
  const int t = blockIdx.x * kNsElemPerBlock + le;
  const bool valid = t < a.count;
  const int ee = a.begin + (valid ? t : 0);
  const int e = a.perm[ee];
  const bool hasP = a.adrP != nullptr;
  int myDof = 0;
  if(lane < 12) myDof = a.adrU[(size_t)ee * 12 + lane];
  else if(lane < 15 && hasP) myDof = a.adrP[(size_t)ee * 3 + lane - 12];
  if(lane < 15) su[le][lane] = (lane < 12 || hasP) ? a.u[myDof] : 0.;
  if(lane < 12 && a.hasMass) {
    double v = a.cn[0] * su[le][lane];
    for(int j = 1; j < a.nsol; ++j) v += a.cn[j] * a.past[j - 1][myDof];
    sud[le][lane] = v;
  }
  if(lane < 6) {
    const int v = a.conn[(size_t)e * 3 + lane / 2];
    cxy[le][lane] = a.xyz[3 * (size_t)v + (lane & 1)];
  }
  __syncthreads();
  if(lane == 0) {
    const double *c = cxy[le];
    const double dxdr0 = -c[0] + c[2], dxdr1 = -c[1] + c[3];
    const double dxds0 = -c[0] + c[4], dxds1 = -c[1] + c[5];
    const double J = dxdr0 * dxds1 - dxdr1 * dxds0;
    geo[le][0] = J;
    geo[le][1] = dxds1 / J;
    geo[le][2] = -dxdr1 / J;
    geo[le][3] = -dxds0 / J;
    geo[le][4] = dxdr0 / J;
  }
  __syncthreads();
  if(!valid || lane == 15) return;
  const double J = geo[le][0], drdx0 = geo[le][1], drdx1 = geo[le][2], drdy0 = geo[le][3], drdy1 = geo[le][4];
  const bool uRow = lane < 12;
  const int ai = lane >> 1, ci = lane & 1, ip = lane - 12;
  const bool doMat = (a.what & 1), doRes = (a.what & 2);
  double Ac[12], As[15], Ad[12], Ag[3], Be = 0.;
#pragma unroll
  for(int j = 0; j < 12; ++j) { Ac[j] = 0.; Ad[j] = 0.; }
#pragma unroll
  for(int j = 0; j < 3; ++j) Ag[j] = 0.;
#pragma unroll
  for(int j = 0; j < 15; ++j) As[j] = 0.;
  const int nq = T.nq;
  for(int k = 0; k < nq; ++k) {
    double gx[6], gy[6], ps[6];
#pragma unroll
    for(int b = 0; b < 6; ++b) {
      gx[b] = T.dLr[k][b] * drdx0 + T.dLs[k][b] * drdx1;
      gy[b] = T.dLr[k][b] * drdy0 + T.dLs[k][b] * drdy1;
      ps[b] = T.L[k][b];
    }
    const double pp[3] = {T.Lg[k][0], T.Lg[k][1], T.Lg[k][2]};
    // u, grad u (grad[m*2+n] = du_n/dx_m), p at the quadrature point
    double ux = 0., uy = 0., g0 = 0., g1 = 0., g2 = 0., g3 = 0.;
#pragma unroll
    for(int b = 0; b < 6; ++b) {
      const double vx = su[le][2 * b], vy = su[le][2 * b + 1];
      ux += vx * ps[b];
      uy += vy * ps[b];
      g0 += vx * gx[b];
      g1 += vy * gx[b];
      g2 += vx * gy[b];
      g3 += vy * gy[b];
    }
    const double pq = su[le][12] * pp[0] + su[le][13] * pp[1] + su[le][14] * pp[2];
    double x[3] = {0., 0., 0.};
#pragma unroll
    for(int j = 0; j < 3; ++j) {
      x[0] += cxy[le][2 * j] * T.Lg[k][j];
      x[1] += cxy[le][2 * j + 1] * T.Lg[k][j];
    }
    const double jw = J * T.w[k];
    if(uRow) {
      const double gi[2] = {gx[ai], gy[ai]};
      const double psi = ps[ai];
      if(a.hasConv) {
        const double cc = feEvalFn(a.cConv, x, e, nq, k);
        const double gu[4] = {g0, g1, g2, g3};
        if(doMat) {
#pragma unroll
          for(int b = 0; b < 6; ++b) {
            const double udg = ux * gx[b] + uy * gy[b];
#pragma unroll
            for(int cj = 0; cj < 2; ++cj) {
              const double v = ((ci == cj) ? psi * udg : 0.) + psi * ps[b] * gu[cj * 2 + ci];
              Ac[2 * b + cj] += cc * v * jw;
            }
          }
        }
        if(doRes) {
          const double ugu = ci == 0 ? (ux * g0 + uy * g2) : (ux * g1 + uy * g3);
          Be -= cc * psi * ugu * jw;
        }
      }
      if(a.hasStr) {
        const double cs = feEvalFn(a.cStr, x, e, nq, k);
        const double mu = feEvalFn(a.mu, x, e, nq, k);
        if(doMat) {
#pragma unroll
          for(int b = 0; b < 6; ++b) {
            const double gg = gi[0] * gx[b] + gi[1] * gy[b];
            const double gb[2] = {gx[b], gy[b]};
#pragma unroll
            for(int cj = 0; cj < 2; ++cj) {
              const double v = ((ci == cj) ? gg : 0.) + gi[cj] * gb[ci];
              As[2 * b + cj] += -cs * mu * v * jw;
            }
          }
#pragma unroll
          for(int j = 0; j < 3; ++j) As[12 + j] += cs * pp[j] * gi[ci] * jw;
        }
        if(doRes) {
          // sym = grad u + grad u^T; contraction with grad phi_i picks column ci
          const double s0c = ci == 0 ? 2. * g0 : g1 + g2;   // sym[0][ci]
          const double s1c = ci == 0 ? g2 + g1 : 2. * g3;   // sym[1][ci]
          const double dcs = gi[0] * s0c + gi[1] * s1c;
          Be -= cs * (pq * gi[ci] - mu * dcs) * jw;
        }
      }
      if(a.hasSrc && doRes) {
        const double S = ci == 0 ? feEvalFn(a.src0, x, e, nq, k) : feEvalFn(a.src1, x, e, nq, k);
        Be -= S * psi * jw;
      }
      if(a.hasVMass) {
        const double cm = feEvalFn(a.cVMass, x, e, nq, k);
        if(doMat) {
#pragma unroll
          for(int b = 0; b < 6; ++b) Ac[2 * b + ci] += cm * psi * ps[b] * jw;
        }
        if(doRes) Be -= cm * (ci == 0 ? ux : uy) * psi * jw;
      }
      if(a.hasGSrc && doRes) Be += feEvalFn(a.gSrc, x, e, nq, k) * gi[ci] * jw;  // S div(phi_i)
      if(a.hasVDiff) {
        // grad(phi_i) : grad(phi_j) = delta(c_i, c_j) grad(psi_a) . grad(psi_b)
        const double cd = feEvalFn(a.cVDiff, x, e, nq, k) * feEvalFn(a.dVDiff, x, e, nq, k);
        if(doMat) {
#pragma unroll
          for(int b = 0; b < 6; ++b) Ac[2 * b + ci] += cd * (gi[0] * gx[b] + gi[1] * gy[b]) * jw;
        }
        if(doRes) {
          const double gdu = gi[0] * (ci == 0 ? g0 : g1) + gi[1] * (ci == 0 ? g2 : g3);
          Be -= cd * gdu * jw;
        }
      }
      if(a.hasGrad) {
        // div(phi_i) = d psi_a / d x_ci
        const double cg = feEvalFn(a.cGrad, x, e, nq, k);
        if(doMat) {
#pragma unroll
          for(int j = 0; j < 3; ++j) Ag[j] -= cg * pp[j] * gi[ci] * jw;
        }
        if(doRes) Be += cg * pq * gi[ci] * jw;
      }
      if(a.hasMass) {
        // phi_i . phi_j = psi_a psi_b delta(c_i, c_j)
        const double cm = feEvalFn(a.cMass, x, e, nq, k);
        if(doMat) {
          const double f = cm * a.cn[0] * psi * jw;
#pragma unroll
          for(int b = 0; b < 6; ++b) Ac[2 * b + ci] += f * ps[b];
        }
        if(doRes) {
          double udq = 0.;
#pragma unroll
          for(int b = 0; b < 6; ++b) udq += ps[b] * sud[le][2 * b + ci];
          Be -= cm * udq * psi * jw;
        }
      }
    } else if(a.hasDiv) {
      const double cd = feEvalFn(a.cDiv, x, e, nq, k);
      if(doMat) {
#pragma unroll
        for(int b = 0; b < 6; ++b) {
          Ad[2 * b] += cd * pp[ip] * gx[b] * jw;
          Ad[2 * b + 1] += cd * pp[ip] * gy[b] * jw;
        }
      }
      if(doRes) Be -= cd * (g0 + g3) * pp[ip] * jw;
    }
  }
  if(doRes && (uRow || hasP) && myDof < a.n) a.rhs[myDof] += Be;
  if(!doMat) return;
  if(uRow) {
    if(a.hasConv || a.hasMass || a.hasVDiff || a.hasVMass)
#pragma unroll
      for(int j = 0; j < 12; ++j) {
        const int p = a.posUU[(size_t)(lane * 12 + j) * a.E + ee];
        if(p >= 0) a.val[p] += Ac[j];
      }
    if(a.hasGrad)
#pragma unroll
      for(int j = 0; j < 3; ++j) {
        const int p = a.posGrad[(size_t)(lane * 3 + j) * a.E + ee];
        if(p >= 0) a.val[p] += Ag[j];
      }
    if(a.hasStr)
#pragma unroll
      for(int j = 0; j < 15; ++j) {
        const int p = a.posStr[(size_t)(lane * 15 + j) * a.E + ee];
        if(p >= 0) a.val[p] += As[j];
      }
  } else if(a.hasDiv) {
#pragma unroll
    for(int j = 0; j < 12; ++j) {
      const int p = a.posDiv[(size_t)(ip * 12 + j) * a.E + ee];
      if(p >= 0) a.val[p] += Ad[j];
    }
  }
}


// ---------------------------------------------------------------------------
// Diffusion (+ source) in GEMM form on the matrix cores.  On a straight
// simplex the local matrix is Ae_ij = sum_{q,t} c_{q,t}(e) S_{q,t,ij} with
// c_{q,t} = k(x_q) J M_t (M_t the unique entries of the inverse-Jacobian
// metric) and S the reference tensors (weights folded in).  A wavefront
// takes 16 elements: A = C (16 elements x K), B = S (K x 16 entries per
// tile), one v_mfma_*_16x16x4 per 4 columns of K and per 16-entry tile.
// The lane holding D[i][j] in register v differs between the two
// instructions: v_mfma_f32_16x16x4f32 has i = 4 (lane/16) + v, while
// v_mfma_f64_16x16x4f64 has i = lane/16 + 4 v (j = lane % 16 for both).  FP64 (value parity with the reference loop <= 1e-12) or FP32
// (element matrices for the FP32 inner solve).
typedef double feDouble4 __attribute__((ext_vector_type(4)));
typedef float feFloat4 __attribute__((ext_vector_type(4)));

template <typename T>
struct FeMfma;
template <>
struct FeMfma<double> {
  using V4 = feDouble4;
  static __device__ __forceinline__ int row(int quad, int v) { return quad + 4 * v; }
  static __device__ __forceinline__ V4 mma(double x, double y, V4 c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
};
template <>
struct FeMfma<float> {
  using V4 = feFloat4;
  static __device__ __forceinline__ int row(int quad, int v) { return 4 * quad + v; }
  static __device__ __forceinline__ V4 mma(float x, float y, V4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
};

struct MfmaAsmArgs {
  int begin, count, E, n, what, K;
  const int *perm, *conn, *adr, *pos;
  const double *xyz, *tab, *u;
  const void *sref;
  FeFn coeff, src;
  int hasSrc;
  void *val;
  double *rhs;
};

template <typename T, int NF, int DIM>
__global__ __launch_bounds__(256) void asmDiffusionMfma(MfmaAsmArgs a)
{
  constexpr int NV = DIM + 1, NE = NF * NF, NT = (NE + 15) / 16, NTERM = DIM == 2 ? 3 : 6;
  __shared__ FeRefTablesFor<NF> Tb;
  __shared__ T aeLds[4][16][NE];
  {
    const double *src = a.tab;
    double *dst = (double *)&Tb;
    for(int i = threadIdx.x; i < (int)(sizeof(Tb) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int elemBase = (blockIdx.x * 4 + wave) * 16;
  const int le = lane & 15, kq = lane >> 4;
  const bool valid = elemBase + le < a.count;
  const int ee = a.begin + (valid ? elemBase + le : 0);
  const int e = a.perm[ee];
  double c[NV][3];
#pragma unroll
  for(int j = 0; j < NV; ++j) {
    const int v = a.conn[(size_t)e * NV + j];
#pragma unroll
    for(int d = 0; d < 3; ++d) c[j][d] = a.xyz[3 * (size_t)v + d];
  }
  double J, drdx[3], drdy[3], drdz[3];
  feGeometry<DIM>(c, J, drdx, drdy, drdz);
  double M[6];
  {
    const double g[3][3] = {{drdx[0], drdy[0], drdz[0]}, {drdx[1], drdy[1], drdz[1]}, {drdx[2], drdy[2], drdz[2]}};
    const int ta[6] = {0, 0, DIM == 2 ? 1 : 0, 1, 1, 2}, tb[6] = {0, 1, DIM == 2 ? 1 : 2, 1, 2, 2};
#pragma unroll
    for(int t = 0; t < NTERM; ++t) {
      double s = 0.;
#pragma unroll
      for(int d = 0; d < DIM; ++d) s += g[ta[t]][d] * g[tb[t]][d];
      M[t] = s;
    }
  }
  const int nq = Tb.nq;
  const T *sref = (const T *)a.sref;
  typename FeMfma<T>::V4 acc[NT];
#pragma unroll
  for(int t = 0; t < NT; ++t) acc[t] = (typename FeMfma<T>::V4){0, 0, 0, 0};
  for(int s = 0; s < a.K / 4; ++s) {
    const int kk = 4 * s + kq;
    const int q = kk / NTERM, tt = kk - q * NTERM;
    T av = (T)0;
    if(valid && q < nq) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * Tb.Lg[q][j];
      av = (T)(feEvalFn(a.coeff, x, e, nq, q) * J * M[tt]);
    }
#pragma unroll
    for(int t = 0; t < NT; ++t) {
      const T bv = sref[(size_t)kk * 16 * NT + t * 16 + le];
      acc[t] = FeMfma<T>::mma(av, bv, acc[t]);
    }
  }
  // scatter the matrix: lane holds entry (t*16 + le) of elements row(kq, v)
  if(a.what & 1) {
    T *val = (T *)a.val;
#pragma unroll
    for(int t = 0; t < NT; ++t) {
      const int entry = t * 16 + le;
      if(entry >= NE) continue;
#pragma unroll
      for(int v = 0; v < 4; ++v) {
        const int el = elemBase + FeMfma<T>::row(kq, v);
        if(el >= a.count) continue;
        const int p = a.pos[(size_t)entry * a.E + a.begin + el];
        if(p >= 0) val[p] += acc[t][v];
      }
    }
  }
  if(!(a.what & 2)) return;
  // residual Be = -Ae u_e - source, through LDS
#pragma unroll
  for(int t = 0; t < NT; ++t) {
    const int entry = t * 16 + le;
    if(entry < NE)
#pragma unroll
      for(int v = 0; v < 4; ++v) aeLds[wave][FeMfma<T>::row(kq, v)][entry] = acc[t][v];
  }
  __syncthreads();
  for(int idx = lane; idx < 16 * NF; idx += 64) {
    const int el = idx / NF, i = idx - el * NF;
    if(elemBase + el >= a.count) continue;
    const int eeL = a.begin + elemBase + el;
    const int *ad = a.adr + (size_t)eeL * NF;
    double be = 0.;
#pragma unroll
    for(int j = 0; j < NF; ++j) be -= (double)aeLds[wave][el][i * NF + j] * a.u[ad[j]];
    if(a.hasSrc) {
      const int eL = a.perm[eeL];
      double cc[NV][3];
#pragma unroll
      for(int j = 0; j < NV; ++j) {
        const int v = a.conn[(size_t)eL * NV + j];
#pragma unroll
        for(int d = 0; d < 3; ++d) cc[j][d] = a.xyz[3 * (size_t)v + d];
      }
      double JL, r0[3], r1[3], r2[3];
      feGeometry<DIM>(cc, JL, r0, r1, r2);
      for(int q = 0; q < nq; ++q) {
        double x[3] = {0., 0., 0.};
#pragma unroll
        for(int j = 0; j < NV; ++j)
#pragma unroll
          for(int d = 0; d < 3; ++d) x[d] += cc[j][d] * Tb.Lg[q][j];
        be -= Tb.L[q][i] * feEvalFn(a.src, x, eL, nq, q) * JL * Tb.w[q];
      }
    }
    const int row = ad[i];
    if(row < a.n) a.rhs[row] += be;
  }
}

// ---------------------------------------------------------------------------
// Taylor-Hood Jacobian of the convection, Newtonian-stress and divergence
// forms on the matrix cores (matrix-only sweeps).  With B_q the 7 x 15 basis
// quantities at quadrature point q -- rows: the two velocity values, the four
// velocity gradients d/dx_m of component n at 2 + 2 m + n, the pressure
// value; columns: the 12 velocity and 3 pressure functions -- and D_q the
// 7 x 7 coupling of the forms there (test quantity r, trial quantity s, the
// quadrature weight and Jacobian folded in), the element matrix is
// A = sum_q B_q^T D_q B_q: one v_mfma_f64_16x16x4f64 tile per element over
// K = 7 nq, A operand L[(q,s)][i] = sum_r D_q[r][s] B_q[r][i] and B operand
// B_q[s][j] formed per lane.  One wavefront per element; lanes < nq build the
// D_q in LDS.
constexpr int kNsMfmaQ = 16;  // quadrature points the LDS couplings hold

__global__ __launch_bounds__(256) void asmNsMfma(NsAsmArgs a)
{
  __shared__ FeRefTables T;
  __shared__ double Dq[4][kNsMfmaQ][7][7];
  __shared__ double su[4][16], sx[4][6], sg[4][5];
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(FeRefTables) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + w;
  const bool valid = t < a.count;
  const int ee = a.begin + (valid ? t : 0);
  const int e = a.perm[ee];
  if(valid) {
    if(l < 12) su[w][l] = a.u[a.adrU[(size_t)ee * 12 + l]];
    else if(l < 15) su[w][l] = a.u[a.adrP[(size_t)ee * 3 + l - 12]];
    if(l < 6) sx[w][l] = a.xyz[3 * (size_t)a.conn[(size_t)e * 3 + l / 2] + (l & 1)];
  }
  __syncthreads();
  if(valid && l == 0) {
    const double *c = sx[w];
    const double dxdr0 = -c[0] + c[2], dxdr1 = -c[1] + c[3];
    const double dxds0 = -c[0] + c[4], dxds1 = -c[1] + c[5];
    const double J = dxdr0 * dxds1 - dxdr1 * dxds0;
    sg[w][0] = J;
    sg[w][1] = dxds1 / J;
    sg[w][2] = -dxdr1 / J;
    sg[w][3] = -dxds0 / J;
    sg[w][4] = dxdr0 / J;
  }
  __syncthreads();
  const int nq = T.nq;
  if(valid && l < nq) {
    const int q = l;
    const double J = sg[w][0], drdx0 = sg[w][1], drdx1 = sg[w][2], drdy0 = sg[w][3], drdy1 = sg[w][4];
    double ux = 0., uy = 0., g[4] = {0., 0., 0., 0.};  // g[m * 2 + n] = d u_n / d x_m
#pragma unroll
    for(int b = 0; b < 6; ++b) {
      const double gx = T.dLr[q][b] * drdx0 + T.dLs[q][b] * drdx1, gy = T.dLr[q][b] * drdy0 + T.dLs[q][b] * drdy1;
      const double vx = su[w][2 * b], vy = su[w][2 * b + 1];
      ux += vx * T.L[q][b];
      uy += vy * T.L[q][b];
      g[0] += vx * gx;
      g[1] += vy * gx;
      g[2] += vx * gy;
      g[3] += vy * gy;
    }
    double x[3] = {0., 0., 0.};
#pragma unroll
    for(int j = 0; j < 3; ++j) {
      x[0] += sx[w][2 * j] * T.Lg[q][j];
      x[1] += sx[w][2 * j + 1] * T.Lg[q][j];
    }
    const double jw = J * T.w[q];
    double D[7][7];
#pragma unroll
    for(int r = 0; r < 7; ++r)
#pragma unroll
      for(int c = 0; c < 7; ++c) D[r][c] = 0.;
    if(a.hasConv) {  // phi_i . ((u . grad) phi_j + (phi_j . grad) u)
      const double cc = feEvalFn(a.cConv, x, e, nq, q) * jw;
#pragma unroll
      for(int ci = 0; ci < 2; ++ci) {
        D[ci][2 + ci] += cc * ux;
        D[ci][4 + ci] += cc * uy;
#pragma unroll
        for(int cj = 0; cj < 2; ++cj) D[ci][cj] += cc * g[cj * 2 + ci];
      }
    }
    {  // -mu grad phi_i : (grad phi_j + grad phi_j^T) and p div phi_i
      const double cs = feEvalFn(a.cStr, x, e, nq, q), f = -cs * feEvalFn(a.mu, x, e, nq, q) * jw;
#pragma unroll
      for(int m = 0; m < 2; ++m)
#pragma unroll
        for(int n = 0; n < 2; ++n) {
          D[2 + 2 * m + n][2 + 2 * m + n] += f;
          D[2 + 2 * n + m][2 + 2 * m + n] += f;
        }
      D[2][6] += cs * jw;
      D[5][6] += cs * jw;
    }
    {  // psi_k div phi_j
      const double cd = feEvalFn(a.cDiv, x, e, nq, q) * jw;
      D[6][2] += cd;
      D[6][5] += cd;
    }
#pragma unroll
    for(int r = 0; r < 7; ++r)
#pragma unroll
      for(int c = 0; c < 7; ++c) Dq[w][q][r][c] = D[r][c];
  }
  __syncthreads();
  if(!valid) return;
  const double drdx0 = sg[w][1], drdx1 = sg[w][2], drdy0 = sg[w][3], drdy1 = sg[w][4];
  const int i = l & 15, kq = l >> 4;  // A row / B column i; k offset kq
  const int ai = i >> 1, ci = i & 1;
  feDouble4 acc = {0., 0., 0., 0.};
  const int K = 7 * nq;
  for(int k0 = 0; k0 < K; k0 += 4) {
    const int kk = k0 + kq;
    double av = 0., bv = 0.;
    if(kk < K && i < 15) {
      const int q = kk / 7, s = kk - 7 * q;
      if(i < 12) {
        const double ps = T.L[q][ai];
        const double gx = T.dLr[q][ai] * drdx0 + T.dLs[q][ai] * drdx1;
        const double gy = T.dLr[q][ai] * drdy0 + T.dLs[q][ai] * drdy1;
        av = Dq[w][q][ci][s] * ps + Dq[w][q][2 + ci][s] * gx + Dq[w][q][4 + ci][s] * gy;
        if(s < 2) bv = ci == s ? ps : 0.;
        else if(s < 6) bv = ci == ((s - 2) & 1) ? ((s - 2) >> 1 == 0 ? gx : gy) : 0.;
      } else {
        const double pp = T.Lg[q][i - 12];
        av = Dq[w][q][6][s] * pp;
        bv = s == 6 ? pp : 0.;
      }
    }
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for(int v = 0; v < 4; ++v) {
    const int row = kq + 4 * v, col = i;
    if(row >= 15 || col >= 15 || (row >= 12 && col >= 12)) continue;
    const int p = row < 12 ? a.posStr[(size_t)(row * 15 + col) * a.E + ee]
                           : a.posDiv[(size_t)((row - 12) * 12 + col) * a.E + ee];
    if(p >= 0) a.val[p] += acc[v];
  }
}

template <typename T, int NF, int DIM>
void launchMfma(fe_hip_ctx *ctx, MfmaAsmArgs a)
{
  for(size_t c = 0; c + 1 < ctx->colorStart.size(); ++c) {
    a.begin = ctx->colorStart[c];
    a.count = ctx->colorStart[c + 1] - a.begin;
    if(a.count == 0) continue;
    hipLaunchKernelGGL((asmDiffusionMfma<T, NF, DIM>), dim3((a.count + 63) / 64), dim3(256), 0, ctx->stream, a);
  }
}

__global__ void toFloat(int64_t n, const double *__restrict__ x, float *__restrict__ y)
{
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

// Per-block max |x| (plain stores), folded by maxFinal after the kernel
// boundary -- no device-scope fences inside the kernels.
__global__ void maxAbsBlocks(int n, const double *__restrict__ x, double *__restrict__ partials)
{
  __shared__ double lds[16];
  double m = 0.;
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, fabs(x[i]));
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if(lane == 0) lds[wid] = m;
  __syncthreads();
  if(threadIdx.x == 0) {
    double b = 0.;
    for(int w = 0; w < (int)(blockDim.x >> 6); ++w) b = fmax(b, lds[w]);
    partials[blockIdx.x] = b;
  }
}

// One wavefront folds the block maxima (strided loads in flight together,
// then a shuffle tree); launched with 64 threads.
__global__ void maxFinal(const double *__restrict__ partials, int G, double *__restrict__ out)
{
  double r = 0.;
  for(int b = threadIdx.x; b < G; b += 64) r = fmax(r, partials[b]);
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
  if(threadIdx.x == 0) *out = r;
}

__global__ void addCorrection(int n, double *__restrict__ u, const double *__restrict__ du)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) u[i] += du[i];
}

template <int NF, int DIM>
void launchScalar(fe_hip_ctx *ctx, ScalarAsmArgs a)
{
  constexpr int LPE = NF <= 4 ? 4 : 8;
  for(size_t c = 0; c + 1 < ctx->colorStart.size(); ++c) {
    a.begin = ctx->colorStart[c];
    a.count = ctx->colorStart[c + 1] - a.begin;
    if(a.count == 0) continue;
    if constexpr(NF <= 6) {
      if(ctx->scalarLanes && !a.nlDiff) {
        hipLaunchKernelGGL((asmScalarLanes<NF, DIM>), dim3((a.count * LPE + 255) / 256), dim3(256), 0, ctx->stream,
                           a);
        continue;
      }
    }
    // the advection, mass-power and nonlinear-diffusion terms are compiled
    // out of the common variant (fewer registers, more waves per SIMD)
    if(a.hasAdv || a.hasReact || a.nlDiff)
      hipLaunchKernelGGL((asmScalarFP64<NF, DIM, true>), dim3((a.count + 255) / 256), dim3(256), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((asmScalarFP64<NF, DIM, false>), dim3((a.count + 255) / 256), dim3(256), 0, ctx->stream,
                         a);
  }
}

template <int NF, int DIM>
void launchMixed(fe_hip_ctx *ctx, MixedAsmArgs a)
{
  for(size_t c = 0; c + 1 < ctx->colorStart.size(); ++c) {
    a.begin = ctx->colorStart[c];
    a.count = ctx->colorStart[c + 1] - a.begin;
    if(a.count == 0) continue;
    hipLaunchKernelGGL((asmMixedScalarFP64<NF, DIM>), dim3((a.count + 255) / 256), dim3(256), 0, ctx->stream, a);
  }
}

}  // namespace

extern "C" int fe_hip_set_to_zero(fe_hip_ctx *ctx, int matrix, int residual)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(matrix) FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_val, 0, sizeof(double) * ctx->nnz, ctx->stream));
  if(matrix && ctx->d_val32) FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_val32, 0, sizeof(float) * ctx->nnz, ctx->stream));
  if(matrix) ctx->matrixZeroed = true;
  if(residual) FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_rhs, 0, sizeof(double) * ctx->n, ctx->stream));
  return FE_STATUS_OK;  // stream-ordered before the next assembly
}

extern "C" int fe_hip_assemble(fe_hip_ctx *ctx, int what, int precision)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  FE_REQUIRE(ctx, what >= 1 && what <= 3, "what must be 1 (matrix), 2 (residual) or 3 (both)");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  // the sweep is bracketed by two events and not waited for: its time is
  // read when fe_hip_kernel_stats asks (the next host read waits anyway)
  if(!ctx->asmEv[0]) {
    FE_HIP_TRY(ctx, hipEventCreate(&ctx->asmEv[0]));
    FE_HIP_TRY(ctx, hipEventCreate(&ctx->asmEv[1]));
  }
  FE_HIP_TRY(ctx, hipEventRecord(ctx->asmEv[0], ctx->stream));
  ctx->asmPending = false;
  // Scalar forms of one field and quadrature rule (at most one DIFFUSION or
  // NONLINEAR_DIFFUSION, TRANSIENT_MASS, MASS, ADVECTION and SOURCE each)
  // are fused: one element sweep computes Ae and Be of all of them; a second
  // form of a kind gets a sweep of its own.
  std::vector<bool> done(ctx->forms.size(), false);
  // FP32 values (FE_PREC_FP32_IR): the final FP32 rounding pass is skipped
  // only when the matrix started from zero and every matrix sweep of this
  // call also wrote the FP32 copy; any sweep that writes FP64 values only
  // clears allFp32.
  const bool fromZero = ctx->matrixZeroed;
  if(what & 1) ctx->matrixZeroed = false;
  bool allFp32 = true;
  bool valClean = fromZero;  // no sweep of this call has written the values yet
  const bool ir32 = precision == FE_PREC_FP32_IR && (what & 1);
  auto ensureVal32 = [&]() -> int {
    if(ctx->d_val32) return FE_STATUS_OK;
    if(feAlloc(ctx, &ctx->d_val32, (size_t)ctx->nnz)) return FE_STATUS_OUT_OF_MEMORY;
    FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_val32, 0, sizeof(float) * ctx->nnz, ctx->stream));
    return FE_STATUS_OK;
  };
  auto scalarForm = [](int t) {
    return t == FE_SYS_DIFFUSION || t == FE_SYS_NONLINEAR_DIFFUSION || t == FE_SYS_TRANSIENT_MASS || t == FE_SYS_MASS || t == FE_SYS_ADVECTION ||
           t == FE_SYS_SOURCE;
  };
  for(size_t fi = 0; fi < ctx->forms.size(); ++fi) {
    if(done[fi] || !scalarForm(ctx->forms[fi].sysType)) continue;
    const int field = ctx->forms[fi].fields[0], degree = ctx->forms[fi].degree;
    int iDiff = -1, iMass = -1, iReact = -1, iAdv = -1, iSrc = -1;
    for(size_t fj = fi; fj < ctx->forms.size(); ++fj) {
      const FeForm &G = ctx->forms[fj];
      if(done[fj] || !scalarForm(G.sysType) || G.fields[0] != field || G.degree != degree) continue;
      int &slot = G.sysType == FE_SYS_DIFFUSION || G.sysType == FE_SYS_NONLINEAR_DIFFUSION ? iDiff
                  : G.sysType == FE_SYS_TRANSIENT_MASS ? iMass
                  : G.sysType == FE_SYS_MASS           ? iReact
                  : G.sysType == FE_SYS_ADVECTION      ? iAdv
                                                       : iSrc;
      if(slot < 0) {
        slot = (int)fj;
        done[fj] = true;
      }
    }
    ScalarAsmArgs a{};
    a.E = ctx->E;
    a.n = ctx->n;
    a.what = what;
    a.perm = ctx->d_perm;
    a.conn = ctx->d_conn;
    a.xyz = ctx->d_xyz;
    a.u = ctx->d_u;
    a.val = ctx->d_val;
    a.rhs = ctx->d_rhs;
    a.tab = ctx->forms[fi].d_tab;
    // Row map and scatter positions: the group of a matrix form of the sweep,
    // else (a lone source) a matrix group on the same field.
    int g = iDiff >= 0 ? ctx->forms[iDiff].group
            : iMass >= 0 ? ctx->forms[iMass].group
            : iReact >= 0 ? ctx->forms[iReact].group
            : iAdv >= 0   ? ctx->forms[iAdv].group
                          : -1;
    for(size_t k = 0; g < 0 && k < ctx->groups.size(); ++k)
      if(ctx->groups[k].layoutI.size() == 1 && ctx->groups[k].layoutI[0] == field) g = (int)k;
    FE_REQUIRE(ctx, g >= 0, "a SOURCE form needs a matrix form (DIFFUSION, TRANSIENT_MASS, MASS) on the same field");
    a.adr = ctx->groups[g].d_adrI;
    a.pos = ctx->groups[g].d_pos;
    if(iDiff >= 0) {
      a.hasDiff = 1;
      a.coeff = ctx->forms[iDiff].coeffs[0];
      if(ctx->forms[iDiff].sysType == FE_SYS_NONLINEAR_DIFFUSION) {
        a.nlDiff = 1;
        a.dcoeff = ctx->forms[iDiff].coeffs[1];
      }
    }
    if(iMass >= 0) {
      a.hasMass = 1;
      a.mass = ctx->forms[iMass].coeffs[0];
      a.nsol = ctx->nsol;
      for(int j = 0; j <= fe_hip_ctx::kMaxPast; ++j) a.cn[j] = ctx->bdf[j];
      for(int j = 0; j < fe_hip_ctx::kMaxPast; ++j) a.past[j] = ctx->d_past[j];
    }
    if(iReact >= 0) {
      a.hasReact = 1;
      a.react = ctx->forms[iReact].coeffs[0];
      a.reactPow = ctx->forms[iReact].coeffs.size() > 1 ? ctx->forms[iReact].coeffs[1].c0 : 1.;
    }
    if(iAdv >= 0) {
      a.hasAdv = 1;
      for(size_t d = 0; d < ctx->forms[iAdv].coeffs.size() && d < 3; ++d) a.adv[d] = ctx->forms[iAdv].coeffs[d];
    }
    if(iSrc >= 0) {
      a.hasSrc = 1;
      a.src = ctx->forms[iSrc].coeffs[0];
    }
    if(!(what & 2) && iDiff < 0 && iMass < 0 && iReact < 0 && iAdv < 0) continue;  // a lone source has no matrix
    const int kind = ctx->fields[field].kind;
    const bool mfmaKind = kind == FE_TRI_P1 || kind == FE_TRI_P2 || kind == FE_TET_P1;
    const bool matSweep = (what & 1) && (iDiff >= 0 || iMass >= 0 || iReact >= 0 || iAdv >= 0);
    if(ctx->asmKernel == 2 && mfmaKind && !a.nlDiff && iReact < 0 && iAdv < 0) {
      bool ok = false;
      if(int rc = feGatherPrepare(ctx, field, &ok)) return rc;
      if(ok) {
        FeGatherSweep gs{};
        gs.field = field;
        gs.what = what;
        gs.nq = ctx->forms[fi].tab.nq;
        gs.tab = ctx->forms[fi].d_tab;
        gs.hasDiff = a.hasDiff;
        gs.hasSrc = a.hasSrc;
        gs.hasMass = a.hasMass;
        gs.coeff = a.coeff;
        gs.src = a.src;
        gs.mass = a.mass;
        gs.val32 = nullptr;
        gs.fp32Math = false;
        // the first matrix sweep after a zeroing stores its sums outright
        // (later sweeps of the call add to them, stream-ordered)
        gs.freshVal = matSweep && valClean;
        if(matSweep) valClean = false;
        if(ir32 && matSweep) {
          if(int rc = ensureVal32()) return rc;
          gs.val32 = ctx->d_val32;
          gs.fp32Math = ctx->asmFp32Math != 0;
        }
        if(int rc = feGatherLaunch(ctx, gs)) return rc;
        continue;
      }
    }
    const bool allBefore = allFp32;
    if(matSweep) valClean = false;
    if(matSweep) allFp32 = false;  // re-set below when the MFMA pair writes both copies
    const bool mfmaHigh = kind == FE_TRI_P3 || kind == FE_TRI_P4;  // cubic / quartic: colored MFMA sweep
    if(iDiff >= 0 && !a.nlDiff && iMass < 0 && iReact < 0 && iAdv < 0 && ctx->asmKernel >= 1 &&
       (mfmaKind || mfmaHigh)) {
      const FeForm &F = ctx->forms[iDiff];
      MfmaAsmArgs m{};
      m.E = a.E; m.n = a.n; m.what = what; m.K = F.mfmaK;
      m.perm = a.perm; m.conn = a.conn; m.adr = a.adr; m.pos = a.pos;
      m.xyz = a.xyz; m.tab = a.tab; m.u = a.u; m.sref = F.d_sref64;
      m.coeff = a.coeff; m.src = a.src; m.hasSrc = a.hasSrc;
      m.val = ctx->d_val; m.rhs = ctx->d_rhs;
      if(kind == FE_TRI_P1) launchMfma<double, 3, 2>(ctx, m);
      else if(kind == FE_TRI_P2) launchMfma<double, 6, 2>(ctx, m);
      else if(kind == FE_TET_P1) launchMfma<double, 4, 3>(ctx, m);
      else if(kind == FE_TRI_P3) launchMfma<double, 10, 2>(ctx, m);
      else launchMfma<double, 15, 2>(ctx, m);
      if(precision == FE_PREC_FP32_IR && (what & 1) && ctx->fp32Values == 1) {
        if(int rc = ensureVal32()) return rc;
        m.what = 1;
        m.sref = F.d_sref32;
        m.val = ctx->d_val32;
        if(kind == FE_TRI_P1) launchMfma<float, 3, 2>(ctx, m);
        else if(kind == FE_TRI_P2) launchMfma<float, 6, 2>(ctx, m);
        else if(kind == FE_TET_P1) launchMfma<float, 4, 3>(ctx, m);
        else if(kind == FE_TRI_P3) launchMfma<float, 10, 2>(ctx, m);
        else launchMfma<float, 15, 2>(ctx, m);
        allFp32 = allBefore;  // this sweep wrote both copies
      }
    } else if(kind == FE_TRI_P1) launchScalar<3, 2>(ctx, a);
    else if(kind == FE_TRI_P2) launchScalar<6, 2>(ctx, a);
    else if(kind == FE_TET_P1) launchScalar<4, 3>(ctx, a);
    else if(kind == FE_TRI_P3) launchScalar<10, 2>(ctx, a);
    else if(kind == FE_TRI_P4) launchScalar<15, 2>(ctx, a);
    else FE_REQUIRE(ctx, false, "scalar forms need a scalar field");
  }
  // Mixed scalar forms of one (test, trial) field pair and quadrature rule
  // (at most one MIXED_MASS, MIXED_MASS_POWER and MIXED_GRADGRAD each) are
  // fused into one sweep; a second form of a kind gets a sweep of its own.
  auto mixedForm = [](int t) {
    return t == FE_SYS_MIXED_MASS || t == FE_SYS_MIXED_MASS_POWER || t == FE_SYS_MIXED_GRADGRAD;
  };
  for(size_t fi = 0; fi < ctx->forms.size(); ++fi) {
    if(done[fi] || !mixedForm(ctx->forms[fi].sysType)) continue;
    const FeForm &F0 = ctx->forms[fi];
    int iM = -1, iP = -1, iG = -1;
    for(size_t fj = fi; fj < ctx->forms.size(); ++fj) {
      const FeForm &G = ctx->forms[fj];
      if(done[fj] || !mixedForm(G.sysType) || G.fields != F0.fields || G.degree != F0.degree) continue;
      int &slot = G.sysType == FE_SYS_MIXED_MASS ? iM : G.sysType == FE_SYS_MIXED_MASS_POWER ? iP : iG;
      if(slot < 0) {
        slot = (int)fj;
        done[fj] = true;
      }
    }
    MixedAsmArgs a{};
    a.E = ctx->E;
    a.n = ctx->n;
    a.what = what;
    a.perm = ctx->d_perm;
    a.conn = ctx->d_conn;
    a.xyz = ctx->d_xyz;
    a.u = ctx->d_u;
    a.val = ctx->d_val;
    a.rhs = ctx->d_rhs;
    a.tab = F0.d_tab;
    const FeGroup &grp = ctx->groups[F0.group];
    a.adrI = grp.d_adrI;
    a.adrJ = grp.d_adrJ;
    a.pos = grp.d_pos;
    if(iM >= 0) {
      a.hasMass = 1;
      a.cMass = ctx->forms[iM].coeffs[0];
    }
    if(iP >= 0) {
      a.hasPow = 1;
      a.cPow = ctx->forms[iP].coeffs[0];
      a.pw = ctx->forms[iP].coeffs[1].c0;
    }
    if(iG >= 0) {
      a.hasGG = 1;
      a.cGG = ctx->forms[iG].coeffs[0];
    }
    const int kind = ctx->fields[F0.fields[0]].kind;
    if(what & 1) allFp32 = false;
    if(kind == FE_TRI_P1) launchMixed<3, 2>(ctx, a);
    else if(kind == FE_TRI_P2) launchMixed<6, 2>(ctx, a);
    else if(kind == FE_TET_P1) launchMixed<4, 3>(ctx, a);
    else if(kind == FE_TRI_P3) launchMixed<10, 2>(ctx, a);
    else if(kind == FE_TRI_P4) launchMixed<15, 2>(ctx, a);
    else FE_REQUIRE(ctx, false, "mixed scalar forms need scalar fields");
  }
  // Taylor-Hood stack: all vector/mixed forms fused into one sweep per color
  // (one quadrature rule: the forms must share their degree).
  {
    NsAsmArgs a{};
    int any = 0, degree = -1, thNq = 0;
    const int *posConv = nullptr, *posMass = nullptr, *posVDiff = nullptr, *posVMass = nullptr;
    const int *adrPgrad = nullptr;
    for(size_t fi = 0; fi < ctx->forms.size(); ++fi) {
      FeForm &F = ctx->forms[fi];
      if(done[fi]) continue;
      switch(F.sysType) {
        case FE_SYS_VECTOR_CONVECTIVE_ACCELERATION:
          FE_REQUIRE(ctx, !a.hasConv, "two convective forms");
          a.hasConv = 1; a.cConv = F.coeffs[0]; posConv = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_TRANSIENT_VECTOR_MASS:
          FE_REQUIRE(ctx, !a.hasMass, "two vector transient mass forms");
          a.hasMass = 1; a.cMass = F.coeffs[0]; posMass = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_VECTOR_MASS:
          FE_REQUIRE(ctx, !a.hasVMass, "two vector mass forms");
          a.hasVMass = 1; a.cVMass = F.coeffs[0]; posVMass = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_GRAD_SOURCE:
          FE_REQUIRE(ctx, !a.hasGSrc, "two gradient sources");
          a.hasGSrc = 1; a.gSrc = F.coeffs[0];
          break;
        case FE_SYS_VECTOR_DIFFUSION:
          FE_REQUIRE(ctx, !a.hasVDiff, "two vector diffusion forms");
          a.hasVDiff = 1; a.cVDiff = F.coeffs[0]; a.dVDiff = F.coeffs[1]; posVDiff = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_MIXED_GRADIENT:
          FE_REQUIRE(ctx, !a.hasGrad, "two gradient forms");
          a.hasGrad = 1; a.cGrad = F.coeffs[0]; a.posGrad = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI; adrPgrad = ctx->groups[F.group].d_adrJ;
          break;
        case FE_SYS_MIXED_DIVERGENCE:
          FE_REQUIRE(ctx, !a.hasDiv, "two divergence forms");
          a.hasDiv = 1; a.cDiv = F.coeffs[0]; a.posDiv = ctx->groups[F.group].d_pos;
          a.adrP = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_DIV_NEWTONIAN_STRESS:
          FE_REQUIRE(ctx, !a.hasStr, "two stress forms");
          a.hasStr = 1; a.cStr = F.coeffs[0]; a.mu = F.coeffs[1]; a.posStr = ctx->groups[F.group].d_pos;
          a.adrU = ctx->groups[F.group].d_adrI;
          break;
        case FE_SYS_VECTOR_SOURCE:
          FE_REQUIRE(ctx, !a.hasSrc, "two vector sources");
          a.hasSrc = 1; a.src0 = F.coeffs[0]; a.src1 = F.coeffs[1];
          break;
        default: continue;
      }
      FE_REQUIRE(ctx, degree < 0 || F.degree == degree, "the vector forms must share one quadrature degree");
      degree = F.degree;
      if(!a.tab) {
        a.tab = F.d_tab;
        thNq = F.tab.nq;
      }
      done[fi] = true;
      any = 1;
    }
    if(any) {
      FE_REQUIRE(ctx, a.adrU, "the vector forms need a velocity matrix form (stress, convection or transient mass)");
      FE_REQUIRE(ctx, !a.hasStr || a.hasDiv, "the stress form needs the divergence form (pressure map)");
      // convection and transient mass share the velocity-velocity layout, hence one group
      const int *uu[4] = {posConv, posMass, posVDiff, posVMass};
      for(const int *p : uu)
        if(p) {
          FE_REQUIRE(ctx, !a.posUU || a.posUU == p, "velocity forms on different groups");
          a.posUU = p;
        }
      // the pressure map: the divergence form's rows, else the gradient form's columns
      if(!a.adrP) a.adrP = adrPgrad;
      if(a.hasMass) {
        a.nsol = ctx->nsol;
        for(int j = 0; j <= fe_hip_ctx::kMaxPast; ++j) a.cn[j] = ctx->bdf[j];
        for(int j = 0; j < fe_hip_ctx::kMaxPast; ++j) a.past[j] = ctx->d_past[j];
      }
      a.E = ctx->E;
      a.n = ctx->n;
      a.what = what;
      a.perm = ctx->d_perm;
      a.conn = ctx->d_conn;
      a.xyz = ctx->d_xyz;
      a.u = ctx->d_u;
      a.val = ctx->d_val;
      a.rhs = ctx->d_rhs;
      if(what & 1) allFp32 = false;
      // the MFMA form of the Jacobian: matrix-only sweeps of exactly the
      // stress + divergence (+ convection) stack
      const bool nsMfma = ctx->nsMfma && what == 1 && a.hasStr && a.hasDiv && a.adrP && !a.hasMass && !a.hasVDiff &&
                          !a.hasGrad && !a.hasVMass && thNq <= kNsMfmaQ;
      for(size_t c = 0; c + 1 < ctx->colorStart.size(); ++c) {
        a.begin = ctx->colorStart[c];
        a.count = ctx->colorStart[c + 1] - a.begin;
        if(a.count == 0) continue;
        if(nsMfma)
          hipLaunchKernelGGL(asmNsMfma, dim3((a.count + 3) / 4), dim3(256), 0, ctx->stream, a);
        else
          hipLaunchKernelGGL(asmNavierStokesFP64, dim3((a.count + kNsElemPerBlock - 1) / kNsElemPerBlock), dim3(256),
                             0, ctx->stream, a);
      }
    }
  }
  for(size_t fi = 0; fi < ctx->forms.size(); ++fi)
    FE_REQUIRE(ctx, done[fi], "form type has no device kernel");
  FE_HIP_TRY(ctx, hipGetLastError());
  if(ir32 && !(fromZero && allFp32 && ctx->d_val32)) {
    if(!ctx->d_val32 && feAlloc(ctx, &ctx->d_val32, (size_t)ctx->nnz)) return FE_STATUS_OUT_OF_MEMORY;
    hipLaunchKernelGGL(toFloat, dim3(2048), dim3(256), 0, ctx->stream, ctx->nnz, ctx->d_val, ctx->d_val32);
  }
  FE_HIP_TRY(ctx, hipEventRecord(ctx->asmEv[1], ctx->stream));
  ctx->asmPending = true;
  return FE_STATUS_OK;
}

extern "C" int fe_hip_rhs_max_norm(fe_hip_ctx *ctx, double *norm)
{
  if(!ctx || !norm) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  double *out = ctx->d_partials + 3 * ctx->maxBlocks;
  const int nb = std::min(1024, (ctx->n + 255) / 256);
  hipLaunchKernelGGL(maxAbsBlocks, dim3(nb), dim3(256), 0, ctx->stream, ctx->n, ctx->d_rhs, ctx->d_partials);
  hipLaunchKernelGGL(maxFinal, dim3(1), dim3(64), 0, ctx->stream, ctx->d_partials, nb, out);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipMemcpyAsync(norm, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

extern "C" int fe_hip_correct_solution(fe_hip_ctx *ctx)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  hipLaunchKernelGGL(addCorrection, dim3((ctx->n + 255) / 256), dim3(256), 0, ctx->stream, ctx->n, ctx->d_u, ctx->d_du);
  FE_HIP_TRY(ctx, hipGetLastError());  // stream-ordered: the next read of u waits for it
  return FE_STATUS_OK;
}

// Times one assembly (zero + sweeps) of the current forms for every kernel
// variant that applies -- the row gather (two slice orders), the colored
// sweep with one thread per element or one lane per local row, the colored
// MFMA sweep (16 elements per wavefront) -- best of three each, and keeps
// the fastest in the context's tuning (asm_kernel, scalar_lanes, gather_xcd).
// The system holds the last variant's assembly afterwards.
extern "C" int fe_hip_autotune_assembly(fe_hip_ctx *ctx, int precision, int *asm_kernel, int *scalar_lanes,
                                        int *gather_xcd, double *best_ms)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  struct Cand {
    int kernel, lanes, xcd;
  };
  const Cand cands[] = {{2, 0, 1}, {2, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 1}};
  const Cand keep{ctx->asmKernel, ctx->scalarLanes, ctx->gatherXcd};
  Cand best = keep;
  double bestMs = 1e300;
  for(const Cand &c : cands) {
    ctx->asmKernel = c.kernel;
    ctx->scalarLanes = c.lanes;
    ctx->gatherXcd = c.xcd;
    double ms = 1e300;
    for(int rep = 0; rep < 3; ++rep) {
      if(int rc = fe_hip_set_to_zero(ctx, 1, 1)) return rc;
      if(int rc = fe_hip_assemble(ctx, 3, precision)) return rc;
      double m = 0.;
      if(int rc = fe_hip_kernel_stats(ctx, 1, &m, nullptr)) return rc;
      ms = std::min(ms, m);
    }
    if(ms < bestMs) {
      bestMs = ms;
      best = c;
    }
  }
  ctx->asmKernel = best.kernel;
  ctx->scalarLanes = best.lanes;
  ctx->gatherXcd = best.xcd;
  if(asm_kernel) *asm_kernel = best.kernel;
  if(scalar_lanes) *scalar_lanes = best.lanes;
  if(gather_xcd) *gather_xcd = best.xcd;
  if(best_ms) *best_ms = bestMs;
  return FE_STATUS_OK;
}
