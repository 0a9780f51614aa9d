// Element-level device helpers shared by the assembly and norm kernels:
// the coefficient callbacks and the straight-simplex geometry.
#pragma once

#include "fe_hip_internal.h"

#include <cmath>

static __device__ __forceinline__ double feKovLambda(double Re)
{
  return Re / 2. - sqrt(Re * Re / 4. + 4. * M_PI * M_PI);
}

// The feFunction callbacks of the configs (feFunction.h:75-110), evaluated at
// a physical quadrature point.
static __device__ __forceinline__ double feEvalFn(const FeFn &f, const double *x, int e, int nq, int q)
{
  switch(f.kind) {
    case FE_FN_CONST: return f.c0;
    case FE_FN_SINSIN: return f.c0 * sin(M_PI * x[0]) * sin(M_PI * x[1]);
    case FE_FN_BLOCKS: {
      const int B = (int)f.c0;
      int bi = (int)floor(x[0] * B), bj = (int)floor(x[1] * B);
      bi = bi >= B ? B - 1 : bi;
      bj = bj >= B ? B - 1 : bj;
      return exp(0.5 * (double)(((7 * bi + 13 * bj) % 11) - 5));
    }
    case FE_FN_SMOOTH3:
      return exp(f.c0 * sin(2. * M_PI * x[0]) * sin(2. * M_PI * x[1]) * sin(2. * M_PI * x[2]));
    case FE_FN_ELEM: return f.data[e];
    case FE_FN_QP: return f.data[(size_t)e * nq + q];
    case FE_FN_KOV_U: {
      const double l = feKovLambda(f.c0);
      if(f.c1 == 0.) return 1. - exp(l * x[0]) * cos(2. * M_PI * x[1]);
      return l / (2. * M_PI) * exp(l * x[0]) * sin(2. * M_PI * x[1]);
    }
    case FE_FN_KOV_P: {
      const double l = feKovLambda(f.c0);
      return 0.5 * (1. - exp(2. * l * x[0]));
    }
  }
  return 0.;
}

// A coefficient that may depend on the solution's value u at the point
// (feFunctionArguments::u): FE_FN_POLY_U by Horner's rule, the other kinds
// as feEvalFn.
static __device__ __forceinline__ double feEvalFnU(const FeFn &f, const double *x, int e, int nq, int q, double u)
{
  if(f.kind == FE_FN_POLY_U) {
    double v = 0.;
    for(int i = f.n - 1; i >= 0; --i) v = v * u + f.data[i];
    return v;
  }
  return feEvalFn(f, x, e, nq, q);
}

// Straight-simplex geometry: detJ (feCncGeo.cpp:318-340 / :380-396) and the
// constant inverse Jacobian (feCncGeo.cpp:651-692).
template <int DIM>
__device__ __forceinline__ void feGeometry(const double (&c)[DIM + 1][3], double &J,
                                           double (&drdx)[3], double (&drdy)[3], double (&drdz)[3])
{
  if(DIM == 2) {
    const double dxdr0 = -c[0][0] + c[1][0], dxdr1 = -c[0][1] + c[1][1];
    const double dxds0 = -c[0][0] + c[2][0], dxds1 = -c[0][1] + c[2][1];
    J = dxdr0 * dxds1 - dxdr1 * dxds0;
    drdx[0] = dxds1 / J;
    drdx[1] = -dxdr1 / J;
    drdy[0] = -dxds0 / J;
    drdy[1] = dxdr0 / J;
    drdx[2] = drdy[2] = drdz[0] = drdz[1] = drdz[2] = 0.;
  } else {
    double m[3][3];
#pragma unroll
    for(int i = 0; i < 3; ++i) {
      m[i][0] = -c[0][i] + c[1][i];
      m[i][1] = -c[0][i] + c[2][i];
      m[i][2] = -c[0][i] + c[3][i];
    }
    J = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
        m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    const double ud = 1. / J;
    drdx[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * ud;
    drdx[1] = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]) * ud;
    drdx[2] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * ud;
    drdy[0] = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]) * ud;
    drdy[1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * ud;
    drdy[2] = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]) * ud;
    drdz[0] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * ud;
    drdz[1] = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]) * ud;
    drdz[2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * ud;
  }
}

