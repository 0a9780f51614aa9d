// Orientation check of straight (P1-geometry) elements, shared by the host
// mesh library and the device context: the reference refuses a mesh with a
// non-positive Jacobian determinant on any element when it builds the
// element transformations of a space (feCncGeo::computeJacobians,
// feCncGeo.cpp:346-353 for triangles, :388-395 for tetrahedra, error at
// :410-412).  For straight elements the determinant is constant per element.
#ifndef FE_JACOBIAN_CHECK_H
#define FE_JACOBIAN_CHECK_H

#include <cstddef>

// Determinant of element e: triangles (x1 - x0, x2 - x0) in the xy plane,
// tetrahedra det[x1 - x0, x2 - x0, x3 - x0].
inline double feElementDetJ(int dim, const double *xyz, const int *v)
{
  const double *a = xyz + 3 * (size_t)v[0], *b = xyz + 3 * (size_t)v[1], *c = xyz + 3 * (size_t)v[2];
  const double r0 = b[0] - a[0], r1 = b[1] - a[1], r2 = b[2] - a[2];
  const double s0 = c[0] - a[0], s1 = c[1] - a[1], s2 = c[2] - a[2];
  if(dim == 2) return r0 * s1 - r1 * s0;
  const double *d = xyz + 3 * (size_t)v[3];
  const double t0 = d[0] - a[0], t1 = d[1] - a[1], t2 = d[2] - a[2];
  return r0 * (s1 * t2 - s2 * t1) - r1 * (s0 * t2 - s2 * t0) + r2 * (s0 * t1 - s1 * t0);
}

// First element (in element order) whose determinant is <= 0, or -1.
inline int feFirstNonPositiveJacobian(int dim, const double *xyz, int E, int nvpe, const int *conn, double *det)
{
  for(int e = 0; e < E; ++e) {
    const double J = feElementDetJ(dim, xyz, conn + (size_t)e * nvpe);
    if(!(J > 0.)) {
      if(det) *det = J;
      return e;
    }
  }
  return -1;
}

#endif
