// Lagrange triangles of degree n on equispaced nodes, host side: the node
// set of the scalar kinds in their local order, and the shape functions by
// the barycentric product formula
//   phi_(a,b,c)(l0, l1, l2) = P_a(l0) P_b(l1) P_c(l2),
//   P_m(l) = prod_{q=0}^{m-1} (n l - q) / (q + 1),
// for the node at barycentric coordinates (a, b, c) / n (a + b + c = n), with
// l0 = 1 - r - s, l1 = r, l2 = s.  Shared by the numbering (DOF locations)
// and the reference tables.
#pragma once

#include <cmath>

#include "fe_mesh.h"

// Degree of a scalar triangle kind (0 for the others).
inline int feTriDegree(int kind)
{
  return kind == FE_TRI_P1 ? 1 : kind == FE_TRI_P2 ? 2 : kind == FE_TRI_P3 ? 3 : kind == FE_TRI_P4 ? 4 : 0;
}

// Reference coordinates of local node j: vertices, then the nodes of edge k
// (vertex k to vertex k+1) at steps 1/n, then the element nodes (P3: the
// centroid; P4: (1/4, 1/4), (1/2, 1/4), (1/4, 1/2)) -- the node orders of
// feSpace_2D.cpp:772-786, :1315-1327, :1700-1715.
inline void feTriNodeRS(int kind, int j, double &r, double &s)
{
  static const double V[3][2] = {{0., 0.}, {1., 0.}, {0., 1.}};
  const int n = feTriDegree(kind);
  if(j < 3) {
    r = V[j][0];
    s = V[j][1];
    return;
  }
  const int perEdge = n - 1;
  if(j < 3 + 3 * perEdge) {
    const int k = (j - 3) / perEdge, i = (j - 3) % perEdge;
    const double t = (double)(i + 1) / n;
    r = (1. - t) * V[k][0] + t * V[(k + 1) % 3][0];
    s = (1. - t) * V[k][1] + t * V[(k + 1) % 3][1];
    return;
  }
  static const double I4[3][2] = {{0.25, 0.25}, {0.5, 0.25}, {0.25, 0.5}};
  if(n == 3) {
    r = s = 1. / 3.;
  } else {
    r = I4[j - 12][0];
    s = I4[j - 12][1];
  }
}

// P_m(l) and its derivative.
inline void feLagrangeFactor(int n, int m, double l, double &p, double &dp)
{
  p = 1.;
  dp = 0.;
  for(int q = 0; q < m; ++q) {
    const double f = (n * l - q) / (q + 1), df = (double)n / (q + 1);
    dp = dp * f + p * df;
    p *= f;
  }
}

// Shape function of local node j of degree-n kind `kind` at (r, s), with its
// r and s derivatives.
inline void feTriShape(int kind, int j, double r, double s, double &L, double &dLr, double &dLs)
{
  const int n = feTriDegree(kind);
  double nr, ns;
  feTriNodeRS(kind, j, nr, ns);
  const int b = (int)std::lround(n * nr), c = (int)std::lround(n * ns), a = n - b - c;
  double p0, d0, p1, d1, p2, d2;
  feLagrangeFactor(n, a, 1. - r - s, p0, d0);
  feLagrangeFactor(n, b, r, p1, d1);
  feLagrangeFactor(n, c, s, p2, d2);
  L = p0 * p1 * p2;
  dLr = -d0 * p1 * p2 + p0 * d1 * p2;
  dLs = -d0 * p1 * p2 + p0 * p1 * d2;
}
