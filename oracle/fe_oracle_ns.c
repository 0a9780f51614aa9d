/* Placeholder until the NS restatement lands. */
#include "fe_oracle.h"
int or_assemble_ns(int V, const double *xyz, int E, const int *conn, const int *adrU,
                   const int *adrP, int n, const int *ia, const int *ja, int quad_degree,
                   double mu, const double *u, const int *elem_order, double *val,
                   double *rhs)
{
  (void)V; (void)xyz; (void)E; (void)conn; (void)adrU; (void)adrP; (void)n; (void)ia;
  (void)ja; (void)quad_degree; (void)mu; (void)u; (void)elem_order; (void)val; (void)rhs;
  return -100;
}
