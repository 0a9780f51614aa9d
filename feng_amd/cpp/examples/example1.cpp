// Usage of exe/example1.cpp on the MI355X back end: the steady heat equation
//   -div(k grad u) + f = 0,  f = -1,  u = 0 on the boundary "Bord",
// on a Gmsh mesh, with Lagrange P1 or P2, the diffusion and source forms,
// createLinearSystem(HIP) and solveNewtonRaphson.  The error is measured at
// the unknown DOFs against the reference example's series solution uSol_f
// (exe/example1.cpp:11-27), truncated at the same 40 terms.
//
//   example1 <mesh.msh> [order=1] [k=1] [degreeQuadrature=2]
// prints one JSON line.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "feng_api.h"

using namespace feng;

static double uSol(double x, double y, double k)
{
  const int nMax = 40;
  const double pi4 = M_PI * M_PI * M_PI * M_PI;
  double res = 0.;
  for(int i = 1; i < nMax; i += 2)
    for(int j = 1; j < nMax; j += 2) res += 16. / (pi4 * i * j * (i * i + j * j)) * sin(i * M_PI * x) * sin(j * M_PI * y);
  return res / k;
}

int main(int argc, char **argv)
{
  if(argc < 2) {
    std::fprintf(stderr, "usage: example1 mesh.msh [order] [k] [degreeQuadrature]\n");
    return 1;
  }
  const std::string meshFile = argv[1];
  const int order = argc > 2 ? std::atoi(argv[2]) : 1;
  const double k = argc > 3 ? std::atof(argv[3]) : 1.;
  // default rule: degree 2 integrates the P1/P2 stiffness exactly on straight
  // triangles; cubic and quartic fields need degree 2 (order - 1) at least
  // (a lower rule leaves the stiffness matrix singular)
  const int degreeQuadrature = argc > 4 ? std::atoi(argv[4]) : (order >= 3 ? 2 * order : 2);
  if(degreeQuadrature < 2 * (order - 1)) {
    std::fprintf(stderr, "example1: quadrature degree %d is below 2 (order - 1) = %d\n", degreeQuadrature,
                 2 * (order - 1));
    return 1;
  }

  feMesh *mesh = nullptr;
  if(feMesh::read(meshFile, mesh)) {
    std::fprintf(stderr, "cannot read %s\n", meshFile.c_str());
    return 1;
  }
  if(mesh->dim() != 2) {
    std::fprintf(stderr, "example1 runs the 2D case\n");
    return 1;
  }
  // one field U, essential on the boundary (spaces {uBord, uDomaine})
  if(order < 1 || order > 4) {
    std::fprintf(stderr, "order must be 1 to 4\n");
    return 1;
  }
  const int kinds[4] = {FE_TRI_P1, FE_TRI_P2, FE_TRI_P3, FE_TRI_P4};
  feMetaNumber numbering(*mesh, {kinds[order - 1]}, {1});
  if(numbering.status()) return 1;

  // solution vector: unknowns and essential DOFs (scalarConstant::zero)
  std::vector<double> sol(numbering.getNbDOFs(), 0.);

  feFunction diffusivity;
  diffusivity.kind = FE_FN_CONST;
  diffusivity.c0 = k;
  feFunction minusOne;
  minusOne.kind = FE_FN_CONST;
  minusOne.c0 = -1.;
  feBilinearForm *diff = nullptr, *source = nullptr;
  if(createBilinearForm(diff, {0}, FE_SYS_DIFFUSION, degreeQuadrature, {diffusivity}) ||
     createBilinearForm(source, {0}, FE_SYS_SOURCE, degreeQuadrature, {minusOne}))
    return 1;
  std::vector<feBilinearForm *> forms = {diff, source};

  feLinearSystem *system = nullptr;
  if(createLinearSystem(system, HIP, *mesh, forms, &numbering)) return 1;
  system->setRelativeTol(1e-12);

  feNLSolverOptions tol;
  int newton = 0;
  if(solveNewtonRaphson(system, sol, tol, &newton)) {
    std::fprintf(stderr, "Newton failed: %s\n", static_cast<feLinearSystemHIP *>(system)->lastError().c_str());
    return 1;
  }
  double emax = 0.;
  const std::vector<double> &xyz = numbering.dofCoordinates();
  for(int i = 0; i < numbering.getNbUnknowns(); ++i)
    emax = std::fmax(emax, std::fabs(sol[i] - uSol(xyz[3 * i], xyz[3 * i + 1], k)));
  double usum = 0.;
  for(int i = 0; i < numbering.getNbUnknowns(); ++i) usum += sol[i];
  std::printf("{\"mesh\": \"%s\", \"order\": %d, \"n\": %d, \"nnz\": %lld, \"newton\": %d, \"max_nodal_error\": %.6e, "
              "\"sum_u\": %.15e}\n",
              meshFile.c_str(), order, numbering.getNbUnknowns(),
              (long long)static_cast<feLinearSystemHIP *>(system)->nnz(), newton, emax, usum);
  delete system;
  delete diff;
  delete source;
  delete mesh;
  return 0;
}
