// Golden-vector dumper linked against the reference feNG sources.
//
// Test infrastructure only: built by oracle/Makefile into oracle/_ref/ (never
// shipped, never loaded by the product).  It drives the reference's own
// mesh reader, FE spaces, DOF numbering, CSR pattern builder and element
// kernels on a mesh file and writes what they produce to a small binary
// container that tests/golden/make_fixtures.py turns into fixtures.
//
// The scatter loop below follows the reference's colored element loop with
// the Pardiso sorted-CSR walk (feLinearSystemMklPardiso.cpp:501-749): every
// element of every matrix form adds Ae[i][j] at (adrI[i], adrJ[j]) for rows and
// columns below nInc, and Be[i] at adrI[i] for rows below nInc.
//
// Usage:
//   ref_dump <mesh.msh> <problem> <order> <degQuad> <out.bin> [param]
// problems: diff2d_sinsin | diff2d_blocks | diff3d_smooth | ns_kovasznay | ns_mms
#include "feAPI.h"

#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

// Defined by the reference's Newton driver (feNonLinearSolver.cpp); it holds
// the solution at the start of the time step and is read by every element
// gather, so it is set here exactly as solveNewtonRaphson does before assembly.
extern std::vector<double> solAtTimeN;

namespace {

FILE *g_out = nullptr;

void putArray(const char *name, char type, const void *data, int64_t count)
{
  int32_t len = (int32_t)strlen(name);
  fwrite(&len, 4, 1, g_out);
  fwrite(name, 1, len, g_out);
  fwrite(&type, 1, 1, g_out);
  fwrite(&count, 8, 1, g_out);
  size_t esz = (type == 'd') ? 8 : 4;
  if(count > 0) fwrite(data, esz, (size_t)count, g_out);
}
void putInts(const char *name, const std::vector<int> &v) { putArray(name, 'i', v.data(), (int64_t)v.size()); }
void putDoubles(const char *name, const std::vector<double> &v) { putArray(name, 'd', v.data(), (int64_t)v.size()); }
void putInt(const char *name, int v) { putArray(name, 'i', &v, 1); }

// Block-constant diffusivity used by the diff2d_blocks fixture: B x B blocks on
// the unit square, value exp(0.5*(((7*bi + 13*bj) % 11) - 5)).
int g_blocks = 4;
double blockK(const feFunctionArguments &args, const std::vector<double> &)
{
  int bi = (int)floor(args.pos[0] * g_blocks);
  int bj = (int)floor(args.pos[1] * g_blocks);
  if(bi >= g_blocks) bi = g_blocks - 1;
  if(bj >= g_blocks) bj = g_blocks - 1;
  return exp(0.5 * (double)(((7 * bi + 13 * bj) % 11) - 5));
}
double sinsinSrc(const feFunctionArguments &args, const std::vector<double> &)
{
  return -2. * M_PI * M_PI * sin(M_PI * args.pos[0]) * sin(M_PI * args.pos[1]);
}
double smoothK3(const feFunctionArguments &args, const std::vector<double> &)
{
  const double x = args.pos[0], y = args.pos[1], z = args.pos[2];
  return exp(1.5 * sin(2. * M_PI * x) * sin(2. * M_PI * y) * sin(2. * M_PI * z));
}

// Kovasznay flow: lambda = Re/2 - sqrt(Re^2/4 + 4 pi^2)
double g_Re = 40.;
double kovLambda() { return g_Re / 2. - sqrt(g_Re * g_Re / 4. + 4. * M_PI * M_PI); }
void kovU(const feFunctionArguments &args, const std::vector<double> &, std::vector<double> &res)
{
  const double l = kovLambda(), x = args.pos[0], y = args.pos[1];
  res[0] = 1. - exp(l * x) * cos(2. * M_PI * y);
  res[1] = l / (2. * M_PI) * exp(l * x) * sin(2. * M_PI * y);
}
double kovP(const feFunctionArguments &args, const std::vector<double> &)
{
  const double l = kovLambda();
  return 0.5 * (1. - exp(2. * l * args.pos[0]));
}
double g_mu = 1. / 40.;
double muFun(const feFunctionArguments &, const std::vector<double> &) { return g_mu; }

// Manufactured NS solution of the reference's navier_stokes.cpp test is not
// reproduced here; the ns_mms problem uses the Kovasznay field with the full
// source-free form stack at Re = param.

// The reference keeps the CSR arrays protected (their public getters exist
// only in MKL builds); a derived class reads them back.
class CsrProbe : public feEZCompressedRowStorage
{
public:
  using feEZCompressedRowStorage::feEZCompressedRowStorage;
  const std::vector<feInt> &ia() const { return ia_Pardiso; }
  const std::vector<feInt> &ja() const { return ja_Pardiso; }
};

} // namespace

int main(int argc, char **argv)
{
  if(argc < 6) {
    fprintf(stderr, "usage: ref_dump mesh problem order degQuad out.bin [param]\n");
    return 2;
  }
  std::string meshFile = argv[1];
  std::string problem = argv[2];
  int order = atoi(argv[3]);
  int degQuad = atoi(argv[4]);
  std::string outFile = argv[5];
  double param = argc > 6 ? atof(argv[6]) : 0.;

  setVerbose(0);
  feMesh2DP1 mesh(meshFile);

  std::vector<feSpace *> spaces, essentialSpaces;
  std::vector<feBilinearForm *> forms;
  std::vector<feSpace *> domainSpaces;

  feFunction fBlockK(blockK), fSinSin(sinsinSrc), fSmoothK(smoothK3), fKovP(kovP), fMu(muFun);
  feVectorFunction fKovU(kovU);

  const bool isDiff = (problem == "diff2d_sinsin" || problem == "diff2d_blocks" || problem == "diff3d_smooth");
  const bool isNS = (problem == "ns_kovasznay");
  if(!isDiff && !isNS) {
    fprintf(stderr, "unknown problem %s\n", problem.c_str());
    return 2;
  }
  feSpace *u = nullptr, *uB = nullptr, *p = nullptr, *pB = nullptr;
  if(isDiff) {
    if(problem == "diff2d_blocks" && param > 0) g_blocks = (int)param;
    feCheck(createFiniteElementSpace(uB, &mesh, elementType::LAGRANGE, order, "U", "Bord", degQuad, &scalarConstant::zero));
    feCheck(createFiniteElementSpace(u, &mesh, elementType::LAGRANGE, order, "U", "Domaine", degQuad, &scalarConstant::zero));
    spaces = {uB, u};
    essentialSpaces = {uB};
    domainSpaces = {u};
  } else {
    if(param > 0) g_Re = param;
    g_mu = 1. / g_Re;
    feCheck(createFiniteElementSpace(u, &mesh, elementType::VECTOR_LAGRANGE, order, "U", "Domaine", degQuad, &vectorConstant::zero));
    feCheck(createFiniteElementSpace(uB, &mesh, elementType::VECTOR_LAGRANGE, order, "U", "Bord", degQuad, &fKovU));
    feCheck(createFiniteElementSpace(p, &mesh, elementType::LAGRANGE, order - 1, "P", "Domaine", degQuad, &scalarConstant::zero));
    feCheck(createFiniteElementSpace(pB, &mesh, elementType::LAGRANGE, order - 1, "P", "Bord", degQuad, &fKovP));
    spaces = {u, uB, p, pB};
    essentialSpaces = {uB, pB};
    domainSpaces = {u, p};
  }

  feMetaNumber numbering(&mesh, spaces, essentialSpaces);
  feSolution sol(numbering.getNbDOFs(), spaces, essentialSpaces);
  sol.initialize(&mesh);
  solAtTimeN = sol.getSolution();

  if(isDiff) {
    feBilinearForm *diff = nullptr, *src = nullptr;
    if(problem == "diff2d_sinsin") {
      feCheck(createBilinearForm(diff, {u}, new feSysElm_Diffusion<2>(&scalarConstant::one)));
      feCheck(createBilinearForm(src, {u}, new feSysElm_Source(&fSinSin)));
    } else if(problem == "diff2d_blocks") {
      feCheck(createBilinearForm(diff, {u}, new feSysElm_Diffusion<2>(&fBlockK)));
      feCheck(createBilinearForm(src, {u}, new feSysElm_Source(&scalarConstant::minusOne)));
    } else {
      feCheck(createBilinearForm(diff, {u}, new feSysElm_Diffusion<3>(&fSmoothK)));
      feCheck(createBilinearForm(src, {u}, new feSysElm_Source(&scalarConstant::minusOne)));
    }
    forms = {diff, src};
  } else {
    feBilinearForm *convU = nullptr, *divU = nullptr, *source = nullptr, *divSigma = nullptr;
    feCheck(createBilinearForm(convU, {u}, new feSysElm_VectorConvectiveAcceleration<2>(&scalarConstant::minusOne)));
    feCheck(createBilinearForm(divU, {p, u}, new feSysElm_MixedDivergence<2>(&scalarConstant::one)));
    feCheck(createBilinearForm(source, {u}, new feSysElm_VectorSource<2>(&vectorConstant::zero)));
    feCheck(createBilinearForm(divSigma, {u, p}, new feSysElm_DivergenceNewtonianStress<2>(&scalarConstant::one, &fMu)));
    forms = {convU, divU, source, divSigma};
  }
  const int nInc = numbering.getNbUnknowns();
  const int nDofs = numbering.getNbDOFs();

  g_out = fopen(outFile.c_str(), "wb");
  if(!g_out) return 3;

  putInt("n", nInc);
  putInt("ndofs", nDofs);

  // Mesh as read: vertices, element connectivity and signed edge tags of the domain
  {
    std::vector<Vertex> &verts = mesh.getVertices();
    std::vector<double> xyz;
    for(auto &v : verts) { xyz.push_back(v.x()); xyz.push_back(v.y()); xyz.push_back(v.z()); }
    putDoubles("xyz", xyz);
    const feCncGeo *cnc = mesh.getCncGeoByName("Domaine");
    putInts("conn", cnc->getVerticesConnectivity());
    putInts("edges", cnc->getEdgeConnectivity());
    putInt("nvpe", cnc->getNumVerticesPerElem());
    const feCncGeo *bnd = mesh.getCncGeoByName("Bord");
    putInts("bconn", bnd->getVerticesConnectivity());
    std::vector<int> colorOf(cnc->getNumElements(), -1);
    for(int c = 0; c < cnc->getNbColor(); ++c)
      for(int e : cnc->getListElmPerColorI(c)) colorOf[e] = c;
    putInts("color", colorOf);
  }

  // Element -> DOF maps, per domain space (initializeAddressingVector)
  for(size_t s = 0; s < domainSpaces.size(); ++s) {
    feSpace *fs = domainSpaces[s];
    int nf = fs->getNumFunctions();
    std::vector<feInt> adr(nf);
    std::vector<int> all;
    for(int e = 0; e < fs->getNumElements(); ++e) {
      fs->initializeAddressingVector(e, adr);
      for(int i = 0; i < nf; ++i) all.push_back(adr[i]);
    }
    std::string nm = "adr" + std::to_string(s);
    putInts(nm.c_str(), all);
  }

  // Initial solution (unknowns from the domain functions, essentials from the
  // boundary functions)
  putDoubles("u0", sol.getSolution());

  // CSR pattern from the matrix forms
  std::vector<feBilinearForm *> matForms;
  for(auto *f : forms)
    if(f->hasMatrix()) matForms.push_back(f);
  CsrProbe crs(nInc, matForms, (int)matForms.size(), &numbering);
  int nnz = (int)crs.getNumNNZ();
  std::vector<int> iav(crs.ia().begin(), crs.ia().end()), jav(crs.ja().begin(), crs.ja().end());
  putInts("ia", iav);
  putInts("ja", jav);

  // Assembly: colored element loop, sorted-row scatter
  std::vector<double> val(nnz, 0.), rhs(nInc, 0.);
  auto findPos = [&](int row, int col) -> int {
    for(int k = iav[row]; k < iav[row + 1]; ++k)
      if(jav[k] == col) return k;
    return -1;
  };
  for(auto *f : forms) {
    const feCncGeo *cnc = f->getCncGeo();
    for(int c = 0; c < cnc->getNbColor(); ++c) {
      for(int e : cnc->getListElmPerColorI(c)) {
        f->computeResidual(&sol, e);
        const std::vector<feInt> &adrI = f->getAdrI();
        const double *Be = f->getBe();
        for(size_t i = 0; i < adrI.size(); ++i)
          if(adrI[i] < nInc) rhs[adrI[i]] += Be[i];
        if(f->hasMatrix()) {
          f->computeMatrix(&sol, e);
          const std::vector<feInt> &aI = f->getAdrI();
          const std::vector<feInt> &aJ = f->getAdrJ();
          const double *const *Ae = f->getAe();
          for(size_t i = 0; i < aI.size(); ++i) {
            if(aI[i] >= nInc) continue;
            for(size_t j = 0; j < aJ.size(); ++j) {
              if(aJ[j] >= nInc) continue;
              int p = findPos(aI[i], aJ[j]);
              if(p < 0) { fprintf(stderr, "pattern miss\n"); return 4; }
              val[p] += Ae[i][j];
            }
          }
        }
      }
    }
  }
  putDoubles("val", val);
  putDoubles("rhs", rhs);

  // Per-form adrI of element 0 (layout check) and form ids
  for(size_t k = 0; k < forms.size(); ++k) {
    forms[k]->initializeAddressingVectors(0);
    std::vector<int> a(forms[k]->getAdrI().begin(), forms[k]->getAdrI().end());
    std::vector<int> b(forms[k]->getAdrJ().begin(), forms[k]->getAdrJ().end());
    std::string na = "form" + std::to_string(k) + "_adrI0";
    std::string nb = "form" + std::to_string(k) + "_adrJ0";
    putInts(na.c_str(), a);
    putInts(nb.c_str(), b);
  }

  fclose(g_out);
  for(auto *f : forms) delete f;
  for(auto *s : spaces) delete s;
  return 0;
}
