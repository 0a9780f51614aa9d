// The reference-side binding of the MI355X back end: the feLinearSystem a
// maintainer adds to the reference (src/feLinearSystemHIP.cpp there), with
// `HIP` added to linearSolverType (feLinearSystem.h:17-22) and
//   case HIP: system = new feLinearSystemHIP(bilinearForms, numbering);
// added to createLinearSystem (feLinearSystem.cpp:7-67).  Every virtual of
// feLinearSystem (feLinearSystem.h:85-168) maps onto the C ABI of
// libfe_hip.so (include/fe_hip.h).
//
// What the reference keeps out of reach of a linear system -- the
// coefficient callbacks of each feSysElm and the quadrature degree of its
// spaces (protected members) -- is registered per form with setFormDevice()
// before the first assembly; a form without registration takes the
// reference's unit coefficient and a degree-2 rule.  The device set-up
// (mesh, coloring, fields, forms, CSR pattern) runs on the first assembly.
//
// This file compiles against the reference's headers
// (tests/test_reference_binding.py runs g++ -fsyntax-only on it with the
// oracle's reference flags); it is not part of libfe_hip.so.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "feLinearSystem.h"
#include "fe_hip.h"

class feLinearSystemHIP : public feLinearSystem
{
  struct FormDevice {
    int quadDegree = 2;
    std::vector<fe_hip_fn> coeffs;
  };
  fe_hip_ctx *_ctx = nullptr;
  int _n = 0, _nDofs = 0;
  bool _ready = false;
  double _normDx = 0.;  // ||du||_inf of the last solve (getResidualMaxNorm)
  std::vector<double> _u;
  std::map<int, FormDevice> _formDevice;
  std::vector<feBilinearForm *> _forms;
  fe_hip_solver_opts _opts;

public:
  feLinearSystemHIP(std::vector<feBilinearForm *> bilinearForms, const feMetaNumber *numbering)
    : feLinearSystem(bilinearForms, numbering), _forms(bilinearForms)
  {
    _n = numbering->getNbUnknowns();
    _nDofs = numbering->getNbDOFs();
    _u.assign(_nDofs, 0.);
    fe_hip_default_opts(&_opts);
    if(fe_hip_ctx_create(0, &_ctx) != 0) {
      feErrorMsg(FE_STATUS_ERROR, "feLinearSystemHIP: no HIP device");
      _ctx = nullptr;
    }
  }
  ~feLinearSystemHIP() { fe_hip_ctx_destroy(_ctx); }

  // Coefficients (in the order the feSysElm constructor takes them) and the
  // quadrature degree of form `iForm` (index in the vector given to
  // createLinearSystem).
  void setFormDevice(int iForm, int quadDegree, const std::vector<fe_hip_fn> &coeffs)
  {
    _formDevice[iForm] = FormDevice{quadDegree, coeffs};
  }
  // The KSP / PC choice PETSc reads from -ksp_type / -pc_type.
  void setSolver(int ksp, int pc, int precision)
  {
    _opts.ksp = ksp;
    _opts.pc = pc;
    _opts.precision = precision;
  }

  feInt getSystemSize() const override { return _n; }

  void getRHSMaxNorm(double *norm) const override { fe_hip_rhs_max_norm(_ctx, norm); }
  // the max norm of the last correction du, as feLinearSystemPETSc.cpp:512-519
  // and feLinearSystemMklPardiso.cpp:352-355 return it
  void getResidualMaxNorm(double *norm) const override { *norm = _normDx; }

  void setToZero() override { fe_hip_set_to_zero(_ctx, 1, 1); }
  void setMatrixToZero() override { fe_hip_set_to_zero(_ctx, 1, 0); }
  void setResidualToZero() override { fe_hip_set_to_zero(_ctx, 0, 1); }

  void assemble(const feSolution *sol, const bool) override { run(sol, 3); }
  void assembleMatrices(const feSolution *sol, const bool) override { run(sol, 1); }
  void assembleResiduals(const feSolution *sol) override { run(sol, 2); }

  // DOF lists of essential components / periodic pairs are handed over once
  // with fe_hip_set_constrained_dofs / fe_hip_set_periodic_dofs
  // (feLinearSystemMklPardiso.cpp:1000-1037, PeriodicDOF())
  void constrainEssentialComponents(const feSolution *) override { fe_hip_constrain_essential_components(_ctx); }
  void applyPeriodicity() override { fe_hip_apply_periodicity(_ctx); }
  void permute() override {}

  bool solve(double *normDx, double *normResidual, double *normAxb, int *nIter) override
  {
    fe_hip_solver_opts o = _opts;
    o.rtol = _rel_tol;
    o.atol = _abs_tol;
    o.dtol = _div_tol;
    o.max_iter = _max_iter;
    fe_hip_solve_stats st;
    if(fe_hip_solve(_ctx, &o, nullptr, &st) != 0) return false;
    _normDx = st.norm_dx_inf;
    *normDx = st.norm_dx_inf;
    *normResidual = st.norm_rhs_inf;
    *normAxb = st.norm_axb_inf;
    *nIter = st.iterations;
    return st.converged || st.reason == -3;  // DIVERGED_ITS counts as success (feLinearSystemPETSc.cpp:1017-1020)
  }

  void correctSolution(feSolution *sol, const bool) override
  {
    fe_hip_correct_solution(_ctx);
    fe_hip_get_solution(_ctx, _u.data());
    for(int i = 0; i < _n; ++i) sol->setSolAtDOF(i, _u[i]);
  }

  // Deferred-correction helpers: the DC schemes are not on the device path.
  void assignResidualToDCResidual(feSolutionContainer *) override {}
  void applyCorrectionToResidual(double, std::vector<double> &) override {}

  void viewMatrix() const override {}
  void viewRHS() const override {}
  void viewResidual() const override {}
  void writeMatrix(const std::string, const double) override {}
  void writeRHS(const std::string, const double) override {}
  void writeResidual(const std::string, const double) override {}

private:
  void run(const feSolution *sol, int what)
  {
    if(!_ready && setUp() != FE_STATUS_OK) return;
    const std::vector<double> &u = sol->getSolution();
    fe_hip_set_solution(_ctx, u.data());
    fe_hip_assemble(_ctx, what, _opts.precision);
  }

  // Device field kind of a reference space (the kinds the device assembles).
  static int deviceKind(const feSpace *s)
  {
    const int dim = s->getDim(), deg = s->getPolynomialDegree(), nc = s->getNumComponents();
    if(dim == 2 && nc == 1 && deg >= 1 && deg <= 4) return deg == 1 ? 1 : deg == 2 ? 2 : deg == 3 ? 5 : 6;
    if(dim == 2 && nc == 2 && deg == 2) return 3;  // FE_TRI_VP2
    if(dim == 3 && nc == 1 && deg == 1) return 4;  // FE_TET_P1
    return -1;
  }

  feStatus setUp()
  {
    if(!_ctx) return FE_STATUS_ERROR;
    if(_forms.empty()) return feErrorMsg(FE_STATUS_ERROR, "feLinearSystemHIP: no forms");
    // geometry of the forms' connectivity (feMesh::getCoord, feMesh.cpp:119-134)
    const feCncGeo *cnc = _forms[0]->getCncGeo();
    feMesh *mesh = const_cast<feMesh *>(cnc->getMesh());
    const std::vector<Vertex> &verts = mesh->getVertices();
    std::vector<double> xyz(3 * verts.size());
    for(size_t v = 0; v < verts.size(); ++v) {
      xyz[3 * v + 0] = verts[v].x();
      xyz[3 * v + 1] = verts[v].y();
      xyz[3 * v + 2] = verts[v].z();
    }
    const int E = cnc->getNumElements(), nvpe = cnc->getNumVerticesPerElem();
    const std::vector<int> &conn = cnc->getVerticesConnectivity();
    if(fe_hip_upload_mesh(_ctx, cnc->getDim(), (int)verts.size(), xyz.data(), E, nvpe, conn.data()) != 0)
      return feErrorMsg(FE_STATUS_ERROR, "fe_hip_upload_mesh: %s", fe_hip_last_error(_ctx));
    // the reference's coloring (feCncGeo::colorElements) orders the device sweeps
    if(cnc->getNbColor() > 0) {
      const std::vector<int> &color = cnc->getColorElm();
      if((int)color.size() == E) fe_hip_set_coloring(_ctx, color.data());
    }
    // one device field per distinct feSpace: its element -> DOF map
    // (feSpace::initializeAddressingVector, unchanged)
    std::map<const feSpace *, int> fieldOf;
    for(feBilinearForm *f : _forms)
      for(feSpace *s : f->_intSpaces) {
        if(fieldOf.count(s)) continue;
        const int kind = deviceKind(s);
        if(kind < 0) return feErrorMsg(FE_STATUS_ERROR, "feLinearSystemHIP: space %s not on the device",
                                       s->getFieldID().c_str());
        const int nloc = s->getNumFunctions();
        std::vector<int> adr((size_t)E * nloc);
        std::vector<feInt> a(nloc);
        for(int e = 0; e < E; ++e) {
          s->initializeAddressingVector(e, a);
          for(int k = 0; k < nloc; ++k) adr[(size_t)e * nloc + k] = (int)a[k];
        }
        int id = -1;
        if(fe_hip_add_field(_ctx, kind, adr.data(), &id) != 0)
          return feErrorMsg(FE_STATUS_ERROR, "fe_hip_add_field: %s", fe_hip_last_error(_ctx));
        fieldOf[s] = id;
      }
    fe_hip_set_sizes(_ctx, _n, _nDofs);
    // one device form per feBilinearForm: getID() is the elementSystemType
    for(size_t i = 0; i < _forms.size(); ++i) {
      std::vector<int> fields;
      for(feSpace *s : _forms[i]->_intSpaces) fields.push_back(fieldOf[s]);
      FormDevice fd;
      auto it = _formDevice.find((int)i);
      if(it != _formDevice.end()) fd = it->second;
      if(fd.coeffs.empty()) fd.coeffs.push_back(fe_hip_fn{FE_FN_CONST, 1., 0., nullptr, 0});
      int id = -1;
      if(fe_hip_add_form(_ctx, (int)_forms[i]->getID(), (int)fields.size(), fields.data(), fd.quadDegree,
                         (int)fd.coeffs.size(), fd.coeffs.data(), &id) != 0)
        return feErrorMsg(FE_STATUS_ERROR, "fe_hip_add_form: %s", fe_hip_last_error(_ctx));
    }
    int64_t nnz = 0;
    if(fe_hip_build_csr(_ctx, &nnz) != 0)  // the feEZCompressedRowStorage pattern, bit-exact
      return feErrorMsg(FE_STATUS_ERROR, "fe_hip_build_csr: %s", fe_hip_last_error(_ctx));
    _ready = true;
    return FE_STATUS_OK;
  }
};

// The factory case a maintainer adds to createLinearSystem (feLinearSystem.cpp:7-67).
feStatus createLinearSystemHIP(feLinearSystem *&system, const std::vector<feBilinearForm *> bilinearForms,
                               const feMetaNumber *numbering)
{
  for(feBilinearForm *f : bilinearForms)
    if(f == nullptr) return feErrorMsg(FE_STATUS_ERROR, "Null bilinear form");
  if(numbering->getNbUnknowns() == 0) return feErrorMsg(FE_STATUS_ERROR, "0 unknowns");
  system = new feLinearSystemHIP(bilinearForms, numbering);
  return FE_STATUS_OK;
}
