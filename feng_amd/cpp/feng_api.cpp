// C++ mirror of the reference's linear-system interface over the C ABIs.
#include "feng_api.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace feng {

// ---- mesh -----------------------------------------------------------------

feMesh::feMesh(fe_mesh *h) : _h(h)
{
  int nB = 0, bnv = 0, nEdges = 0;
  fe_mesh_info(h, &_dim, &_V, &_E, &_nv, &nB, &bnv, &nEdges);
  _xyz.resize(3 * (size_t)_V);
  _conn.resize((size_t)_E * _nv);
  std::vector<int> bconn((size_t)std::max(1, nB * bnv)), edges((size_t)std::max(1, 3 * _E)), bedge(std::max(1, nB));
  fe_mesh_get(h, _xyz.data(), _conn.data(), bconn.data(), edges.data(), bedge.data());
}

feMesh::~feMesh()
{
  if(_h) fe_mesh_destroy(_h);
}

feStatus feMesh::square(int N, double jitter, uint64_t seed, feMesh *&out)
{
  fe_mesh *h = nullptr;
  if(int rc = fe_mesh_square(N, jitter, seed, &h)) return rc;
  out = new feMesh(h);
  return FE_STATUS_OK;
}

feStatus feMesh::cube(int N, feMesh *&out)
{
  fe_mesh *h = nullptr;
  if(int rc = fe_mesh_cube(N, &h)) return rc;
  out = new feMesh(h);
  return FE_STATUS_OK;
}

feStatus feMesh::read(const std::string &path, feMesh *&out)
{
  fe_mesh *h = nullptr;
  if(int rc = fe_mesh_read_msh(path.c_str(), &h)) return rc;
  out = new feMesh(h);
  return FE_STATUS_OK;
}

// ---- numbering ------------------------------------------------------------

feMetaNumber::feMetaNumber(const feMesh &mesh, const std::vector<int> &kinds, const std::vector<int> &essential)
  : _kinds(kinds)
{
  if(kinds.size() != essential.size()) {
    _status = FE_STATUS_ERROR;
    return;
  }
  _status = fe_numbering_create(mesh.handle(), (int)kinds.size(), kinds.data(), essential.data(), &_h);
  if(_status) return;
  fe_numbering_info(_h, &_nInc, &_nDofs);
  for(size_t f = 0; f < kinds.size(); ++f) {
    const int k = kinds[f];
    const int nloc = k == FE_TRI_P1 ? 3 : k == FE_TRI_P2 ? 6 : k == FE_TRI_VP2 ? 12 : k == FE_TET_P1 ? 4
                     : k == FE_TRI_P3 ? 10 : k == FE_TRI_P4 ? 15 : 0;
    if(nloc == 0) {
      _status = FE_STATUS_ERROR;
      return;
    }
    _adr.emplace_back((size_t)mesh.nElements() * nloc);
    if((_status = fe_numbering_adr(_h, (int)f, _adr.back().data()))) return;
  }
  _dofXyz.resize(3 * (size_t)_nDofs);
  _dofField.resize(_nDofs);
  _dofComp.resize(_nDofs);
  _status = fe_numbering_dof_info(_h, _dofXyz.data(), _dofField.data(), _dofComp.data());
}

feMetaNumber::~feMetaNumber()
{
  if(_h) fe_numbering_destroy(_h);
}

feStatus createBilinearForm(feBilinearForm *&form, const std::vector<int> &fields, int elementSystemType,
                            int degreeQuadrature, const std::vector<feFunction> &coefficients)
{
  if(fields.empty()) return FE_STATUS_ERROR;
  form = new feBilinearForm{elementSystemType, fields, degreeQuadrature, coefficients};
  return FE_STATUS_OK;
}

// ---- linear system --------------------------------------------------------

feLinearSystemHIP::feLinearSystemHIP(const feMesh &mesh, const std::vector<feBilinearForm *> &forms,
                                     const feMetaNumber *numbering, int device)
{
  _n = numbering->getNbUnknowns();
  _nDofs = numbering->getNbDOFs();
  if((_status = fe_hip_ctx_create(device, &_ctx))) return;
  if((_status = fe_hip_upload_mesh(_ctx, mesh.dim(), mesh.nVertices(), mesh.coordinates().data(), mesh.nElements(),
                                   mesh.nVerticesPerElem(), mesh.connectivity().data())))
    return;
  for(int f = 0; f < numbering->nFields(); ++f) {
    int id = -1;
    if((_status = fe_hip_add_field(_ctx, numbering->kind(f), numbering->addressing(f).data(), &id))) return;
  }
  if((_status = fe_hip_set_sizes(_ctx, _n, _nDofs))) return;
  for(const feBilinearForm *F : forms) {
    std::vector<fe_hip_fn> c(F->coefficients.size());
    for(size_t k = 0; k < c.size(); ++k) {
      const feFunction &fn = F->coefficients[k];
      c[k].kind = fn.kind;
      c[k].c0 = fn.c0;
      c[k].c1 = fn.c1;
      c[k].data = fn.table.empty() ? nullptr : fn.table.data();
      c[k].count = (int64_t)fn.table.size();
    }
    int id = -1;
    if((_status = fe_hip_add_form(_ctx, F->elementSystemType, (int)F->fields.size(), F->fields.data(),
                                  F->degreeQuadrature, (int)c.size(), c.empty() ? nullptr : c.data(), &id)))
      return;
  }
  _status = fe_hip_build_csr(_ctx, &_nnz);
}

feLinearSystemHIP::~feLinearSystemHIP()
{
  if(_ctx) fe_hip_ctx_destroy(_ctx);
}

std::string feLinearSystemHIP::lastError() const { return _ctx ? fe_hip_last_error(_ctx) : "no context"; }

void feLinearSystemHIP::upload(const std::vector<double> &sol) { _status = fe_hip_set_solution(_ctx, sol.data()); }
void feLinearSystemHIP::setToZero() { _status = fe_hip_set_to_zero(_ctx, 1, 1); }
void feLinearSystemHIP::setMatrixToZero() { _status = fe_hip_set_to_zero(_ctx, 1, 0); }
void feLinearSystemHIP::setResidualToZero() { _status = fe_hip_set_to_zero(_ctx, 0, 1); }

void feLinearSystemHIP::assemble(const std::vector<double> &sol)
{
  upload(sol);
  if(!_status) _status = fe_hip_assemble(_ctx, 3, _prec);
}

void feLinearSystemHIP::assembleMatrices(const std::vector<double> &sol)
{
  upload(sol);
  if(!_status) _status = fe_hip_assemble(_ctx, 1, _prec);
}

void feLinearSystemHIP::assembleResiduals(const std::vector<double> &sol)
{
  upload(sol);
  if(!_status) _status = fe_hip_assemble(_ctx, 2, FE_PREC_FP64);
}

void feLinearSystemHIP::getRHSMaxNorm(double *norm) const { fe_hip_rhs_max_norm(_ctx, norm); }

bool feLinearSystemHIP::solve(double *normDx, double *normResidual, double *normAxb, int *nIter)
{
  fe_hip_solver_opts o;
  fe_hip_default_opts(&o);
  o.ksp = _ksp;
  o.pc = _pc;
  o.precision = _prec;
  o.rtol = _rel_tol;
  o.atol = _abs_tol;
  o.dtol = _div_tol;
  o.max_iter = _max_iter;
  fe_hip_solve_stats st;
  if((_status = fe_hip_solve(_ctx, &o, nullptr, &st))) return false;
  *normDx = st.norm_dx_inf;
  *normResidual = st.norm_rhs_inf;
  *normAxb = st.norm_axb_inf;
  *nIter = st.iterations;
  // DIVERGED_ITS counts as success (feLinearSystemPETSc.cpp:1017-1020)
  return st.converged || st.reason == -3;
}

void feLinearSystemHIP::correctSolution(std::vector<double> &sol)
{
  if((_status = fe_hip_correct_solution(_ctx))) return;
  _status = fe_hip_get_solution(_ctx, sol.data());
}

// The DOF lists come through fe_hip_set_constrained_dofs /
// fe_hip_set_periodic_dofs on context(); without them both are no-ops.
void feLinearSystemHIP::constrainEssentialComponents() { _status = fe_hip_constrain_essential_components(_ctx); }

void feLinearSystemHIP::applyPeriodicity() { _status = fe_hip_apply_periodicity(_ctx); }

feStatus createLinearSystem(feLinearSystem *&system, linearSolverType type, const feMesh &mesh,
                            const std::vector<feBilinearForm *> &forms, const feMetaNumber *numbering)
{
  system = nullptr;
  for(const feBilinearForm *f : forms)
    if(!f) return FE_STATUS_ERROR;
  if(!numbering || numbering->getNbUnknowns() == 0) return FE_STATUS_ERROR;
  if(type != HIP) return FE_STATUS_ERROR;  // PETSc / MKL Pardiso are not part of this build
  feLinearSystemHIP *s = new feLinearSystemHIP(mesh, forms, numbering);
  if(s->status()) {
    std::fprintf(stderr, "createLinearSystem: %s\n", s->lastError().c_str());
    const feStatus rc = s->status();
    delete s;
    return rc;
  }
  system = s;
  return FE_STATUS_OK;
}

feStatus solveNewtonRaphson(feLinearSystem *system, std::vector<double> &sol, const feNLSolverOptions &tol,
                            int *iterations)
{
  int it = 0;
  for(;;) {
    system->setToZero();
    system->assembleResiduals(sol);
    double nres = 0.;
    system->getRHSMaxNorm(&nres);
    if(it > 0 && nres <= tol.tolResidual) break;
    system->assembleMatrices(sol);
    system->constrainEssentialComponents();
    system->applyPeriodicity();
    double ndx = 0., nrhs = 0., naxb = 0.;
    int lin = 0;
    if(!system->solve(&ndx, &nrhs, &naxb, &lin)) return FE_STATUS_FAILED;
    if(nrhs > tol.tolDivergence) return FE_STATUS_FAILED;
    system->correctSolution(sol);
    ++it;
    if(nrhs <= tol.tolResidual || ndx <= tol.tolCorrection) break;
    if(it > tol.maxIter) return FE_STATUS_FAILED;
  }
  if(iterations) *iterations = it;
  return FE_STATUS_OK;
}

}  // namespace feng
