// C++ mirror of the reference's linear-system interface for the MI355X path.
//
// Names and argument meanings follow the reference so a driver written
// against feNG reads the same: createLinearSystem (feLinearSystem.cpp:7-67)
// with one more enumerator, HIP; the feLinearSystem virtuals
// (feLinearSystem.h:85-168) that solveNewtonRaphson uses; feNLSolverOptions
// and solveNewtonRaphson (feNonLinearSolver.h:12-40,64-68,
// feNonLinearSolver.cpp:38-182).  Meshes, numbering and forms are the host
// library (include/fe_mesh.h) and form descriptors of include/fe_hip.h.
// Status codes are feStatus values (feMessage.h:44-62).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fe_hip.h"
#include "fe_mesh.h"

namespace feng {

typedef int feStatus;

enum linearSolverType { MKLPARDISO, PETSC, PETSC_MUMPS, HIP };

// Mesh of the host library (generated or read from a .msh v4.1 file).
class feMesh {
public:
  ~feMesh();
  static feStatus square(int N, double jitter, uint64_t seed, feMesh *&out);
  static feStatus cube(int N, feMesh *&out);
  static feStatus read(const std::string &path, feMesh *&out);
  int dim() const { return _dim; }
  int nVertices() const { return _V; }
  int nElements() const { return _E; }
  int nVerticesPerElem() const { return _nv; }
  const std::vector<double> &coordinates() const { return _xyz; }
  const std::vector<int> &connectivity() const { return _conn; }
  fe_mesh *handle() const { return _h; }

private:
  explicit feMesh(fe_mesh *h);
  fe_mesh *_h = nullptr;
  int _dim = 0, _V = 0, _E = 0, _nv = 0;
  std::vector<double> _xyz;
  std::vector<int> _conn;
};

// feMetaNumber counterpart: one field per entry of `kinds` (FE_TRI_P1 ...),
// essential[f] != 0 makes field f Dirichlet on the boundary.
class feMetaNumber {
public:
  feMetaNumber(const feMesh &mesh, const std::vector<int> &kinds, const std::vector<int> &essential);
  ~feMetaNumber();
  feStatus status() const { return _status; }
  int getNbUnknowns() const { return _nInc; }
  int getNbDOFs() const { return _nDofs; }
  const std::vector<int> &addressing(int field) const { return _adr[field]; }
  int kind(int field) const { return _kinds[field]; }
  int nFields() const { return (int)_kinds.size(); }
  // DOF positions (3 per DOF), field and component of every DOF
  const std::vector<double> &dofCoordinates() const { return _dofXyz; }
  const std::vector<int> &dofField() const { return _dofField; }

private:
  fe_numbering *_h = nullptr;
  feStatus _status = 0;
  int _nInc = 0, _nDofs = 0;
  std::vector<int> _kinds;
  std::vector<std::vector<int>> _adr;
  std::vector<double> _dofXyz;
  std::vector<int> _dofField, _dofComp;
};

// Coefficient of a form: one of the device function kinds (FE_FN_*), with a
// table for FE_FN_ELEM / FE_FN_QP / FE_FN_POLY_U (the polynomial's
// coefficients).
struct feFunction {
  int kind = FE_FN_CONST;
  double c0 = 0., c1 = 0.;
  std::vector<double> table;
};

// A weak form: the reference's elementSystemType id (feSysElm.h:13-77), the
// fields it acts on (numbering field indices, reference space order), the
// quadrature degree and its coefficients.
struct feBilinearForm {
  int elementSystemType;
  std::vector<int> fields;
  int degreeQuadrature;
  std::vector<feFunction> coefficients;
};

feStatus createBilinearForm(feBilinearForm *&form, const std::vector<int> &fields, int elementSystemType,
                            int degreeQuadrature, const std::vector<feFunction> &coefficients);

// The part of feLinearSystem (feLinearSystem.h:85-168) that the Newton
// driver and the examples use.
class feLinearSystem {
public:
  virtual ~feLinearSystem() {}
  virtual int getSystemSize() const = 0;
  void setRelativeTol(double t) { _rel_tol = t; }
  void setAbsoluteTol(double t) { _abs_tol = t; }
  void setDivergenceTol(double t) { _div_tol = t; }
  void setMaxIter(int m) { _max_iter = m; }
  virtual void setToZero() = 0;
  virtual void setMatrixToZero() = 0;
  virtual void setResidualToZero() = 0;
  virtual void assemble(const std::vector<double> &sol) = 0;
  virtual void assembleMatrices(const std::vector<double> &sol) = 0;
  virtual void assembleResiduals(const std::vector<double> &sol) = 0;
  virtual void getRHSMaxNorm(double *norm) const = 0;
  virtual bool solve(double *normDx, double *normResidual, double *normAxb, int *nIter) = 0;
  virtual void correctSolution(std::vector<double> &sol) = 0;
  virtual void constrainEssentialComponents() = 0;
  virtual void applyPeriodicity() = 0;

protected:
  double _rel_tol = 1e-8, _abs_tol = 1e-14, _div_tol = 1e6;
  int _max_iter = 10000;
};

class feLinearSystemHIP : public feLinearSystem {
public:
  feLinearSystemHIP(const feMesh &mesh, const std::vector<feBilinearForm *> &forms, const feMetaNumber *numbering,
                    int device = 0);
  ~feLinearSystemHIP() override;
  feStatus status() const { return _status; }
  std::string lastError() const;
  int getSystemSize() const override { return _n; }
  void setToZero() override;
  void setMatrixToZero() override;
  void setResidualToZero() override;
  void assemble(const std::vector<double> &sol) override;
  void assembleMatrices(const std::vector<double> &sol) override;
  void assembleResiduals(const std::vector<double> &sol) override;
  void getRHSMaxNorm(double *norm) const override;
  bool solve(double *normDx, double *normResidual, double *normAxb, int *nIter) override;
  void correctSolution(std::vector<double> &sol) override;
  void constrainEssentialComponents() override;
  void applyPeriodicity() override;
  // Krylov method / preconditioner / precision (what PETSc reads from
  // -ksp_type / -pc_type in the reference)
  void setSolverOptions(int ksp, int pc, int precision) { _ksp = ksp; _pc = pc; _prec = precision; }
  int64_t nnz() const { return _nnz; }
  fe_hip_ctx *context() const { return _ctx; }

private:
  void upload(const std::vector<double> &sol);
  fe_hip_ctx *_ctx = nullptr;
  feStatus _status = 0;
  int _n = 0, _nDofs = 0;
  int64_t _nnz = 0;
  int _ksp = FE_KSP_CG, _pc = FE_PC_JACOBI, _prec = FE_PREC_FP64;
};

// createLinearSystem (feLinearSystem.cpp:7-67): HIP builds the device system;
// the reference's own back ends are not part of this build.
feStatus createLinearSystem(feLinearSystem *&system, linearSolverType type, const feMesh &mesh,
                            const std::vector<feBilinearForm *> &forms, const feMetaNumber *numbering);

typedef struct feNLSolverOptionsStruct {
  double tolResidual = 1e-10;
  double tolCorrection = 1e-10;
  double tolDivergence = 1e4;
  double maxIter = 20;
} feNLSolverOptions;

// solveNewtonRaphson (feNonLinearSolver.cpp:38-182) on a steady problem:
// residual, stopping tests on ||NL(u)||_inf and ||du||_inf, matrix, solve,
// u += du.  *iterations: Newton steps taken.
feStatus solveNewtonRaphson(feLinearSystem *system, std::vector<double> &sol, const feNLSolverOptions &tol,
                            int *iterations = nullptr);

}  // namespace feng
