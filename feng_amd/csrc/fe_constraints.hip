// Constraints on the assembled system (the reference's linear-system hooks
// called by Newton after the matrix assembly, feNonLinearSolver.cpp:94-95):
//   constrainEssentialComponents -- rows of essential components of vector
//     spaces: their columns and rows zeroed, 1 on the diagonal, 0 in the RHS
//     (MatZeroRowsColumns with zero values, feLinearSystemPETSc.cpp:816-853;
//     the same loops in feLinearSystemMklPardiso.cpp:994-1097);
//   applyPeriodicity -- the slave row of each periodic pair becomes
//     u_slave - u_master = 0 (feLinearSystemMklPardiso.cpp:1099-1127).
// The first is one pass over the rows: an entry (i, j) becomes 1 / 0 when
// row i is constrained and 0 when only column j is, which is the final state
// of the reference's sequential column-then-row loops whatever their order.
#include "fe_hip_internal.h"

#include <algorithm>

namespace {

__global__ void zeroRowsColumns(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                const int *__restrict__ flag, double *__restrict__ val, float *__restrict__ val32,
                                double *__restrict__ rhs)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= n) return;
  const bool rowC = flag[r] != 0;
  for(int p = ia[r]; p < ia[r + 1]; ++p) {
    const int j = ja[p];
    if(rowC || flag[j]) {
      const double v = (rowC && j == r) ? 1. : 0.;
      val[p] = v;
      if(val32) val32[p] = (float)v;
    }
  }
  if(rowC) rhs[r] = 0.;
}

// One thread per slave row: zero, 1 at (s, s), then -1 at (s, m) with m the
// last master the reference's map order assigns to this slave.
__global__ void periodicRows(int count, const int *__restrict__ slave, const int *__restrict__ master,
                             const int *__restrict__ ia, const int *__restrict__ ja, double *__restrict__ val,
                             float *__restrict__ val32, double *__restrict__ rhs)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= count) return;
  const int s = slave[k], m = master[k];
  for(int p = ia[s]; p < ia[s + 1]; ++p) {
    const int j = ja[p];
    double v = 0.;
    if(j == s) v = 1.;
    if(j == m) v = -1.;
    val[p] = v;
    if(val32) val32[p] = (float)v;
  }
  rhs[s] = 0.;
}

}  // namespace

extern "C" int fe_hip_set_constrained_dofs(fe_hip_ctx *ctx, int count, const int *dofs)
{
  if(!ctx || (count > 0 && !dofs)) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  FE_REQUIRE(ctx, count >= 0, "negative count");
  // DOFs shared with an essential space carry numbers >= nInc and are skipped
  // (feLinearSystemMklPardiso.cpp:1019-1033)
  std::vector<int> rows;
  rows.reserve(count);
  for(int k = 0; k < count; ++k) {
    FE_REQUIRE(ctx, dofs[k] >= 0 && dofs[k] < ctx->nDofs, "constrained DOF out of range");
    if(dofs[k] < ctx->n) rows.push_back(dofs[k]);
  }
  std::sort(rows.begin(), rows.end());
  rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
  ctx->conRows = rows;
  if(!ctx->d_conFlag && feAlloc(ctx, &ctx->d_conFlag, (size_t)ctx->n)) return FE_STATUS_OUT_OF_MEMORY;
  std::vector<int> flag(ctx->n, 0);
  for(int r : rows) flag[r] = 1;
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_conFlag, flag.data(), sizeof(int) * ctx->n, hipMemcpyHostToDevice,
                                 ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

extern "C" int fe_hip_constrain_essential_components(fe_hip_ctx *ctx)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(ctx->conRows.empty()) return FE_STATUS_OK;
  hipLaunchKernelGGL(zeroRowsColumns, dim3((ctx->n + 255) / 256), dim3(256), 0, ctx->stream, ctx->n, ctx->d_ia,
                     ctx->d_ja, ctx->d_conFlag, ctx->d_val, ctx->d_val32, ctx->d_rhs);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

extern "C" int fe_hip_set_periodic_dofs(fe_hip_ctx *ctx, int count, const int *master, const int *slave)
{
  if(!ctx || (count > 0 && (!master || !slave))) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, !ctx->csrBuilt, "periodic pairs enter the CSR pattern: set them before fe_hip_build_csr");
  FE_REQUIRE(ctx, ctx->n > 0, "set the system sizes first");
  FE_REQUIRE(ctx, count >= 0, "negative count");
  // the reference's std::map<master, slave>: one slave per master, masters
  // ascending; only pairs of two unknowns enter the system
  std::vector<std::pair<int, int>> pairs;
  for(int k = 0; k < count; ++k) {
    FE_REQUIRE(ctx, master[k] >= 0 && master[k] < ctx->nDofs && slave[k] >= 0 && slave[k] < ctx->nDofs,
               "periodic DOF out of range");
    pairs.push_back({master[k], slave[k]});
  }
  std::sort(pairs.begin(), pairs.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  for(size_t k = 1; k < pairs.size(); ++k)
    FE_REQUIRE(ctx, pairs[k].first != pairs[k - 1].first, "a master DOF appears twice (one slave per master)");
  ctx->perMaster.clear();
  ctx->perSlave.clear();
  std::vector<int> last(ctx->n, -1);  // the master the last pair of each slave sets
  for(const auto &p : pairs)
    if(p.first < ctx->n && p.second < ctx->n) {
      ctx->perMaster.push_back(p.first);
      ctx->perSlave.push_back(p.second);
      last[p.second] = p.first;
    }
  std::vector<int> hs, hm;
  for(int s = 0; s < ctx->n; ++s)
    if(last[s] >= 0) {
      hs.push_back(s);
      hm.push_back(last[s]);
    }
  ctx->nPerRows = (int)hs.size();
  if(ctx->nPerRows) {
    if(feAlloc(ctx, &ctx->d_perSlave, hs.size()) || feAlloc(ctx, &ctx->d_perMaster, hm.size()))
      return FE_STATUS_OUT_OF_MEMORY;
    FE_HIP_TRY(ctx, hipMemcpy(ctx->d_perSlave, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice));
    FE_HIP_TRY(ctx, hipMemcpy(ctx->d_perMaster, hm.data(), sizeof(int) * hm.size(), hipMemcpyHostToDevice));
  }
  return FE_STATUS_OK;
}

extern "C" int fe_hip_apply_periodicity(fe_hip_ctx *ctx)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(ctx->nPerRows == 0) return FE_STATUS_OK;
  hipLaunchKernelGGL(periodicRows, dim3((ctx->nPerRows + 255) / 256), dim3(256), 0, ctx->stream, ctx->nPerRows,
                     ctx->d_perSlave, ctx->d_perMaster, ctx->d_ia, ctx->d_ja, ctx->d_val, ctx->d_val32, ctx->d_rhs);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}
