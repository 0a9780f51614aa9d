// Preconditioners of the device Krylov layer.
//
// ILU(0) / IC(0) factorization (the PC_ILU / PC_ICC the reference selects
// through PETSc options, feLinearSystemPETSc.cpp:337-352): exact IKJ
// incomplete factorization on the CSR pattern of A, row-level scheduled: the
// host computes the dependency level of every row (level(i) = 1 + max level
// of the columns j < i of row i) once per pattern; the device factors one
// level per launch, one wavefront per row, with the row held in LDS.
//
// Split factored approximate inverse on the IC(0) pattern:
//   M^-1 = G^T D^-1 G,  G unit lower triangular on the lower pattern of A.
// From IC(0) (A ~ L D L^T): G = I - L_strict (the first Neumann term of
// L^-1), D = the IC(0) pivots.  G (lower CSR) and G^T (upper CSR, with the
// transpose map of the symmetric pattern) are stored explicitly so the
// learned correction (G + dG) uses the same apply kernels.
#include "fe_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int kRowCap = 256;  // max row length held in LDS by the factor kernel

// One wavefront per row of the level: IKJ ILU(0) of row i, in place in lu.
__global__ __launch_bounds__(64) void iluFactorLevel(const int *__restrict__ rows, int count,
                                                     const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const int *__restrict__ diag, double *__restrict__ lu,
                                                     double shiftRel, int *__restrict__ nShift)
{
  __shared__ int cols[kRowCap];
  __shared__ double vals[kRowCap];
  if((int)blockIdx.x >= count) return;
  const int i = rows[blockIdx.x];
  const int b = ia[i], e = ia[i + 1], len = e - b;
  const int lane = threadIdx.x;
  if(len > kRowCap) {
    if(lane == 0) atomicAdd(nShift, 1 << 20);  // flagged: row too long
    return;
  }
  double rowAbs = 0.;
  for(int t = lane; t < len; t += 64) {
    cols[t] = ja[b + t];
    vals[t] = lu[b + t];
    rowAbs = fmax(rowAbs, fabs(vals[t]));
  }
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) rowAbs = fmax(rowAbs, __shfl_xor(rowAbs, o, 64));
  __syncthreads();
  for(int t = 0; t < len; ++t) {
    const int k = cols[t];
    if(k >= i) break;
    const double piv = lu[diag[k]];
    const double l = vals[t] / piv;
    __syncthreads();
    if(lane == 0) vals[t] = l;
    // row k beyond its diagonal: subtract l * u_kj where (i, j) is in the pattern
    const int kb = diag[k] + 1, ke = ia[k + 1];
    for(int q = kb + lane; q < ke; q += 64) {
      const int j = ja[q];
      int lo = t + 1, hi = len - 1;
      while(lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = cols[mid];
        if(c == j) {
          vals[mid] -= l * lu[q];
          break;
        }
        if(c < j) lo = mid + 1; else hi = mid - 1;
      }
    }
    __syncthreads();
  }
  // zero pivot (saddle-point rows): shift it, relative to the row scale
  for(int t = lane; t < len; t += 64) {
    if(cols[t] == i && fabs(vals[t]) <= 1e-14 * rowAbs) {
      vals[t] = shiftRel * (rowAbs > 0. ? rowAbs : 1.);
      atomicAdd(nShift, 1);
    }
  }
  __syncthreads();
  for(int t = lane; t < len; t += 64) lu[b + t] = vals[t];
}

// Row counts of the lower (j <= i) and upper (j >= i) parts.
__global__ void countParts(int n, const int *__restrict__ ia, const int *__restrict__ ja, const int *__restrict__ diag,
                           int *__restrict__ cL, int *__restrict__ cU)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  cL[i] = diag[i] - ia[i] + 1;
  cU[i] = ia[i + 1] - diag[i];
}

// Lower part with the unit diagonal and values g_ij = -l_ij (G = I - L_strict);
// upper part G^T with values -u_ij / u_ii and unit diagonal; D = u_ii.  One
// thread per entry of A (row from rowOf): the lower entry (i, j) goes to
// position iaL[i] + (p - ia[i]) of G (the diagonal last), the upper entry to
// iaU[i] + (p - diag[i]) of G^T (the diagonal first).
template <typename T>
__global__ void buildSplitFactor(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ia,
                                 const int *__restrict__ ja, const int *__restrict__ diag,
                                 const double *__restrict__ lu, const int *__restrict__ iaL, int *__restrict__ jaL,
                                 T *__restrict__ gL, const int *__restrict__ iaU, int *__restrict__ jaU,
                                 T *__restrict__ gU, T *__restrict__ dinv, double *__restrict__ dval)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= nnz) return;
  const int i = rowOf[p], j = ja[p], d = diag[i];
  if(j <= i) {
    const int o = iaL[i] + (int)(p - ia[i]);
    jaL[o] = j;
    gL[o] = (T)(j == i ? 1. : -lu[p]);
  }
  if(j >= i) {
    const int o = iaU[i] + (int)(p - d);
    jaU[o] = j;
    gU[o] = (T)(j == i ? 1. : -lu[p] / lu[d]);
  }
  if(j == i) {
    dinv[i] = (T)(1. / lu[p]);
    dval[i] = lu[p];
  }
}

// Row of every entry of A.
__global__ void rowIndex(int n, const int *__restrict__ ia, int *__restrict__ rowOf)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  for(int p = ia[i]; p < ia[i + 1]; ++p) rowOf[p] = i;
}


// Position of the mirrored entry (j, i) of every entry (i, j); -1 (and the
// flag) when the pattern is not structurally symmetric there.
__global__ void transposeMap(int n, const int *__restrict__ ia, const int *__restrict__ ja, int *__restrict__ tpos,
                             int *__restrict__ missing)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  for(int p = ia[i]; p < ia[i + 1]; ++p) {
    const int j = ja[p];
    int lo = ia[j], hi = ia[j + 1] - 1, q = -1;
    while(lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int c = ja[mid];
      if(c == i) {
        q = mid;
        break;
      }
      if(c < i) lo = mid + 1;
      else hi = mid - 1;
    }
    tpos[p] = q;
    if(q < 0) missing[0] = 1;
  }
}

// Fine-grained parallel IC(0) (Chow & Patel, SIAM J. Sci. Comput. 37 (2015)
// C169-C193) in the form A = L D L^T, L unit lower on the strict lower
// pattern of A: Jacobi-style sweeps of the fixed-point equations
//   l_ij = (a_ij - sum_{k<j} l_ik d_k l_jk) / d_j   (j < i),
//   d_i  = a_ii - sum_{k<i} l_ik^2 d_k,
// from l_ij = a_ij / a_jj, d_i = a_ii.  Its fixed point is the exact IC(0)
// factor; a few sweeps give a factor of the same quality as preconditioner.
__global__ void icInit(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ja,
                       const int *__restrict__ diag, const double *__restrict__ val, double *__restrict__ l,
                       double *__restrict__ d)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= nnz) return;
  const int i = rowOf[p], j = ja[p];
  if(j < i) l[p] = val[p] / val[diag[j]];
  else if(j == i) d[i] = val[p];
}

// One sweep, 8 lanes per row: lane q handles the strict lower entries
// ia[i] + q, + 8, ...; lane 0 also the pivot.
__global__ __launch_bounds__(256) void icSweep(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const int *__restrict__ diag, const double *__restrict__ val,
                                               const double *__restrict__ lOld, const double *__restrict__ dOld,
                                               double *__restrict__ lNew, double *__restrict__ dNew)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> 3, lane = t & 7;
  if(i >= n) return;
  const int b = ia[i], pd = diag[i];
  for(int e = b + lane; e < pd; e += 8) {
    const int j = ja[e];
    double s = val[e];
    // sum over k < j of l_ik d_k l_jk: merge row i's lower entries (cols < j)
    // with row j's lower entries
    int a = b, c = ia[j];
    const int ce = diag[j];
    while(a < e && c < ce) {
      const int ka = ja[a], kc = ja[c];
      if(ka == kc) {
        s -= lOld[a] * dOld[ka] * lOld[c];
        ++a;
        ++c;
      } else if(ka < kc) {
        ++a;
      } else {
        ++c;
      }
    }
    lNew[e] = s / dOld[j];
  }
  if(lane == 0) {
    double s = val[pd];
    for(int e = b; e < pd; ++e) s -= lOld[e] * lOld[e] * dOld[ja[e]];
    dNew[i] = s > 1e-8 * val[pd] ? s : 1e-8 * val[pd];
  }
}

// ILU(0) storage of the LDL^T factor: lu = l on the strict lower part, d on
// the diagonal, u_ij = d_i l_ji on the strict upper part (one thread per entry).
__global__ void icToLu(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ja,
                       const int *__restrict__ tpos, const double *__restrict__ l, const double *__restrict__ d,
                       double *__restrict__ lu)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= nnz) return;
  const int i = rowOf[p], j = ja[p];
  lu[p] = j < i ? l[p] : (j == i ? d[i] : d[i] * l[tpos[p]]);
}

}  // namespace

// Host: dependency levels of the lower-triangular DAG of the pattern.
int feComputeLevels(fe_hip_ctx *ctx, std::vector<int> &levelPtr, std::vector<int> &levelRows)
{
  const int n = ctx->n;
  std::vector<int> ia(n + 1), ja(ctx->nnz);
  FE_HIP_TRY(ctx, hipMemcpy(ia.data(), ctx->d_ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  FE_HIP_TRY(ctx, hipMemcpy(ja.data(), ctx->d_ja, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost));
  std::vector<int> lev(n, 0);
  int nl = 0;
  for(int i = 0; i < n; ++i) {
    int l = 0;
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      const int j = ja[p];
      if(j >= i) break;
      l = std::max(l, lev[j] + 1);
    }
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  levelPtr.assign(nl + 1, 0);
  for(int i = 0; i < n; ++i) levelPtr[lev[i] + 1]++;
  for(int l = 0; l < nl; ++l) levelPtr[l + 1] += levelPtr[l];
  std::vector<int> pos(levelPtr.begin(), levelPtr.end() - 1);
  levelRows.assign(n, 0);
  for(int i = 0; i < n; ++i) levelRows[pos[lev[i]]++] = i;
  return FE_STATUS_OK;
}

// Host: dependency levels of the upper-triangular DAG (level(i) = 1 + max
// level of the columns j > i of row i), rows grouped by level.
int feComputeLevelsUpper(fe_hip_ctx *ctx, std::vector<int> &levelPtr, std::vector<int> &levelRows)
{
  const int n = ctx->n;
  std::vector<int> ia(n + 1), ja(ctx->nnz);
  FE_HIP_TRY(ctx, hipMemcpy(ia.data(), ctx->d_ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  FE_HIP_TRY(ctx, hipMemcpy(ja.data(), ctx->d_ja, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost));
  std::vector<int> lev(n, 0);
  int nl = 0;
  for(int i = n - 1; i >= 0; --i) {
    int l = 0;
    for(int p = ia[i + 1] - 1; p >= ia[i]; --p) {
      const int j = ja[p];
      if(j <= i) break;
      l = std::max(l, lev[j] + 1);
    }
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  levelPtr.assign(nl + 1, 0);
  for(int i = 0; i < n; ++i) levelPtr[lev[i] + 1]++;
  for(int l = 0; l < nl; ++l) levelPtr[l + 1] += levelPtr[l];
  std::vector<int> pos(levelPtr.begin(), levelPtr.end() - 1);
  levelRows.assign(n, 0);
  for(int i = 0; i < n; ++i) levelRows[pos[lev[i]]++] = i;
  return FE_STATUS_OK;
}

int fePrecondSetupPattern(fe_hip_ctx *ctx)
{
  FePrecond &P = ctx->pc;
  if(P.patternReady) return FE_STATUS_OK;
  std::vector<int> lrows, urows;
  if(int rc = feComputeLevels(ctx, P.levelPtr, lrows)) return rc;
  if(int rc = feComputeLevelsUpper(ctx, P.levelPtrU, urows)) return rc;
  const int n = ctx->n;
  if(feAlloc(ctx, &P.d_levelRows, n) || feAlloc(ctx, &P.d_lu, (size_t)ctx->nnz) || feAlloc(ctx, &P.d_iaL, n + 1) ||
     feAlloc(ctx, &P.d_iaU, n + 1) || feAlloc(ctx, &P.d_nShift, 4) || feAlloc(ctx, &P.d_levelRowsU, n) ||
     feAlloc(ctx, &P.d_levelPtr, P.levelPtr.size()) || feAlloc(ctx, &P.d_levelPtrU, P.levelPtrU.size()))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(P.d_levelRows, lrows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(P.d_levelRowsU, urows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(P.d_levelPtr, P.levelPtr.data(), sizeof(int) * P.levelPtr.size(), hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(P.d_levelPtrU, P.levelPtrU.data(), sizeof(int) * P.levelPtrU.size(),
                            hipMemcpyHostToDevice));
  if(feAlloc(ctx, &P.d_tposA, (size_t)ctx->nnz) || feAlloc(ctx, &P.d_rowOf, (size_t)ctx->nnz))
    return FE_STATUS_OUT_OF_MEMORY;
  hipLaunchKernelGGL(rowIndex, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_ia, P.d_rowOf);
  FE_HIP_TRY(ctx, hipMemsetAsync(P.d_nShift, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(transposeMap, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja,
                     P.d_tposA, P.d_nShift);
  int missing = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&missing, P.d_nShift, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  P.symmetricPattern = (missing == 0);
  int *cL = nullptr, *cU = nullptr;
  if(feAlloc(ctx, &cL, n) || feAlloc(ctx, &cU, n)) return FE_STATUS_OUT_OF_MEMORY;
  hipLaunchKernelGGL(countParts, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, ctx->d_diag,
                     cL, cU);
  int64_t nnzL = 0, nnzU = 0;
  if(int rc = feExclusiveScan(ctx, cL, P.d_iaL, n, &nnzL)) return rc;
  if(int rc = feExclusiveScan(ctx, cU, P.d_iaU, n, &nnzU)) return rc;
  feFree(ctx, cL);
  feFree(ctx, cU);
  P.nnzL = nnzL;
  P.nnzU = nnzU;
  if(feAlloc(ctx, &P.d_jaL, nnzL) || feAlloc(ctx, &P.d_jaU, nnzU) || feAlloc(ctx, &P.d_gL64, nnzL) ||
     feAlloc(ctx, &P.d_gU64, nnzU) || feAlloc(ctx, &P.d_gL32, nnzL) || feAlloc(ctx, &P.d_gU32, nnzU) ||
     feAlloc(ctx, &P.d_dinv64, n) || feAlloc(ctx, &P.d_dinv32, n) || feAlloc(ctx, &P.d_dval, n))
    return FE_STATUS_OUT_OF_MEMORY;
  P.patternReady = true;
  return FE_STATUS_OK;
}

// ILU(0) of the current FP64 values into P.d_lu -- exact, level by level
// (sweeps = 0), or `sweeps` sweeps of the parallel IC(0) iteration (symmetric
// matrices) -- then the split factors G, G^T and D^-1 in both precisions.
int fePrecondFactor(fe_hip_ctx *ctx, int sweeps)
{
  if(int rc = fePrecondSetupPattern(ctx)) return rc;
  FePrecond &P = ctx->pc;
  const int n = ctx->n;
  FE_HIP_TRY(ctx, hipMemsetAsync(P.d_nShift, 0, sizeof(int), ctx->stream));
  if(sweeps <= 0) {
    FE_HIP_TRY(ctx, hipMemcpyAsync(P.d_lu, ctx->d_val, sizeof(double) * ctx->nnz, hipMemcpyDeviceToDevice, ctx->stream));
    const int nl = (int)P.levelPtr.size() - 1;
    for(int l = 0; l < nl; ++l) {
      const int b = P.levelPtr[l], cnt = P.levelPtr[l + 1] - b;
      hipLaunchKernelGGL(iluFactorLevel, dim3(cnt), dim3(64), 0, ctx->stream, P.d_levelRows + b, cnt, ctx->d_ia,
                         ctx->d_ja, ctx->d_diag, P.d_lu, 1e-8, P.d_nShift);
    }
  } else {
    FE_REQUIRE(ctx, P.symmetricPattern, "the parallel IC(0) needs a structurally symmetric matrix");
    if(!P.d_icL) {
      if(feAlloc(ctx, &P.d_icL, 2 * (size_t)ctx->nnz) || feAlloc(ctx, &P.d_icD, 2 * (size_t)n))
        return FE_STATUS_OUT_OF_MEMORY;
    }
    double *L[2] = {P.d_icL, P.d_icL + ctx->nnz}, *D[2] = {P.d_icD, P.d_icD + n};
    const int gE = (int)((ctx->nnz + 255) / 256);
    hipLaunchKernelGGL(icInit, dim3(gE), dim3(256), 0, ctx->stream, ctx->nnz, P.d_rowOf, ctx->d_ja, ctx->d_diag,
                       ctx->d_val, L[0], D[0]);
    for(int s = 0; s < sweeps; ++s)
      hipLaunchKernelGGL(icSweep, dim3((int)(((int64_t)n * 8 + 255) / 256)), dim3(256), 0, ctx->stream, n, ctx->d_ia,
                         ctx->d_ja, ctx->d_diag, ctx->d_val, L[s & 1], D[s & 1], L[(s + 1) & 1], D[(s + 1) & 1]);
    hipLaunchKernelGGL(icToLu, dim3(gE), dim3(256), 0, ctx->stream, ctx->nnz, P.d_rowOf, ctx->d_ja, P.d_tposA,
                       L[sweeps & 1], D[sweeps & 1], P.d_lu);
  }
  FE_HIP_TRY(ctx, hipGetLastError());
  const int gE2 = (int)((ctx->nnz + 255) / 256);
  hipLaunchKernelGGL(buildSplitFactor<double>, dim3(gE2), dim3(256), 0, ctx->stream, ctx->nnz, P.d_rowOf, ctx->d_ia,
                     ctx->d_ja, ctx->d_diag, P.d_lu, P.d_iaL, P.d_jaL, P.d_gL64, P.d_iaU, P.d_jaU, P.d_gU64,
                     P.d_dinv64, P.d_dval);
  hipLaunchKernelGGL(buildSplitFactor<float>, dim3(gE2), dim3(256), 0, ctx->stream, ctx->nnz, P.d_rowOf, ctx->d_ia,
                     ctx->d_ja, ctx->d_diag, P.d_lu, P.d_iaL, P.d_jaL, P.d_gL32, P.d_iaU, P.d_jaU, P.d_gU32,
                     P.d_dinv32, P.d_dval);
  FE_HIP_TRY(ctx, hipGetLastError());
  int shifts = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&shifts, P.d_nShift, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, shifts < (1 << 20), "ILU(0): a row is longer than the 256 entries held in LDS");
  P.lastShifts = shifts;
  P.factored = true;
  return FE_STATUS_OK;
}

extern "C" int fe_hip_pc_info(fe_hip_ctx *ctx, int *levels, int *shifts)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  if(levels) *levels = ctx->pc.levelPtr.empty() ? 0 : (int)ctx->pc.levelPtr.size() - 1;
  if(shifts) *shifts = ctx->pc.lastShifts;
  return FE_STATUS_OK;
}

extern "C" int fe_hip_get_factor(fe_hip_ctx *ctx, double *lu)
{
  if(!ctx || !lu) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->pc.factored, "no factorization yet");
  FE_HIP_TRY(ctx, hipMemcpy(lu, ctx->pc.d_lu, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost));
  return FE_STATUS_OK;
}
