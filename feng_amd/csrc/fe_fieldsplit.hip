// Block upper-triangular Schur preconditioner for saddle-point systems
//
//   A = [ A_uu  A_up ]      P = [ A_uu  A_up ]      P^-1 r:  z_p = S^-1 r_p
//       [ A_pu  A_pp ]          [  0     S   ]               z_u = A_uu^-1 (r_u - A_up z_p)
//
// p: the unknowns of the one field whose diagonal entries are all zero (the
// pressure of incompressible flow), u: the others.  This is the role of
// PETSc's PCFIELDSPLIT with -pc_fieldsplit_type schur
// -pc_fieldsplit_schur_fact_type upper -pc_fieldsplit_schur_precondition selfp,
// which a PETSc KSP of the reference can be given through its options
// (feLinearSystemPETSc.cpp:337-352 reads -pc_type; the reference's own
// Navier-Stokes tests factor directly, navier_stokes.cpp:101-108).
//
//   A_uu^-1 ~ one aggregation-multigrid V-cycle (fe_amg*.hip) of the block
//             restricted to entries that couple unknowns of the same field
//             and component (each velocity component a scalar
//             convection-diffusion operator), negated when its diagonal is
//             negative (momentum rows that carry the negated operator); it
//             runs in a sub-context on the parent's stream with its own CSR,
//             whose values are gathered from A at every set-up;
//   S       ~ the diagonal of A_pp - A_pu diag(A_uu)^-1 A_up (selfp), its
//             sign whatever the sign convention of the divergence rows.
#include "fe_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct FeFieldSplit {
  int schurField = -1;
  int nu = 0, np = 0;
  int64_t nnzF = 0;
  int *d_pOf = nullptr;    // unknown -> pressure index, -1 for the other unknowns
  int *d_uRow = nullptr;   // velocity index -> unknown
  int *d_fmap = nullptr;   // entry of the velocity block (sub-context CSR) -> entry of A
  int *d_bia = nullptr, *d_bja = nullptr, *d_bmap = nullptr;  // A_up: velocity rows x pressure columns (CSR,
                                                              // unknown columns) and its entries' positions in A
  double *d_bval = nullptr;
  int64_t nnzB = 0;
  int *d_sia = nullptr;    // per pressure row: its selfp terms
  int4 *d_sent = nullptr;  // (position of a_ik, of a_ki, of a_kk, 0), k a velocity unknown
  int *d_pDiag = nullptr;  // position of each pressure row's diagonal entry (-1: none)
  double *d_sdinv = nullptr;
  double *d_ru = nullptr, *d_zu = nullptr;
  fe_hip_ctx *sub = nullptr;
  bool nested = false;
  double sign = 1.;  // -1 when the velocity block's diagonal is negative: its V-cycle runs on -A_uu
};

namespace {

constexpr int kFsLPR = 8;  // lanes per row of the residual of the velocity rows

__global__ void fsGather(int64_t cnt, const int *__restrict__ map, const double *__restrict__ val, double sign,
                         double *__restrict__ out)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(k < cnt) out[k] = sign * val[map[k]];
}

// 1 / diag(A_pp - A_pu diag(A_uu)^-1 A_up) per pressure row (1 where it is 0).
__global__ void fsSchurDiag(int np, const int *__restrict__ sia, const int4 *__restrict__ sent,
                            const double *__restrict__ val, const int *__restrict__ pDiag, double *__restrict__ sdinv)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if(q >= np) return;
  double s = pDiag[q] >= 0 ? val[pDiag[q]] : 0.;
  for(int t = sia[q]; t < sia[q + 1]; ++t) {
    const int4 e = sent[t];
    s -= val[e.x] * val[e.y] / val[e.z];
  }
  sdinv[q] = (s != 0. && s == s) ? 1. / s : 1.;
}

// z_p = S^-1 r_p, z_u = 0.
__global__ void fsInit(int n, const int *__restrict__ pOf, const double *__restrict__ sdinv,
                       const double *__restrict__ r, double *__restrict__ z)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int q = pOf[i];
  z[i] = q >= 0 ? sdinv[q] * r[i] : 0.;
}

// ru = r_u - A_up z_p over the velocity rows (A_up with its own CSR: the
// pressure columns of each velocity row).
__global__ __launch_bounds__(256) void fsResidualU(int nu, const int *__restrict__ uRow, const int *__restrict__ bia,
                                                   const int *__restrict__ bja, const double *__restrict__ bval,
                                                   const double *__restrict__ z, const double *__restrict__ r,
                                                   double *__restrict__ ru)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = t / kFsLPR, sub = t % kFsLPR;
  const bool on = a < nu;
  const int i = on ? uRow[a] : 0;
  double s = 0.;
  if(on)
    for(int k = bia[a] + sub; k < bia[a + 1]; k += kFsLPR) s += bval[k] * z[bja[k]];
#pragma unroll
  for(int o = kFsLPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kFsLPR);
  if(on && sub == 0) ru[a] = r[i] - s;
}

__global__ void fsScatterU(int nu, const int *__restrict__ uRow, double sign, const double *__restrict__ zu,
                           double *__restrict__ z)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a < nu) z[uRow[a]] = sign * zu[a];
}

template <typename T>
int upload(fe_hip_ctx *ctx, T **d, const std::vector<T> &h)
{
  if(int rc = feAlloc(ctx, d, std::max<size_t>(1, h.size()))) return rc;
  if(!h.empty()) FE_HIP_TRY(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return FE_STATUS_OK;
}

int fsPattern(fe_hip_ctx *ctx)
{
  const int n = ctx->n;
  std::vector<int> ia(n + 1), ja(ctx->nnz), dg(n);
  FE_HIP_TRY(ctx, hipMemcpy(ia.data(), ctx->d_ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  FE_HIP_TRY(ctx, hipMemcpy(ja.data(), ctx->d_ja, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost));
  FE_HIP_TRY(ctx, hipMemcpy(dg.data(), ctx->d_diag, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> val(ctx->nnz);
  FE_HIP_TRY(ctx, hipMemcpy(val.data(), ctx->d_val, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost));
  // field and component of every unknown
  const int nf = (int)ctx->fields.size();
  std::vector<int> fieldOf(n, -1), compOf(n, 0);
  for(int f = 0; f < nf; ++f) {
    const FeField &F = ctx->fields[f];
    std::vector<int> adr((size_t)ctx->E * F.nloc);
    FE_HIP_TRY(ctx, hipMemcpy(adr.data(), F.adr, sizeof(int) * adr.size(), hipMemcpyDeviceToHost));
    for(size_t t = 0; t < adr.size(); ++t) {
      const int d = adr[t];
      if(d < 0 || d >= n) continue;
      fieldOf[d] = f;
      compOf[d] = F.kind == FE_TRI_VP2 ? (int)(t % F.nloc) % 2 : 0;
    }
  }
  // the Schur field: every diagonal entry of its unknowns is zero
  std::vector<int> nonzero(nf, 0), count(nf, 0);
  for(int i = 0; i < n; ++i) {
    FE_REQUIRE(ctx, fieldOf[i] >= 0, "field split: an unknown belongs to no field");
    count[fieldOf[i]]++;
    if(dg[i] >= 0 && val[dg[i]] != 0.) nonzero[fieldOf[i]]++;
  }
  int s = -1;
  for(int f = 0; f < nf; ++f)
    if(count[f] > 0 && nonzero[f] == 0) {
      FE_REQUIRE(ctx, s < 0, "field split: more than one field with a zero diagonal block");
      s = f;
    }
  FE_REQUIRE(ctx, s >= 0 && nf >= 2, "field split: no field with a zero diagonal block (not a saddle-point system)");
  FeFieldSplit &P = *ctx->fieldSplit;
  P.schurField = s;
  std::vector<int> uIdx(n, -1), pOf(n, -1), uRow, pRow;
  for(int i = 0; i < n; ++i) {
    if(fieldOf[i] == s) {
      pOf[i] = (int)pRow.size();
      pRow.push_back(i);
    } else {
      uIdx[i] = (int)uRow.size();
      uRow.push_back(i);
    }
  }
  P.nu = (int)uRow.size();
  P.np = (int)pRow.size();
  // the multigrid wants a positive diagonal: the sign of the velocity block
  // (the momentum rows may carry the negated operator)
  double dsum = 0.;
  for(int i : uRow)
    if(dg[i] >= 0) dsum += val[dg[i]];
  P.sign = dsum < 0. ? -1. : 1.;
  // velocity block on same-field, same-component entries
  std::vector<int> fia(P.nu + 1, 0), fja, fmap;
  for(int a = 0; a < P.nu; ++a) {
    const int i = uRow[a];
    for(int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if(uIdx[j] < 0 || fieldOf[j] != fieldOf[i] || compOf[j] != compOf[i]) continue;
      fja.push_back(uIdx[j]);
      fmap.push_back(k);
    }
    fia[a + 1] = (int)fja.size();
  }
  P.nnzF = (int64_t)fja.size();
  // A_up: the pressure columns of each velocity row
  std::vector<int> bia(P.nu + 1, 0), bja, bmap;
  for(int a = 0; a < P.nu; ++a) {
    const int i = uRow[a];
    for(int k = ia[i]; k < ia[i + 1]; ++k)
      if(pOf[ja[k]] >= 0) {
        bja.push_back(ja[k]);
        bmap.push_back(k);
      }
    bia[a + 1] = (int)bja.size();
  }
  P.nnzB = (int64_t)bja.size();
  // selfp terms of each pressure row: a_ik a_ki / a_kk over its velocity columns k
  std::vector<int> sia(P.np + 1, 0), pDiag(P.np, -1);
  std::vector<int4> sent;
  for(int q = 0; q < P.np; ++q) {
    const int i = pRow[q];
    pDiag[q] = dg[i];
    for(int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if(uIdx[j] < 0 || dg[j] < 0) continue;
      const int *b = ja.data() + ia[j], *e = ja.data() + ia[j + 1];
      const int *it = std::lower_bound(b, e, i);
      if(it == e || *it != i) continue;  // no mirrored entry: no term
      sent.push_back(make_int4(k, (int)(it - ja.data()), dg[j], 0));
    }
    sia[q + 1] = (int)sent.size();
  }
  if(upload(ctx, &P.d_pOf, pOf) || upload(ctx, &P.d_uRow, uRow) || upload(ctx, &P.d_fmap, fmap) ||
     upload(ctx, &P.d_sia, sia) || upload(ctx, &P.d_sent, sent) || upload(ctx, &P.d_pDiag, pDiag) ||
     feAlloc(ctx, &P.d_sdinv, std::max(1, P.np)) || feAlloc(ctx, &P.d_ru, std::max(1, P.nu)) ||
     feAlloc(ctx, &P.d_zu, std::max(1, P.nu)) || upload(ctx, &P.d_bia, bia) || upload(ctx, &P.d_bja, bja) ||
     upload(ctx, &P.d_bmap, bmap) || feAlloc(ctx, &P.d_bval, std::max<int64_t>(1, P.nnzB)))
    return FE_STATUS_OUT_OF_MEMORY;
  // the velocity block's context: CSR of its own, the parent's stream, the
  // parent's multigrid settings
  fe_hip_ctx *sub = nullptr;
  FE_REQUIRE(ctx, fe_hip_ctx_create(ctx->device, &sub) == FE_STATUS_OK, "field split: sub-context creation failed");
  P.sub = sub;
  (void)hipStreamDestroy(sub->stream);
  sub->stream = ctx->stream;
  sub->borrowedStream = true;
  sub->n = P.nu;
  sub->nnz = P.nnzF;
  if(upload(sub, &sub->d_ia, fia) || upload(sub, &sub->d_ja, fja) || feAlloc(sub, &sub->d_val, std::max<int64_t>(1, P.nnzF)))
    return FE_STATUS_OUT_OF_MEMORY;
  sub->csrBuilt = true;
  sub->amgOmega = ctx->amgOmega;
  sub->amgOverCorrect = ctx->amgOverCorrect;
  sub->amgBlockRows = ctx->amgBlockRows;
  sub->amgXcd = ctx->amgXcd;
  sub->amgLanes = ctx->amgLanes;
  sub->amgPadTiles = ctx->amgPadTiles;
  sub->amgKernel = ctx->amgKernel;
  sub->pcValues32 = ctx->pcValues32;
  return FE_STATUS_OK;
}

}  // namespace

int feFieldSplitSetup(fe_hip_ctx *ctx)
{
  if(!ctx->fieldSplit) {
    ctx->fieldSplit = new FeFieldSplit();
    if(int rc = fsPattern(ctx)) return rc;
  }
  FeFieldSplit &P = *ctx->fieldSplit;
  fe_hip_ctx *sub = P.sub;
  if(P.nnzF > 0)
    hipLaunchKernelGGL(fsGather, dim3((unsigned)((P.nnzF + 255) / 256)), dim3(256), 0, ctx->stream, P.nnzF, P.d_fmap,
                       ctx->d_val, P.sign, sub->d_val);
  if(P.nnzB > 0)
    hipLaunchKernelGGL(fsGather, dim3((unsigned)((P.nnzB + 255) / 256)), dim3(256), 0, ctx->stream, P.nnzB, P.d_bmap,
                       ctx->d_val, 1., P.d_bval);
  if(P.np > 0)
    hipLaunchKernelGGL(fsSchurDiag, dim3((P.np + 255) / 256), dim3(256), 0, ctx->stream, P.np, P.d_sia, P.d_sent,
                       ctx->d_val, P.d_pDiag, P.d_sdinv);
  FE_HIP_TRY(ctx, hipGetLastError());
  if(int rc = feAmgSetup(sub)) {
    ctx->setError("field split, velocity block: " + sub->err, __FILE__, __LINE__);
    return rc;
  }
  P.nested = feAmgNestedUsable(sub);
  const int rc = P.nested ? feAmgPcgPrepare(sub) : feAmgNumeric(sub);
  if(rc) {
    ctx->setError("field split, velocity block: " + sub->err, __FILE__, __LINE__);
    return rc;
  }
  return FE_STATUS_OK;
}

int feFieldSplitApply(fe_hip_ctx *ctx, const double *r, double *z)
{
  FeFieldSplit &P = *ctx->fieldSplit;
  const int n = ctx->n;
  hipLaunchKernelGGL(fsInit, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, P.d_pOf, P.d_sdinv, r, z);
  if(P.nu > 0) {
    hipLaunchKernelGGL(fsResidualU, dim3((unsigned)(((int64_t)P.nu * kFsLPR + 255) / 256)), dim3(256), 0, ctx->stream,
                       P.nu, P.d_uRow, P.d_bia, P.d_bja, P.d_bval, z, r, P.d_ru);
    const int rc = P.nested ? feAmgPcgVcycle(P.sub, P.d_ru, P.d_zu)
                            : feAmgApply(P.sub, (const double *)P.d_ru, P.d_zu, nullptr, P.sub->pcValues32 != 0);
    if(rc) {
      ctx->setError("field split, velocity block: " + P.sub->err, __FILE__, __LINE__);
      return rc;
    }
    hipLaunchKernelGGL(fsScatterU, dim3((P.nu + 255) / 256), dim3(256), 0, ctx->stream, P.nu, P.d_uRow, P.sign,
                       (const double *)P.d_zu, z);
  }
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

void feFieldSplitRelease(fe_hip_ctx *ctx)
{
  if(!ctx->fieldSplit) return;
  if(ctx->fieldSplit->sub) fe_hip_ctx_destroy(ctx->fieldSplit->sub);  // device buffers of the parent are its own
  delete ctx->fieldSplit;
  ctx->fieldSplit = nullptr;
}
