// Direct solve of the assembled system on the device: LU with partial
// pivoting of the dense copy of the CSR system (rocSOLVER getrf / getrs), the
// role of the reference's direct back ends -- MKL Pardiso
// (feLinearSystemMklPardiso.cpp:1154-1222) and PETSc's -pc_type lu with
// MUMPS, which its Navier-Stokes and mixed-form tests select
// (tests/withLinearSolver/navier_stokes.cpp:101-108).  Only the role is
// alike: a dense factor in HBM (n^2 doubles), so up to kDirectMax unknowns;
// the FP64 true residual is checked as for the Krylov solves.
#include "fe_hip_internal.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

struct FeDirect {
  rocblas_handle h = nullptr;
  double *A = nullptr, *x = nullptr, *r = nullptr;
  rocblas_int *ipiv = nullptr, *info = nullptr;
  int cap = 0;  // unknowns the buffers hold
};

namespace {

constexpr int kDirectMax = 46000;  // n^2 < 2^31 for the 32-bit indexing of the factorization

// Column-major A[j n + i] = a_ij (A zeroed before), one thread per row.
__global__ void denseScatter(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ val, double *__restrict__ A)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for(int k = ia[i]; k < ia[i + 1]; ++k) A[(size_t)ja[k] * n + i] = val[k];
}

int ensureDirect(fe_hip_ctx *ctx, int n)
{
  if(!ctx->direct) ctx->direct = new FeDirect();
  FeDirect &d = *ctx->direct;
  if(!d.h) {
    FE_REQUIRE(ctx, rocblas_create_handle(&d.h) == rocblas_status_success, "direct solve: rocBLAS handle");
  }
  FE_REQUIRE(ctx, rocblas_set_stream(d.h, ctx->stream) == rocblas_status_success, "direct solve: rocBLAS stream");
  if(d.cap < n) {
    for(void *p : {(void *)d.A, (void *)d.x, (void *)d.r, (void *)d.ipiv, (void *)d.info})
      if(p) feFree(ctx, p);
    d = FeDirect{d.h};
    if(int rc = feAlloc(ctx, &d.A, (size_t)n * n)) return rc;
    if(int rc = feAlloc(ctx, &d.x, (size_t)n)) return rc;
    if(int rc = feAlloc(ctx, &d.r, (size_t)n)) return rc;
    if(int rc = feAlloc(ctx, &d.ipiv, (size_t)n)) return rc;
    if(int rc = feAlloc(ctx, &d.info, 1)) return rc;
    d.cap = n;
  }
  return FE_STATUS_OK;
}

}  // namespace

void feDirectRelease(fe_hip_ctx *ctx)
{
  if(ctx->direct && ctx->direct->h) (void)rocblas_destroy_handle(ctx->direct->h);
  delete ctx->direct;  // device buffers are owned by the context
  ctx->direct = nullptr;
}

int feSolveDirect(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double *du, fe_hip_solve_stats *stats)
{
  const int n = ctx->n;
  FE_REQUIRE(ctx, n >= 1 && n <= kDirectMax,
             "direct solve: the dense factor holds at most " + std::to_string(kDirectMax) + " unknowns");
  if(int rc = ensureDirect(ctx, n)) return rc;
  FeDirect &d = *ctx->direct;
  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(d.A, 0, sizeof(double) * (size_t)n * n, ctx->stream));
  const int g = std::min(4096, (n + 255) / 256);
  hipLaunchKernelGGL(denseScatter, dim3(g), dim3(256), 0, ctx->stream, n, (const int *)ctx->d_ia,
                     (const int *)ctx->d_ja, (const double *)ctx->d_val, d.A);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_REQUIRE(ctx, rocsolver_dgetrf(d.h, n, n, d.A, n, d.ipiv, d.info) == rocblas_status_success,
             "direct solve: rocsolver_dgetrf");
  FE_HIP_TRY(ctx, hipMemcpyAsync(d.x, ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FE_REQUIRE(ctx,
             rocsolver_dgetrs(d.h, rocblas_operation_none, n, 1, d.A, n, d.ipiv, d.x, n) == rocblas_status_success,
             "direct solve: rocsolver_dgetrs");
  rocblas_int info = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&info, d.info, sizeof(info), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, d.x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  FE_REQUIRE(ctx, info == 0, "direct solve: the matrix is singular (zero pivot in column " + std::to_string(info) + ")");
  // FP64 true residual and infinity norms
  double rr = 0., bb = 0.;
  if(int rc = feResidualNorm2(ctx, d.x, d.r, &rr)) return rc;
  std::vector<double> hx(n), hb(n), hr(n);
  FE_HIP_TRY(ctx, hipMemcpyAsync(hx.data(), d.x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipMemcpyAsync(hb.data(), ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipMemcpyAsync(hr.data(), d.r, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  double a = 0., b = 0., c = 0.;
  for(int i = 0; i < n; ++i) {
    a = std::max(a, std::fabs(hx[i]));
    b = std::max(b, std::fabs(hb[i]));
    c = std::max(c, std::fabs(hr[i]));
    bb += hb[i] * hb[i];
  }
  const double res = std::sqrt(rr), bn = std::sqrt(bb);
  const int converged = res <= std::max(o->rtol * bn, o->atol) ? 1 : 0;
  if(stats) {
    memset(stats, 0, sizeof(*stats));
    stats->iterations = 1;
    stats->outer_iterations = 1;
    stats->converged = converged;
    stats->reason = converged ? 1 : -3;
    stats->relres_fp64 = bn > 0. ? res / bn : 0.;
    stats->norm_dx_inf = a;
    stats->norm_rhs_inf = b;
    stats->norm_axb_inf = c;
    stats->ms_solve = ms;
  }
  if(du) std::memcpy(du, hx.data(), sizeof(double) * n);
  return FE_STATUS_OK;
}
