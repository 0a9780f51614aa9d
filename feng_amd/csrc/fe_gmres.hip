// Restarted GMRES(m) on the device (the KSPGMRES the reference's
// Navier-Stokes runs select, feLinearSystemPETSc.cpp:900-1031; restart 30 by
// PETSc default, 50 in the configs), FP64, right preconditioned so the
// monitored residual is the true residual of A x = b.
//
// Arnoldi uses classical Gram-Schmidt (PETSc's default orthogonalization,
// no refinement), or with one reorthogonalization (CGS2, `gmres_orth` 2): all j+1
// projections of a step come from one multi-dot kernel and one multi-axpy
// kernel, instead of j+1 dependent dot products.  The small
// Hessenberg least-squares problem (Givens rotations) lives on the host; the
// host reads j+2 scalars per Arnoldi step.
//
// Preconditioners: Jacobi, or ILU(0) (exact incomplete factors from
// fe_precond.hip, PETSc's PCILU) applied by exact triangular solves
// scheduled by dependency levels -- y = L^-1 r over the levels of the lower
// triangle, z = U^-1 y over those of the upper one; the rows of a level are
// independent, a level of many rows is one launch, a run of consecutive
// levels of few rows is one workgroup with a barrier between levels -- or,
// with `ilu_sweeps` = k > 0, by k Jacobi iterations per factor (the parallel
// approximate triangular solve: y <- r - L_strict y, z <- D^-1 (y - U_strict z)).
#include "fe_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kTB = 256;
constexpr int kG = 1024;

template <int LPR>
__global__ __launch_bounds__(kTB) void spmv64(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                              const double *__restrict__ val, const double *__restrict__ x,
                                              double *__restrict__ y)
{
  constexpr int RPB = kTB / LPR;
  const int sub = threadIdx.x % LPR;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = blockIdx.x; chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    double s = 0.;
    if(row < n)
      s = feRowDot<FeKs<LPR>::value, LPR>(ja, val, ia[row] + sub, ia[row + 1], [&](int j) { return x[j]; });
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
    if(sub == 0 && row < n) y[row] = s;
  }
}

// out_i = scale_i * in0_i + sum_{j != i} g_ij v_j over a unit-diagonal factor.
__global__ __launch_bounds__(kTB) void factorSweep(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                   const double *__restrict__ g, const double *__restrict__ scale,
                                                   const double *__restrict__ in0, const double *__restrict__ v,
                                                   double *__restrict__ out)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = scale ? scale[i] * in0[i] : in0[i];
    for(int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if(j != i) s += g[k] * (scale ? v[j] : v[j]);
    }
    out[i] = s;
  }
}

constexpr int kTriLPR = 8;      // lanes per row of the exact triangular solves
constexpr int kTriRunTB = 1024;  // workgroup of a run of small levels
constexpr int kTriSmall = 256;   // levels of at most this many rows join a run

// Row i of an exact triangular solve: out_i = (scale ? scale_i in0_i : in0_i)
// + sum_{j != i} g_ij out_j, the out_j of earlier levels.  kTriLPR lanes per
// row; every lane of the group takes part in the reduction.
__device__ __forceinline__ void triRow(bool on, int i, int sub, const int *__restrict__ ia, const int *__restrict__ ja,
                                       const double *__restrict__ g, const double *__restrict__ scale,
                                       const double *__restrict__ in0, double *out)
{
  double s = 0.;
  if(on)
    for(int k = ia[i] + sub; k < ia[i + 1]; k += kTriLPR) {
      const int j = ja[k];
      if(j != i) s += g[k] * out[j];
    }
#pragma unroll
  for(int o = kTriLPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kTriLPR);
  if(on && sub == 0) out[i] = (scale ? scale[i] * in0[i] : in0[i]) + s;
}

// One level of many rows.
__global__ __launch_bounds__(256) void triLevel(const int *__restrict__ rows, int count, const int *__restrict__ ia,
                                                const int *__restrict__ ja, const double *__restrict__ g,
                                                const double *__restrict__ scale, const double *__restrict__ in0,
                                                double *out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = t / kTriLPR, sub = t % kTriLPR;
  const bool on = r < count;
  triRow(on, on ? rows[r] : 0, sub, ia, ja, g, scale, in0, out);
}

// Levels l0 .. l1 - 1, each of few rows, in one workgroup: a barrier between
// levels orders the writes of one level before the reads of the next.
__global__ __launch_bounds__(kTriRunTB) void triLevelRun(const int *__restrict__ levelPtr, int l0, int l1,
                                                         const int *__restrict__ rows, const int *__restrict__ ia,
                                                         const int *__restrict__ ja, const double *__restrict__ g,
                                                         const double *__restrict__ scale,
                                                         const double *__restrict__ in0, double *out)
{
  constexpr int RPP = kTriRunTB / kTriLPR;  // rows per pass
  const int sub = threadIdx.x % kTriLPR;
  for(int l = l0; l < l1; ++l) {
    const int b = levelPtr[l], cnt = levelPtr[l + 1] - b;
    for(int r0 = 0; r0 < cnt; r0 += RPP) {
      const int r = r0 + (int)threadIdx.x / kTriLPR;
      const bool on = r < cnt;
      triRow(on, on ? rows[b + r] : 0, sub, ia, ja, g, scale, in0, out);
    }
    __syncthreads();
  }
}

// Segments of the exact solves over the levels in levelPtr.
std::vector<FePrecond::TriSeg> triPlan(const std::vector<int> &levelPtr)
{
  std::vector<FePrecond::TriSeg> plan;
  const int nl = (int)levelPtr.size() - 1;
  int runStart = -1;
  for(int l = 0; l < nl; ++l) {
    const int cnt = levelPtr[l + 1] - levelPtr[l];
    if(cnt <= kTriSmall) {
      if(runStart < 0) runStart = l;
      continue;
    }
    if(runStart >= 0) plan.push_back({runStart, l, true});
    runStart = -1;
    plan.push_back({l, l + 1, false});
  }
  if(runStart >= 0) plan.push_back({runStart, nl, true});
  return plan;
}

__global__ void scaleVec(int n, const double *__restrict__ d, const double *__restrict__ x, double *__restrict__ y)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = d[i] * x[i];
}

__global__ void jacobiDiag(int n, const int *__restrict__ diag, const double *__restrict__ val,
                           double *__restrict__ dinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double d = diag[i] >= 0 ? val[diag[i]] : 0.;
  dinv[i] = d != 0. ? 1. / d : 1.;
}

// partials[k][block] = sum_i V_k[i] w[i], k <= j
__global__ __launch_bounds__(kTB) void multiDot(int n, int nv, const double *__restrict__ V, size_t ld,
                                                const double *__restrict__ w, double *__restrict__ partials)
{
  __shared__ double lds[kTB / 64 + 1];
  for(int k = 0; k < nv; ++k) {
    double s = 0.;
    const double *vk = V + (size_t)k * ld;
    for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += vk[i] * w[i];
    feStoreBlockSum(s, partials + (size_t)k * gridDim.x, lds);
  }
}

__global__ __launch_bounds__(kTB) void foldMany(const double *__restrict__ partials, int G, double *__restrict__ out)
{
  __shared__ double lds[kTB / 64 + 1];
  const double s = feFoldPartials(partials + (size_t)blockIdx.x * G, G, lds);
  if(threadIdx.x == 0) out[blockIdx.x] = s;
}

// w -= sum_k h_k V_k (k < nv); optionally accumulate h into hAcc
__global__ __launch_bounds__(kTB) void multiAxpy(int n, int nv, const double *__restrict__ V, size_t ld,
                                                 const double *__restrict__ h, double *__restrict__ w)
{
  __shared__ double hs[128];
  for(int k = threadIdx.x; k < nv; k += blockDim.x) hs[k] = h[k];
  __syncthreads();
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = w[i];
    for(int k = 0; k < nv; ++k) s -= hs[k] * V[(size_t)k * ld + i];
    w[i] = s;
  }
}

// x += sum_k y_k V_k
__global__ __launch_bounds__(kTB) void multiCombine(int n, int nv, const double *__restrict__ V, size_t ld,
                                                    const double *__restrict__ y, double *__restrict__ x)
{
  __shared__ double ys[128];
  for(int k = threadIdx.x; k < nv; k += blockDim.x) ys[k] = y[k];
  __syncthreads();
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.;
    for(int k = 0; k < nv; ++k) s += ys[k] * V[(size_t)k * ld + i];
    x[i] = s;
  }
}

__global__ void scaleInto(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[i];
}

__global__ void axpyKernel(int n, double a, const double *__restrict__ x, double *__restrict__ y)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] += a * x[i];
}

__global__ void residualKernel(int n, const double *__restrict__ b, const double *__restrict__ ax,
                               double *__restrict__ r)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r[i] = b[i] - ax[i];
}

struct Gm {
  fe_hip_ctx *ctx;
  int n, m;
  size_t ld;
  double *V, *w, *z, *t1, *t2, *x, *r, *dinvJ, *part, *h;
  int pc;
  int sweeps;

  void spmv(const double *x_, double *y_)
  {
    const int L = ctx->spmvLanes;
#define FE_SP(LL) hipLaunchKernelGGL(spmv64<LL>, dim3(kG), dim3(kTB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, ctx->d_val, x_, y_)
    switch(L) {
      case 1: FE_SP(1); break;
      case 2: FE_SP(2); break;
      case 4: FE_SP(4); break;
      case 8: FE_SP(8); break;
      case 16: FE_SP(16); break;
      default: FE_SP(32); break;
    }
#undef FE_SP
  }

  // z = M^-1 v
  int precond(const double *v, double *out)
  {
    if(pc == FE_PC_JACOBI) {
      hipLaunchKernelGGL(scaleVec, dim3(kG), dim3(kTB), 0, ctx->stream, n, dinvJ, v, out);
      return FE_STATUS_OK;
    }
    if(pc == FE_PC_NONE) {
      FE_HIP_TRY(ctx, hipMemcpyAsync(out, v, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
      return FE_STATUS_OK;
    }
    if(pc == FE_PC_FIELDSPLIT) return feFieldSplitApply(ctx, v, out);
    FePrecond &P = ctx->pc;
    if(sweeps == 0) {
      // exact: y = L^-1 v (G_L = I - L_strict: y_i = v_i + sum_j<i g_ij y_j), then
      // z = U^-1 y (z_i = y_i / u_ii + sum_j>i g_ij z_j with g_ij = -u_ij / u_ii)
      if(P.triL.empty()) P.triL = triPlan(P.levelPtr);
      if(P.triU.empty()) P.triU = triPlan(P.levelPtrU);
      auto solve = [&](const std::vector<FePrecond::TriSeg> &plan, const std::vector<int> &lp, const int *dLp,
                       const int *rows, const int *tia, const int *tja, const double *tg, const double *scale,
                       const double *in0, double *o) {
        for(const FePrecond::TriSeg &sg : plan) {
          if(sg.run) {
            hipLaunchKernelGGL(triLevelRun, dim3(1), dim3(kTriRunTB), 0, ctx->stream, dLp, sg.l0, sg.l1, rows, tia,
                               tja, tg, scale, in0, o);
          } else {
            const int b = lp[sg.l0], cnt = lp[sg.l0 + 1] - b;
            hipLaunchKernelGGL(triLevel, dim3((cnt * kTriLPR + 255) / 256), dim3(256), 0, ctx->stream, rows + b, cnt,
                               tia, tja, tg, scale, in0, o);
          }
        }
      };
      solve(P.triL, P.levelPtr, P.d_levelPtr, P.d_levelRows, P.d_iaL, P.d_jaL, P.d_gL64, nullptr, v, t1);
      solve(P.triU, P.levelPtrU, P.d_levelPtrU, P.d_levelRowsU, P.d_iaU, P.d_jaU, P.d_gU64, P.d_dinv64, t1, out);
      return FE_STATUS_OK;
    }
    // L sweeps: y_{k+1} = v + G_L,strict y_k (G_L = -L_strict), y_0 = v
    const double *cur = v;
    double *bufs[2] = {t1, t2};
    int bi = 0;
    for(int s = 0; s < sweeps; ++s) {
      hipLaunchKernelGGL(factorSweep, dim3(kG), dim3(kTB), 0, ctx->stream, n, P.d_iaL, P.d_jaL, P.d_gL64,
                         (const double *)nullptr, v, cur, bufs[bi]);
      cur = bufs[bi];
      bi ^= 1;
    }
    // keep y in t_y (copy if it aliases the U-sweep buffers)
    const double *y = cur;
    // U sweeps: z_{k+1} = D^-1 y + G_U,strict z_k (G_U = -D^-1 U_strict), z_0 = D^-1 y
    double *zb[2] = {out, (y == t1) ? t2 : t1};
    hipLaunchKernelGGL(scaleVec, dim3(kG), dim3(kTB), 0, ctx->stream, n, P.d_dinv64, y, zb[0]);
    int zi = 0;
    for(int s = 0; s < sweeps; ++s) {
      hipLaunchKernelGGL(factorSweep, dim3(kG), dim3(kTB), 0, ctx->stream, n, P.d_iaU, P.d_jaU, P.d_gU64,
                         P.d_dinv64, y, zb[zi], zb[zi ^ 1]);
      zi ^= 1;
    }
    if(zb[zi] != out)
      FE_HIP_TRY(ctx, hipMemcpyAsync(out, zb[zi], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    return FE_STATUS_OK;
  }

  // profiling: an event pair around a launch (timed: the A products of the
  // Arnoldi steps, the dominant kernel, and whole steps)
  int mark(FeKernelTimer &t)
  {
    if(t.used + 1 > t.ev.size())
      for(int q = 0; q < 64; ++q) {
        hipEvent_t e;
        FE_HIP_TRY(ctx, hipEventCreate(&e));
        t.ev.push_back(e);
      }
    FE_HIP_TRY(ctx, hipEventRecord(t.ev[t.used++], ctx->stream));
    return FE_STATUS_OK;
  }
  // complete pairs into the totals (the stream has been synchronized); an
  // open pair's first event moves to the front
  int collect(FeKernelTimer &t)
  {
    size_t q = 0;
    for(; q + 1 < t.used; q += 2) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t.ev[q], t.ev[q + 1]));
      t.total_ms += ms;
      t.launches++;
    }
    if(q < t.used) std::swap(t.ev[0], t.ev[q]);
    t.used -= q;
    return FE_STATUS_OK;
  }

  // host copies of the first nv folded values
  int dots(int nv, const double *Vb, const double *w_, double *hHost, bool timed = false)
  {
    if(timed && ctx->profiling)
      if(int rc = mark(ctx->spmvTimer)) return rc;
    hipLaunchKernelGGL(multiDot, dim3(kG), dim3(kTB), 0, ctx->stream, n, nv, Vb, ld, w_, part);
    if(timed && ctx->profiling)
      if(int rc = mark(ctx->spmvTimer)) return rc;
    hipLaunchKernelGGL(foldMany, dim3(nv), dim3(kTB), 0, ctx->stream, part, kG, h);
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hHost, h, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(ctx->profiling) {
      if(int rc = collect(ctx->spmvTimer)) return rc;
      if(int rc = collect(ctx->iterTimer)) return rc;
    }
    return FE_STATUS_OK;
  }
};

}  // namespace

int feSolveGmres(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double *du, fe_hip_solve_stats *stats)
{
  FE_REQUIRE(ctx, o->precision == FE_PREC_FP64, "GMRES runs in FP64");
  FE_REQUIRE(ctx,
             o->pc == FE_PC_NONE || o->pc == FE_PC_JACOBI || o->pc == FE_PC_ILU0 || o->pc == FE_PC_BJ_ILU0 ||
                 o->pc == FE_PC_FIELDSPLIT,
             "GMRES takes the NONE, JACOBI, ILU0 or FIELDSPLIT preconditioner");
  const int n = ctx->n;
  const int m = std::max(2, std::min(o->restart, 120));
  Gm g;
  g.ctx = ctx;
  g.n = n;
  g.m = m;
  g.ld = ((size_t)n + 63) & ~(size_t)63;
  g.pc = (o->pc == FE_PC_BJ_ILU0) ? FE_PC_ILU0 : o->pc;
  g.sweeps = std::max(0, ctx->pc.sweeps);
  double *work = nullptr;
  const size_t need = g.ld * (m + 1) + 7 * g.ld + (size_t)kG * 128 + 256;
  if(feAlloc(ctx, &work, need)) return FE_STATUS_OUT_OF_MEMORY;
  g.V = work;
  g.w = g.V + g.ld * (m + 1);
  g.z = g.w + g.ld;
  g.t1 = g.z + g.ld;
  g.t2 = g.t1 + g.ld;
  g.x = g.t2 + g.ld;
  g.r = g.x + g.ld;
  g.dinvJ = g.r + g.ld;
  g.part = g.dinvJ + g.ld;
  g.h = g.part + (size_t)kG * 128;

  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  ctx->spmvTimer = FeKernelTimer{std::move(ctx->spmvTimer.ev)};
  ctx->iterTimer = FeKernelTimer{std::move(ctx->iterTimer.ev)};
  double setupMs = 0.;
  if(g.pc == FE_PC_ILU0) {
    if(int rc = fePrecondFactor(ctx)) { feFree(ctx, work); return rc; }
  } else if(g.pc == FE_PC_FIELDSPLIT) {
    if(int rc = feFieldSplitSetup(ctx)) { feFree(ctx, work); return rc; }
  } else if(g.pc == FE_PC_JACOBI) {
    hipLaunchKernelGGL(jacobiDiag, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_diag, ctx->d_val, g.dinvJ);
  }
  {
    float sm = 0.f;
    FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
    FE_HIP_TRY(ctx, hipEventSynchronize(t1));
    FE_HIP_TRY(ctx, hipEventElapsedTime(&sm, t0, t1));
    setupMs = sm;
  }

  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), hh(m + 2), hh2(m + 2);
  FE_HIP_TRY(ctx, hipMemsetAsync(g.x, 0, sizeof(double) * n, ctx->stream));
  double bn = 0.;
  g.dots(1, ctx->d_rhs, ctx->d_rhs, &bn);
  bn = std::sqrt(bn);
  const double target = std::max(o->rtol * bn, o->atol);
  int iters = 0, outer = 0, converged = 0, reason = 0;
  double res = bn;
  if(bn == 0.) {
    converged = 1;
    reason = 2;
  }
  while(!converged && iters < o->max_iter) {
    // r = b - A x
    g.spmv(g.x, g.w);
    hipLaunchKernelGGL(residualKernel, dim3(kG), dim3(kTB), 0, ctx->stream, n, ctx->d_rhs, g.w, g.r);
    double beta = 0.;
    if(int rc = g.dots(1, g.r, g.r, &beta)) return rc;
    beta = std::sqrt(beta);
    res = beta;
    outer++;
    if(beta <= target) { converged = 1; reason = 2; break; }
    if(!(beta == beta) || beta / bn > o->dtol) { reason = -4; break; }
    hipLaunchKernelGGL(scaleInto, dim3(kG), dim3(kTB), 0, ctx->stream, n, 1. / beta, g.r, g.V);
    std::fill(gv.begin(), gv.end(), 0.);
    gv[0] = beta;
    int j = 0;
    bool cycleDone = false;
    for(; j < m && iters < o->max_iter; ++j) {
      double *vj = g.V + g.ld * j;
      if(ctx->profiling)
        if(int rc = g.mark(ctx->iterTimer)) return rc;
      if(int rc = g.precond(vj, g.z)) { feFree(ctx, work); return rc; }
      // profiling: events around the A product (the dominant kernel of the
      // Arnoldi steps), collected at the next host read
      if(ctx->profiling)
        if(int rc = g.mark(ctx->spmvTimer)) return rc;
      g.spmv(g.z, g.w);
      if(ctx->profiling)
        if(int rc = g.mark(ctx->spmvTimer)) return rc;
      // classical Gram-Schmidt, twice (CGS2) when gmres_orth = 2
      if(int rc = g.dots(j + 1, g.V, g.w, hh.data())) return rc;
      hipLaunchKernelGGL(multiAxpy, dim3(kG), dim3(kTB), 0, ctx->stream, n, j + 1, g.V, g.ld, g.h, g.w);
      if(ctx->gmresOrth == 2) {
        if(int rc = g.dots(j + 1, g.V, g.w, hh2.data())) return rc;
        hipLaunchKernelGGL(multiAxpy, dim3(kG), dim3(kTB), 0, ctx->stream, n, j + 1, g.V, g.ld, g.h, g.w);
      } else {
        std::fill(hh2.begin(), hh2.begin() + j + 1, 0.);
      }
      double wn = 0.;
      if(int rc = g.dots(1, g.w, g.w, &wn)) return rc;
      wn = std::sqrt(wn);
      for(int k = 0; k <= j; ++k) H[(size_t)k * m + j] = hh[k] + hh2[k];
      H[(size_t)(j + 1) * m + j] = wn;
      if(wn > 0.)
        hipLaunchKernelGGL(scaleInto, dim3(kG), dim3(kTB), 0, ctx->stream, n, 1. / wn, g.w, g.V + g.ld * (j + 1));
      if(ctx->profiling)
        if(int rc = g.mark(ctx->iterTimer)) return rc;
      for(int k = 0; k < j; ++k) {
        const double a = H[(size_t)k * m + j], b = H[(size_t)(k + 1) * m + j];
        H[(size_t)k * m + j] = cs[k] * a + sn[k] * b;
        H[(size_t)(k + 1) * m + j] = -sn[k] * a + cs[k] * b;
      }
      const double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
      const double rr = std::sqrt(a * a + b * b);
      cs[j] = rr > 0. ? a / rr : 1.;
      sn[j] = rr > 0. ? b / rr : 0.;
      H[(size_t)j * m + j] = rr;
      H[(size_t)(j + 1) * m + j] = 0.;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      iters++;
      if(std::fabs(gv[j + 1]) <= target || wn == 0.) {
        ++j;
        cycleDone = true;
        break;
      }
    }
    (void)cycleDone;
    // y = H^-1 g, x += M^-1 (V y)
    for(int k = j - 1; k >= 0; --k) {
      double s = gv[k];
      for(int l = k + 1; l < j; ++l) s -= H[(size_t)k * m + l] * y[l];
      y[k] = s / H[(size_t)k * m + k];
    }
    FE_HIP_TRY(ctx, hipMemcpyAsync(g.h, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(multiCombine, dim3(kG), dim3(kTB), 0, ctx->stream, n, j, g.V, g.ld, g.h, g.w);
    if(int rc = g.precond(g.w, g.z)) { feFree(ctx, work); return rc; }
    hipLaunchKernelGGL(axpyKernel, dim3(kG), dim3(kTB), 0, ctx->stream, n, 1., g.z, g.x);
    FE_HIP_TRY(ctx, hipGetLastError());
  }
  if(!converged) {
    g.spmv(g.x, g.w);
    hipLaunchKernelGGL(residualKernel, dim3(kG), dim3(kTB), 0, ctx->stream, n, ctx->d_rhs, g.w, g.r);
    double rr = 0.;
    if(int rc = g.dots(1, g.r, g.r, &rr)) return rc;
    res = std::sqrt(rr);
    if(res <= target) { converged = 1; reason = 2; }
    else if(reason == 0) reason = -3;
  }
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, g.x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if(ctx->profiling) {
    if(int rc = g.collect(ctx->spmvTimer)) return rc;
    if(int rc = g.collect(ctx->iterTimer)) return rc;
    ctx->spmvTimer.used = ctx->iterTimer.used = 0;
  }
  if(stats) {
    memset(stats, 0, sizeof(*stats));
    // infinity norms: du, b, A du - b
    g.spmv(g.x, g.w);
    std::vector<double> hx(n), hb(n), hw(n);
    FE_HIP_TRY(ctx, hipMemcpyAsync(hx.data(), g.x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(hb.data(), ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(hw.data(), g.w, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    double a = 0., b = 0., c = 0.;
    for(int i = 0; i < n; ++i) {
      a = std::max(a, std::fabs(hx[i]));
      b = std::max(b, std::fabs(hb[i]));
      c = std::max(c, std::fabs(hw[i] - hb[i]));
    }
    stats->iterations = iters;
    stats->outer_iterations = outer;
    stats->converged = converged;
    stats->reason = reason;
    stats->relres_fp64 = bn > 0. ? res / bn : 0.;
    stats->norm_dx_inf = a;
    stats->norm_rhs_inf = b;
    stats->norm_axb_inf = c;
    stats->ms_solve = ms;
    stats->ms_setup = setupMs;
  }
  if(du) FE_HIP_TRY(ctx, hipMemcpy(du, g.x, sizeof(double) * n, hipMemcpyDeviceToHost));
  feFree(ctx, work);
  return FE_STATUS_OK;
}
