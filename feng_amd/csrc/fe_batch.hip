// Batched independent solves: many diffusion problems assembled into one
// block-diagonal system (problem p owns the contiguous unknown rows
// [rows[p], rows[p+1]) and couples to no other problem), solved by
// per-problem pipelined Jacobi-PCG in single launches over all problems.
//
// Each block owns one row tile that lies inside one problem; the partial
// sums of iteration i are stored per tile, and a block folds the partials of
// its own problem's tiles only, so every problem has its own alpha, beta,
// stopping test and iteration count.  Tiles of a problem that has stopped
// return at once.  `kBatchGraphIters` iterations are captured in a HIP graph
// whose kernels read the iteration base from device memory (advanced by the
// last node), so the host only relaunches the graph and polls the stop
// flags.  Outer FP64 iterative refinement per problem: r = b - A x in FP64,
// inner solve of A d = r in FP32 (values from the FP32 assembly) to
// inner_rtol ||r_p||, x += d; a problem whose FP64 residual an FP32
// correction does not reduce gets that correction undone, and the remaining
// corrections of the batch are solved in FP64.
#include "fe_hip_internal.h"
#include "fe_tile.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace fe_tile;

struct FeBatch {
  int nprob = 0, TB = 0, nTiles = 0;
  std::vector<int> rows;  // nprob + 1
  int *tRow0 = nullptr, *tRows = nullptr, *tProb = nullptr, *pTile = nullptr;
  int *done = nullptr, *iter = nullptr, *outerDone = nullptr, *itBase = nullptr;
  double *gamma = nullptr, *alpha = nullptr, *tol2 = nullptr, *rr64 = nullptr, *part = nullptr;
  double *pg[2] = {}, *pd[2] = {}, *prho[2] = {};
  double *x64 = nullptr, *r64 = nullptr, *bak64 = nullptr;
  TileVecs<double> v64;
  TileVecs<float> v32;
  double *sp[9] = {};  // per-problem AMG-PCG vectors x, r, u, w, z, q, s, p, m (allocated on first use)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graphKey[5] = {-1, -1, -1, -1, -1};  // TB, lanes, precision, maxit, active tiles of the captured graph
  // the per-problem multigrid PCG: kBatchGraphIters iterations (V-cycle +
  // update each) and the iteration-base advance captured as one graph
  hipGraph_t ampGraph = nullptr;
  hipGraphExec_t ampExec = nullptr;
  int ampKey[4] = {-1, -1, -1, -1};  // active tiles, iteration budget, block, FP32 level values
  size_t cap = 0;  // allocated tiles
  std::vector<int> hPtile;  // host copy of the per-problem tile ranges
  int *act = nullptr;       // tiles of the problems still iterating (the iteration grid)
  int nAct = 0;
};

void feBatchRelease(fe_hip_ctx *ctx)
{
  if(!ctx->batch) return;
  if(ctx->batch->exec) (void)hipGraphExecDestroy(ctx->batch->exec);
  if(ctx->batch->graph) (void)hipGraphDestroy(ctx->batch->graph);
  if(ctx->batch->ampExec) (void)hipGraphExecDestroy(ctx->batch->ampExec);
  if(ctx->batch->ampGraph) (void)hipGraphDestroy(ctx->batch->ampGraph);
  delete ctx->batch;
  ctx->batch = nullptr;
}

namespace {

constexpr int kBatchGraphIters = 16;

struct BTiles {
  const int *row0, *rows, *prob, *ptile;
};

struct BState {
  double *gamma, *alpha;  // [nprob][2] by parity
  const double *tol2;
  int *done, *iter;
  const int *outerDone, *itBase;
};

struct BParts {
  double *g[2], *d[2], *rho[2];
};

// Sum of three values over the block; every thread gets the totals.
template <int TB>
__device__ __forceinline__ void blockSum3(double &a, double &b, double &c)
{
  __shared__ double lds[3][TB / 64];
  a = feWaveSum(a);
  b = feWaveSum(b);
  c = feWaveSum(c);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if(lane == 0) {
    lds[0][wid] = a;
    lds[1][wid] = b;
    lds[2][wid] = c;
  }
  __syncthreads();
  a = b = c = 0.;
#pragma unroll
  for(int i = 0; i < TB / 64; ++i) {
    a += lds[0][i];
    b += lds[1][i];
    c += lds[2][i];
  }
  __syncthreads();
}

// Start of an inner solve: r = b (FP64 residual of the outer iteration), x = 0,
// w = A D^-1 r, m = D^-1 w, s = z = p = 0; partials of iteration 0.  Problems
// already converged only clear x and stop.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void bInit(const int *__restrict__ ia, const int *__restrict__ ja,
                                            const T *__restrict__ val, const double *__restrict__ b, TileVecs<T> v,
                                            BTiles bt, BParts P, BState S)
{
  __shared__ T ldsY[TB];
  const int tile = blockIdx.x, p = bt.prob[tile];
  const int base = bt.row0[tile], nEnd = base + bt.rows[tile];
  const int row = base + threadIdx.x;
  if(S.outerDone[p]) {
    if(row < nEnd) v.x[row] = (T)0;
    if(threadIdx.x == 0) {
      P.g[0][tile] = P.d[0][tile] = P.rho[0][tile] = 0.;
      if(tile == bt.ptile[p]) {
        S.done[p] = 1;
        S.iter[p] = 0;
      }
    }
    return;
  }
  int kb[LPR], ke[LPR];
  tileBounds<LPR, TB>(nEnd, base, ia, kb, ke);
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR, gg = threadIdx.x / LPR;
#pragma unroll
  for(int u = 0; u < LPR; ++u) {
    T acc = feRowDot<FeKs<LPR>::value, LPR>(ja, val, kb[u], ke[u], [&](int j) { return v.dinv[j] * (T)b[j]; });
#pragma unroll
    for(int q = LPR / 2; q > 0; q >>= 1) acc += __shfl_xor(acc, q, LPR);
    if(sub == u) ldsY[u * RPB + gg] = acc;
  }
  __syncthreads();
  double g = 0., d = 0., rho = 0.;
  if(row < nEnd) {
    const T di = v.dinv[row];
    const T ri = (T)b[row];
    const T wi = ldsY[threadIdx.x];
    const T ui = di * ri;
    v.r[row] = ri;
    v.x[row] = (T)0;
    v.w[row] = wi;
    v.m[0][row] = di * wi;
    v.s[row] = v.z[row] = v.p[row] = (T)0;
    g = (double)ri * (double)ui;
    d = (double)wi * (double)ui;
    rho = (double)ri * (double)ri;
  }
  blockSum3<TB>(g, d, rho);
  if(threadIdx.x == 0) {
    P.g[0][tile] = g;
    P.d[0][tile] = d;
    P.rho[0][tile] = rho;
    if(tile == bt.ptile[p]) {
      S.gamma[2 * p + 1] = 1.;
      S.alpha[2 * p + 1] = __builtin_inf();
      S.done[p] = 0;
      S.iter[p] = 0;
    }
  }
}

// Iteration it = itBase + k of every problem that has not stopped.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void bIter(const int *__restrict__ ia, const int *__restrict__ ja,
                                            const T *__restrict__ val, TileVecs<T> v, BTiles bt, BParts P, BState S,
                                            const int *__restrict__ act, int k, int maxit)
{
  __shared__ T ldsY[TB];
  const int tile = act[blockIdx.x], p = bt.prob[tile];
  if(S.done[p]) return;
  const int it = S.itBase[0] + k;
  const int base = bt.row0[tile], nEnd = base + bt.rows[tile];
  int kb[LPR], ke[LPR];
  TileRow<T> me;
  tileBounds<LPR, TB>(nEnd, base, ia, kb, ke);
  tileLoadRow(v, base + threadIdx.x, nEnd, me);
  const int c = it & 1, o = c ^ 1;
  const double gPrev = S.gamma[2 * p + o], aPrev = S.alpha[2 * p + o];
  const int t0 = bt.ptile[p], t1 = bt.ptile[p + 1];
  double gamma = 0., delta = 0., rho = 0.;
  for(int q = t0 + threadIdx.x; q < t1; q += TB) {
    gamma += P.g[c][q];
    delta += P.d[c][q];
    rho += P.rho[c][q];
  }
  blockSum3<TB>(gamma, delta, rho);
  const bool lead = (tile == t0 && threadIdx.x == 0);
  if(!(rho > S.tol2[p]) || it >= maxit) {
    if(lead) {
      S.iter[p] = it;
      S.done[p] = (rho <= S.tol2[p]) ? 1 : (rho == rho ? 2 : 4);
    }
    return;
  }
  const double betaD = gamma / gPrev;
  const double den = delta - betaD * gamma / aPrev;
  if(!(den > 0.)) {
    if(lead) {
      S.iter[p] = it;
      S.done[p] = 3;
    }
    return;
  }
  const double alphaD = gamma / den;
  if(lead) {
    S.gamma[2 * p + c] = gamma;
    S.alpha[2 * p + c] = alphaD;
  }
  const T alpha = (T)alphaD, beta = (T)betaD;
  tileProduct<T, LPR, TB, false>(nEnd, base, ia, ja, val, v.m[c], nullptr, kb, ke, ldsY);
  __syncthreads();
  double g = 0., d = 0., rh = 0.;
  const int row = base + threadIdx.x;
  if(row < nEnd) {
    const T y = ldsY[threadIdx.x];
    const T z = y + beta * me.z;
    const T s = me.w + beta * me.s;
    const T pp = me.d * me.r + beta * me.p;
    const T rn = me.r - alpha * s;
    const T wn = me.w - alpha * z;
    const T un = me.d * rn;
    v.x[row] = me.x + alpha * pp;
    v.p[row] = pp;
    v.s[row] = s;
    v.z[row] = z;
    v.r[row] = rn;
    v.w[row] = wn;
    v.m[o][row] = me.d * wn;
    g = (double)rn * (double)un;
    d = (double)wn * (double)un;
    rh = (double)rn * (double)rn;
  }
  blockSum3<TB>(g, d, rh);
  if(threadIdx.x == 0) {
    P.g[o][tile] = g;
    P.d[o][tile] = d;
    P.rho[o][tile] = rh;
  }
}

__global__ void bAdvance(int *itBase, int by)
{
  if(threadIdx.x == 0 && blockIdx.x == 0) itBase[0] += by;
}

// r = b - A x (FP64), per-tile partials of ||r||^2 (x may be null: r = b).
template <int TB>
__global__ __launch_bounds__(TB) void bResidual64(const int *__restrict__ ia, const int *__restrict__ ja,
                                                  const double *__restrict__ val, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ r, BTiles bt,
                                                  double *__restrict__ part)
{
  const int tile = blockIdx.x;
  const int row = bt.row0[tile] + threadIdx.x;
  double rr = 0.;
  if(threadIdx.x < bt.rows[tile]) {
    double s = b[row];
    if(x)
      for(int k = ia[row]; k < ia[row + 1]; ++k) s -= val[k] * x[ja[k]];
    r[row] = s;
    rr = s * s;
  }
  double z1 = 0., z2 = 0.;
  blockSum3<TB>(rr, z1, z2);
  if(threadIdx.x == 0) part[tile] = rr;
}

// Per-problem sums of the tile partials.
__global__ void bFoldProblems(int nprob, const int *__restrict__ ptile, const double *__restrict__ part,
                              double *__restrict__ out)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= nprob) return;
  double s = 0.;
  for(int t = ptile[p]; t < ptile[p + 1]; ++t) s += part[t];
  out[p] = s;
}

template <typename T>
__global__ void bAddCorrection(int n, double *__restrict__ x, const T *__restrict__ d)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += (double)d[i];
}

template <typename T>
__global__ void bDinv(int n, const int *__restrict__ diag, const double *__restrict__ val, T *__restrict__ dinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double dd = val[diag[i]];
  dinv[i] = (T)(dd != 0. ? 1. / dd : 1.);
}

// ---------------------------------------------------------------------------
// Per-problem pipelined PCG with an external preconditioner (the union AMG
// V-cycle, block diagonal, so problem p sees its own V-cycle): the split form
// of the single-system solver (Ghysels & Vanroose, Alg. 3) with the scalars,
// stopping test and iteration count of each problem, u = M r and m = M w
// given as vectors.

struct BSpVecs {
  double *x, *r, *u, *w, *z, *q, *s, *p, *m;
};

// r = b (the outer FP64 residual), x = 0; rows of problems that have stopped
// get r = 0.
__global__ void bSpStart(const double *__restrict__ b, BSpVecs v, BTiles bt, BState S)
{
  const int tile = blockIdx.x, p = bt.prob[tile];
  const int row = bt.row0[tile] + threadIdx.x;
  if(threadIdx.x < bt.rows[tile]) {
    v.x[row] = 0.;
    v.r[row] = S.outerDone[p] ? 0. : b[row];
  }
  // the stop flags of this pass's first V-cycle (which skips the tiles of
  // stopped problems); bSpInit sets them again
  if(threadIdx.x == 0 && tile == bt.ptile[p]) S.done[p] = S.outerDone[p];
}

// w = A u, z = q = s = p = 0, partials of iteration 0, per-problem start values.
template <int LPR, int TB>
__global__ __launch_bounds__(TB) void bSpInit(const int *__restrict__ ia, const int *__restrict__ ja,
                                              const double *__restrict__ val, BSpVecs v, BTiles bt, BParts P,
                                              BState S)
{
  __shared__ double ldsY[TB];
  const int tile = blockIdx.x, p = bt.prob[tile];
  const int base = bt.row0[tile], nEnd = base + bt.rows[tile];
  const int row = base + threadIdx.x;
  if(S.outerDone[p]) {
    if(threadIdx.x == 0) {
      P.g[0][tile] = P.d[0][tile] = P.rho[0][tile] = 0.;
      if(tile == bt.ptile[p]) {
        S.done[p] = 1;
        S.iter[p] = 0;
      }
    }
    return;
  }
  int kb[LPR], ke[LPR];
  tileBounds<LPR, TB>(nEnd, base, ia, kb, ke);
  tileProduct<double, LPR, TB, false>(nEnd, base, ia, ja, val, v.u, nullptr, kb, ke, ldsY);
  __syncthreads();
  double g = 0., d = 0., rho = 0.;
  if(row < nEnd) {
    const double wi = ldsY[threadIdx.x], ri = v.r[row], ui = v.u[row];
    v.w[row] = wi;
    v.z[row] = v.q[row] = v.s[row] = v.p[row] = 0.;
    g = ri * ui;
    d = wi * ui;
    rho = ri * ri;
  }
  blockSum3<TB>(g, d, rho);
  if(threadIdx.x == 0) {
    P.g[0][tile] = g;
    P.d[0][tile] = d;
    P.rho[0][tile] = rho;
    if(tile == bt.ptile[p]) {
      S.gamma[2 * p + 1] = 1.;
      S.alpha[2 * p + 1] = __builtin_inf();
      S.done[p] = 0;
      S.iter[p] = 0;
    }
  }
}

// Iteration `it` of every problem still iterating (active tiles), m = M w.
template <int LPR, int TB>
__global__ __launch_bounds__(TB) void bSpIter(const int *__restrict__ ia, const int *__restrict__ ja,
                                              const double *__restrict__ val, BSpVecs v, BTiles bt, BParts P,
                                              BState S, const int *__restrict__ act,
                                              const int *__restrict__ itBase, int k, int maxit)
{
  __shared__ double ldsY[TB];
  const int it = *itBase + k;  // the iteration: a device base (advanced by bAdvance) + the offset in the chunk
  const int tile = act[blockIdx.x], p = bt.prob[tile];
  if(S.done[p]) return;
  const int base = bt.row0[tile], nEnd = base + bt.rows[tile];
  const int row = base + threadIdx.x;
  int kb[LPR], ke[LPR];
  tileBounds<LPR, TB>(nEnd, base, ia, kb, ke);
  double r = 0., u = 0., w = 0., z = 0., q = 0., s = 0., pv = 0., x = 0., m = 0.;
  if(row < nEnd) {
    r = v.r[row]; u = v.u[row]; w = v.w[row]; z = v.z[row]; q = v.q[row];
    s = v.s[row]; pv = v.p[row]; x = v.x[row]; m = v.m[row];
  }
  const int c = it & 1, o = c ^ 1;
  const double gPrev = S.gamma[2 * p + o], aPrev = S.alpha[2 * p + o];
  const int t0 = bt.ptile[p], t1 = bt.ptile[p + 1];
  double gamma = 0., delta = 0., rho = 0.;
  for(int k = t0 + threadIdx.x; k < t1; k += TB) {
    gamma += P.g[c][k];
    delta += P.d[c][k];
    rho += P.rho[c][k];
  }
  blockSum3<TB>(gamma, delta, rho);
  const bool lead = (tile == t0 && threadIdx.x == 0);
  if(!(rho > S.tol2[p]) || it >= maxit) {
    if(lead) {
      S.iter[p] = it;
      S.done[p] = (rho <= S.tol2[p]) ? 1 : (rho == rho ? 2 : 4);
    }
    return;
  }
  const double beta = gamma / gPrev;
  const double den = delta - beta * gamma / aPrev;
  if(!(den > 0.)) {
    if(lead) {
      S.iter[p] = it;
      S.done[p] = 3;
    }
    return;
  }
  const double alpha = gamma / den;
  if(lead) {
    S.gamma[2 * p + c] = gamma;
    S.alpha[2 * p + c] = alpha;
  }
  tileProduct<double, LPR, TB, false>(nEnd, base, ia, ja, val, v.m, nullptr, kb, ke, ldsY);
  __syncthreads();
  double g = 0., d = 0., rh = 0.;
  if(row < nEnd) {
    z = ldsY[threadIdx.x] + beta * z;
    q = m + beta * q;
    s = w + beta * s;
    pv = u + beta * pv;
    x += alpha * pv;
    r -= alpha * s;
    u -= alpha * q;
    w -= alpha * z;
    v.x[row] = x; v.r[row] = r; v.u[row] = u; v.w[row] = w;
    v.z[row] = z; v.q[row] = q; v.s[row] = s; v.p[row] = pv;
    g = r * u;
    d = w * u;
    rh = r * r;
  }
  blockSum3<TB>(g, d, rh);
  if(threadIdx.x == 0) {
    P.g[o][tile] = g;
    P.d[o][tile] = d;
    P.rho[o][tile] = rh;
  }
}

inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

// Tiles of TB rows that never straddle two problems.
int buildTiles(fe_hip_ctx *ctx, FeBatch &B, int TB)
{
  if(B.TB == TB && B.tRow0) return FE_STATUS_OK;
  std::vector<int> row0, rows, prob, ptile(B.nprob + 1, 0);
  for(int p = 0; p < B.nprob; ++p) {
    ptile[p] = (int)row0.size();
    for(int r = B.rows[p]; r < B.rows[p + 1]; r += TB) {
      row0.push_back(r);
      rows.push_back(std::min(TB, B.rows[p + 1] - r));
      prob.push_back(p);
    }
  }
  ptile[B.nprob] = (int)row0.size();
  const size_t nt = row0.size();
  if(nt > B.cap) {
    for(void *q : {(void *)B.tRow0, (void *)B.tRows, (void *)B.tProb, (void *)B.part})
      if(q) feFree(ctx, q);
    for(int k = 0; k < 2; ++k)
      for(void *q : {(void *)B.pg[k], (void *)B.pd[k], (void *)B.prho[k]})
        if(q) feFree(ctx, q);
    if(B.act) feFree(ctx, B.act);
    if(feAlloc(ctx, &B.tRow0, nt) || feAlloc(ctx, &B.tRows, nt) || feAlloc(ctx, &B.tProb, nt) ||
       feAlloc(ctx, &B.part, nt) || feAlloc(ctx, &B.act, nt))
      return FE_STATUS_OUT_OF_MEMORY;
    for(int k = 0; k < 2; ++k)
      if(feAlloc(ctx, &B.pg[k], nt) || feAlloc(ctx, &B.pd[k], nt) || feAlloc(ctx, &B.prho[k], nt))
        return FE_STATUS_OUT_OF_MEMORY;
    B.cap = nt;
  }
  FE_HIP_TRY(ctx, hipMemcpy(B.tRow0, row0.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(B.tRows, rows.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(B.tProb, prob.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(B.pTile, ptile.data(), sizeof(int) * (B.nprob + 1), hipMemcpyHostToDevice));
  B.nTiles = (int)nt;
  B.TB = TB;
  B.hPtile = ptile;
  B.nAct = 0;
  B.graphKey[0] = -1;  // tiles changed: recapture
  return FE_STATUS_OK;
}

int setupBatch(fe_hip_ctx *ctx, int nprob, const int *rows, FeBatch *&out)
{
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR pattern first");
  FE_REQUIRE(ctx, nprob >= 1 && rows && rows[0] == 0 && rows[nprob] == ctx->n,
             "problem row ranges must start at 0 and end at the number of unknowns");
  for(int p = 0; p < nprob; ++p) FE_REQUIRE(ctx, rows[p + 1] > rows[p], "every problem needs at least one unknown");
  if(ctx->batch && ctx->batch->nprob == nprob && std::equal(rows, rows + nprob + 1, ctx->batch->rows.begin())) {
    out = ctx->batch;
    return FE_STATUS_OK;
  }
  feBatchRelease(ctx);
  ctx->batch = new FeBatch();
  FeBatch &B = *ctx->batch;
  B.nprob = nprob;
  B.rows.assign(rows, rows + nprob + 1);
  const size_t n = ctx->n;
  if(feAlloc(ctx, &B.pTile, nprob + 1) || feAlloc(ctx, &B.done, nprob) || feAlloc(ctx, &B.iter, nprob) ||
     feAlloc(ctx, &B.outerDone, nprob) || feAlloc(ctx, &B.itBase, 1) || feAlloc(ctx, &B.gamma, 2 * nprob) ||
     feAlloc(ctx, &B.alpha, 2 * nprob) || feAlloc(ctx, &B.tol2, nprob) || feAlloc(ctx, &B.rr64, nprob) ||
     feAlloc(ctx, &B.x64, n) || feAlloc(ctx, &B.r64, n) || feAlloc(ctx, &B.bak64, n))
    return FE_STATUS_OUT_OF_MEMORY;
  double **a[] = {&B.v64.x, &B.v64.r, &B.v64.w, &B.v64.s, &B.v64.z, &B.v64.p, &B.v64.dinv, &B.v64.m[0], &B.v64.m[1]};
  for(auto pp : a)
    if(feAlloc(ctx, pp, n)) return FE_STATUS_OUT_OF_MEMORY;
  float **b[] = {&B.v32.x, &B.v32.r, &B.v32.w, &B.v32.s, &B.v32.z, &B.v32.p, &B.v32.dinv, &B.v32.m[0], &B.v32.m[1]};
  for(auto pp : b)
    if(feAlloc(ctx, pp, n)) return FE_STATUS_OUT_OF_MEMORY;
  out = ctx->batch;
  return FE_STATUS_OK;
}

BTiles tilesOf(const FeBatch &B) { return BTiles{B.tRow0, B.tRows, B.tProb, B.pTile}; }
BState stateOf(const FeBatch &B) { return BState{B.gamma, B.alpha, B.tol2, B.done, B.iter, B.outerDone, B.itBase}; }
BParts partsOf(const FeBatch &B) { return BParts{{B.pg[0], B.pg[1]}, {B.pd[0], B.pd[1]}, {B.prho[0], B.prho[1]}}; }

template <typename T, int LPR, int TB>
void launchInit(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v)
{
  hipLaunchKernelGGL((bInit<T, LPR, TB>), dim3(B.nTiles), dim3(TB), 0, ctx->stream, ctx->d_ia, ctx->d_ja, val, B.r64,
                     v, tilesOf(B), partsOf(B), stateOf(B));
}

template <typename T, int LPR, int TB>
void launchIter(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v, int k, int maxit)
{
  hipLaunchKernelGGL((bIter<T, LPR, TB>), dim3(B.nAct), dim3(TB), 0, ctx->stream, ctx->d_ia, ctx->d_ja, val, v,
                     tilesOf(B), partsOf(B), stateOf(B), (const int *)B.act, k, maxit);
}

// Dispatch over (lanes, block) of a templated launcher.
template <typename T, template <typename, int, int> class L, typename... Args>
int dispatch(fe_hip_ctx *ctx, int lanes, int TB, Args &&...args)
{
#define FE_B_CASE(LL, BB)                    \
  if(lanes == LL && TB == BB) {              \
    L<T, LL, BB>::run(ctx, args...);         \
    return FE_STATUS_OK;                     \
  }
  FE_B_CASE(2, 256) FE_B_CASE(4, 256) FE_B_CASE(8, 256)
  FE_B_CASE(2, 512) FE_B_CASE(4, 512) FE_B_CASE(8, 512)
  FE_B_CASE(2, 1024) FE_B_CASE(4, 1024) FE_B_CASE(8, 1024)
#undef FE_B_CASE
  FE_REQUIRE(ctx, false, "batched solver: lanes must be 2, 4 or 8 and the block 256, 512 or 1024");
}

template <typename T, int LPR, int TB>
struct InitL {
  static void run(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v) { launchInit<T, LPR, TB>(ctx, B, val, v); }
};
template <typename T, int LPR, int TB>
struct IterL {
  static void run(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v, int k, int maxit)
  {
    launchIter<T, LPR, TB>(ctx, B, val, v, k, maxit);
  }
};

// Captures kBatchGraphIters iterations + the base advance into a graph.
template <typename T>
int captureGraph(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v, int maxit, int prec)
{
  const int key[5] = {B.TB, ctx->spmvLanes, prec, maxit, B.nAct};
  if(B.exec && std::equal(key, key + 5, B.graphKey)) return FE_STATUS_OK;
  if(B.exec) (void)hipGraphExecDestroy(B.exec);
  if(B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
  FE_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = FE_STATUS_OK;
  for(int k = 0; k < kBatchGraphIters && rc == FE_STATUS_OK; ++k)
    rc = dispatch<T, IterL>(ctx, ctx->spmvLanes, B.TB, B, val, v, k, maxit);
  if(rc == FE_STATUS_OK) hipLaunchKernelGGL(bAdvance, dim3(1), dim3(64), 0, ctx->stream, B.itBase, kBatchGraphIters);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if(rc) return rc;
  FE_HIP_TRY(ctx, e);
  B.graph = g;
  FE_HIP_TRY(ctx, hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
  std::copy(key, key + 5, B.graphKey);
  return FE_STATUS_OK;
}

// Uploads the tiles of the problems with done[p] == 0 as the iteration grid.
int setActive(fe_hip_ctx *ctx, FeBatch &B, const std::vector<int> &done)
{
  std::vector<int> act;
  act.reserve(B.nTiles);
  for(int p = 0; p < B.nprob; ++p)
    if(!done[p])
      for(int t = B.hPtile[p]; t < B.hPtile[p + 1]; ++t) act.push_back(t);
  B.nAct = std::max(1, (int)act.size());
  if(act.empty()) act.push_back(0);  // a stopped problem's tile returns at once
  FE_HIP_TRY(ctx, hipMemcpyAsync(B.act, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, ctx->stream));
  return FE_STATUS_OK;
}

// Runs inner iterations until every problem has stopped (or maxit).  The
// iteration grid covers only the tiles of problems still iterating: when the
// active tiles fall below 3/4 of the captured grid, the list is rebuilt and
// the graph recaptured with the smaller grid.
template <typename T>
int runInner(fe_hip_ctx *ctx, FeBatch &B, const T *val, TileVecs<T> &v, int maxit, int prec, std::vector<int> &hDone,
             std::vector<int> &hIter, const std::vector<int> &outerDone)
{
  if(int rc = setActive(ctx, B, outerDone)) return rc;
  if(int rc = captureGraph<T>(ctx, B, val, v, maxit, prec)) return rc;
  FE_HIP_TRY(ctx, hipMemsetAsync(B.itBase, 0, sizeof(int), ctx->stream));
  if(int rc = dispatch<T, InitL>(ctx, ctx->spmvLanes, B.TB, B, val, v)) return rc;
  FE_HIP_TRY(ctx, hipGetLastError());
  const int maxLaunch = maxit / kBatchGraphIters + 2;
  for(int l = 0; l < maxLaunch; ++l) {
    FE_HIP_TRY(ctx, hipGraphLaunch(B.exec, ctx->stream));
    if(l % 2 == 1 || l + 1 == maxLaunch) {
      FE_HIP_TRY(ctx, hipMemcpyAsync(hDone.data(), B.done, sizeof(int) * B.nprob, hipMemcpyDeviceToHost, ctx->stream));
      FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      if(std::all_of(hDone.begin(), hDone.end(), [](int d) { return d != 0; })) break;
      int active = 0;
      for(int p = 0; p < B.nprob; ++p)
        if(!hDone[p]) active += B.hPtile[p + 1] - B.hPtile[p];
      if(active < (3 * B.nAct) / 4) {
        if(int rc = setActive(ctx, B, hDone)) return rc;
        if(int rc = captureGraph<T>(ctx, B, val, v, maxit, prec)) return rc;
      }
    }
  }
  FE_HIP_TRY(ctx, hipMemcpyAsync(hDone.data(), B.done, sizeof(int) * B.nprob, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipMemcpyAsync(hIter.data(), B.iter, sizeof(int) * B.nprob, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int residuals(fe_hip_ctx *ctx, FeBatch &B, const double *x, std::vector<double> &rn)
{
  const BTiles bt = tilesOf(B);
#define FE_RES(TBV) \
  hipLaunchKernelGGL(bResidual64<TBV>, dim3(B.nTiles), dim3(TBV), 0, ctx->stream, ctx->d_ia, ctx->d_ja, ctx->d_val, x, \
                     ctx->d_rhs, B.r64, bt, B.part)
  if(B.TB == 256) FE_RES(256);
  else if(B.TB == 512) FE_RES(512);
  else FE_RES(1024);
#undef FE_RES
  hipLaunchKernelGGL(bFoldProblems, dim3(cdiv(B.nprob, 256)), dim3(256), 0, ctx->stream, B.nprob, B.pTile, B.part,
                     B.rr64);
  FE_HIP_TRY(ctx, hipGetLastError());
  std::vector<double> rr(B.nprob);
  FE_HIP_TRY(ctx, hipMemcpyAsync(rr.data(), B.rr64, sizeof(double) * B.nprob, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  rn.resize(B.nprob);
  for(int p = 0; p < B.nprob; ++p) rn[p] = std::sqrt(rr[p]);
  return FE_STATUS_OK;
}

// AMG-PCG on the union system.  The system is block diagonal, so the
// aggregation never joins two problems and one V-cycle applies every
// problem's own V-cycle in shared launches.  The CG scalars are global: the
// solve runs until ||r|| <= min_p max(rtol ||b_p||, atol), which bounds every
// problem's own residual, and the per-problem FP64 residuals are then
// checked.
int solveBatchedAmg(fe_hip_ctx *ctx, FeBatch &B, const fe_hip_solver_opts *o, fe_hip_batch_stats *stats,
                    double *ms_solve)
{
  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  std::vector<double> bn, rn;
  if(int rc = residuals(ctx, B, nullptr, bn)) return rc;
  std::vector<double> target(B.nprob);
  double tmin = 1e300;
  for(int p = 0; p < B.nprob; ++p) {
    target[p] = std::max(o->rtol * bn[p], o->atol);
    if(bn[p] > 0.) tmin = std::min(tmin, target[p]);
  }
  fe_hip_solver_opts o2 = *o;
  o2.rtol = 0.;
  o2.atol = tmin < 1e300 ? 0.99 * tmin : o->atol;
  fe_hip_solve_stats st{};
  if(int rc = feSolve(ctx, &o2, nullptr, &st)) return rc;
  if(int rc = residuals(ctx, B, ctx->d_du, rn)) return rc;
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if(ms_solve) *ms_solve = ms;
  if(stats)
    for(int p = 0; p < B.nprob; ++p) {
      stats[p].iterations = st.iterations;
      stats[p].relres_fp64 = bn[p] > 0. ? rn[p] / bn[p] : 0.;
      stats[p].converged = rn[p] <= target[p];
    }
  return FE_STATUS_OK;
}

// Per-problem AMG-PCG: each problem its own scalars and stop, the union
// V-cycle as the preconditioner (FP64 vectors), per-problem FP64 refinement
// restarts from the true residual when the recursive one met its target
// first.  Iteration grid = tiles of the problems still iterating.
int solveBatchedAmgPerProblem(fe_hip_ctx *ctx, FeBatch &B, const fe_hip_solver_opts *o, fe_hip_batch_stats *stats,
                              double *ms_solve)
{
  const int n = ctx->n, nprob = B.nprob, TB = B.TB;
  FE_REQUIRE(ctx, TB == 256 || TB == 512 || TB == 1024, "batched AMG-PCG: block must be 256, 512 or 1024");
  for(double *&q : B.sp)
    if(!q && feAlloc(ctx, &q, n)) return FE_STATUS_OUT_OF_MEMORY;
  BSpVecs v{B.sp[0], B.sp[1], B.sp[2], B.sp[3], B.sp[4], B.sp[5], B.sp[6], B.sp[7], B.sp[8]};
  if(int rc = feAmgSetup(ctx)) return rc;
  // the V-cycle skips the tiles of stopped problems (their done flags) while
  // this solve runs
  if(int rc = feAmgProblemTiles(ctx, nprob, B.rows.data(), ctx->batchSkip ? B.done : nullptr)) return rc;
  struct Detach {
    fe_hip_ctx *c;
    ~Detach() { (void)feAmgProblemTiles(c, 0, nullptr, nullptr); }
  } detach{ctx};
  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  if(int rc = feAmgNumeric(ctx)) return rc;
  const bool v32 = ctx->pcValues32 != 0;
  FE_HIP_TRY(ctx, hipMemsetAsync(B.x64, 0, sizeof(double) * n, ctx->stream));
  std::vector<double> bn, rn;
  if(int rc = residuals(ctx, B, nullptr, bn)) return rc;  // r = b
  rn = bn;
  std::vector<double> target(nprob), tol2(nprob);
  for(int p = 0; p < nprob; ++p) target[p] = std::max(o->rtol * bn[p], o->atol);
  std::vector<int> its(nprob, 0), outerDone(nprob, 0), hDone(nprob), hIter(nprob);
  const BTiles bt = tilesOf(B);
  const BState S = stateOf(B);
  const BParts P = partsOf(B);
  for(int outer = 0; outer < 20; ++outer) {
    bool any = false;
    for(int p = 0; p < nprob; ++p) {
      outerDone[p] = (rn[p] <= target[p] || its[p] >= o->max_iter) ? 1 : 0;
      any |= !outerDone[p];
      // recursive-residual target below the true one (restart aims lower)
      const double tt = (outer == 0 ? 0.5 : 0.05) * target[p];
      tol2[p] = tt * tt;
    }
    if(!any) break;
    FE_HIP_TRY(ctx, hipMemcpyAsync(B.outerDone, outerDone.data(), sizeof(int) * nprob, hipMemcpyHostToDevice, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(B.tol2, tol2.data(), sizeof(double) * nprob, hipMemcpyHostToDevice, ctx->stream));
    if(int rc = setActive(ctx, B, outerDone)) return rc;
    hipLaunchKernelGGL(bSpStart, dim3(B.nTiles), dim3(TB), 0, ctx->stream, B.r64, v, bt, S);
    if(int rc = feAmgApply(ctx, (const double *)v.r, v.u, nullptr, v32)) return rc;
#define FE_SP_B(KER, GRID, ...)                                                                                     \
  do {                                                                                                           \
    if(TB == 256) hipLaunchKernelGGL((KER<4, 256>), dim3(GRID), dim3(256), 0, ctx->stream, __VA_ARGS__);         \
    else if(TB == 512) hipLaunchKernelGGL((KER<4, 512>), dim3(GRID), dim3(512), 0, ctx->stream, __VA_ARGS__);    \
    else hipLaunchKernelGGL((KER<4, 1024>), dim3(GRID), dim3(1024), 0, ctx->stream, __VA_ARGS__);                \
  } while(0)
    FE_SP_B(bSpInit, B.nTiles, ctx->d_ia, ctx->d_ja, (const double *)ctx->d_val, v, bt, P, S);
    int budget = 0;
    for(int p = 0; p < nprob; ++p)
      if(!outerDone[p]) budget = std::max(budget, o->max_iter - its[p]);
    // one chunk: kBatchGraphIters iterations and the advance of the device
    // iteration base; replayed from a captured graph unless profiling (the
    // same launches either way)
    auto chunk = [&]() -> int {
      for(int k = 0; k < kBatchGraphIters; ++k) {
        if(int rc = feAmgApply(ctx, (const double *)v.w, v.m, nullptr, v32)) return rc;
        FE_SP_B(bSpIter, B.nAct, ctx->d_ia, ctx->d_ja, (const double *)ctx->d_val, v, bt, P, S, (const int *)B.act,
                (const int *)B.itBase, k, budget);
      }
      hipLaunchKernelGGL(bAdvance, dim3(1), dim3(64), 0, ctx->stream, B.itBase, kBatchGraphIters);
      return FE_STATUS_OK;
    };
    const bool useGraph = ctx->batchGraph && !ctx->profiling;
    FE_HIP_TRY(ctx, hipMemsetAsync(B.itBase, 0, sizeof(int), ctx->stream));
    for(int it = 0; it <= budget; it += kBatchGraphIters) {
      if(useGraph) {
        const int key[4] = {B.nAct, budget, TB, v32 ? 1 : 0};
        if(!B.ampExec || !std::equal(key, key + 4, B.ampKey)) {
          if(B.ampExec) (void)hipGraphExecDestroy(B.ampExec);
          if(B.ampGraph) (void)hipGraphDestroy(B.ampGraph);
          B.ampExec = nullptr;
          B.ampGraph = nullptr;
          FE_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
          const int rc = chunk();
          hipGraph_t g = nullptr;
          const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
          if(rc) return rc;
          FE_HIP_TRY(ctx, e);
          B.ampGraph = g;
          FE_HIP_TRY(ctx, hipGraphInstantiate(&B.ampExec, g, nullptr, nullptr, 0));
          std::copy(key, key + 4, B.ampKey);
        }
        FE_HIP_TRY(ctx, hipGraphLaunch(B.ampExec, ctx->stream));
      } else if(int rc = chunk()) {
        return rc;
      }
      FE_HIP_TRY(ctx, hipGetLastError());
      FE_HIP_TRY(ctx, hipMemcpyAsync(hDone.data(), B.done, sizeof(int) * nprob, hipMemcpyDeviceToHost, ctx->stream));
      FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      if(std::all_of(hDone.begin(), hDone.end(), [](int d) { return d != 0; })) break;
      int active = 0;
      for(int p = 0; p < nprob; ++p)
        if(!hDone[p]) active += B.hPtile[p + 1] - B.hPtile[p];
      // the iteration grid shrinks (a recapture) below 3/4 of its tiles
      if(active < (3 * B.nAct) / 4)
        if(int rc = setActive(ctx, B, hDone)) return rc;
    }
#undef FE_SP_B
    FE_HIP_TRY(ctx, hipMemcpyAsync(hIter.data(), B.iter, sizeof(int) * nprob, hipMemcpyDeviceToHost, ctx->stream));
    hipLaunchKernelGGL(bAddCorrection<double>, dim3(1024), dim3(256), 0, ctx->stream, n, B.x64, (const double *)v.x);
    std::vector<double> rnew;
    if(int rc = residuals(ctx, B, B.x64, rnew)) return rc;
    for(int p = 0; p < nprob; ++p) {
      if(outerDone[p]) continue;
      its[p] += hIter[p];
      rn[p] = rnew[p];
    }
  }
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, B.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  if(ctx->profiling) {
    if(int rc = feTimerCollect(ctx, ctx->spmvTimer)) return rc;
    ctx->spmvTimer.used = 0;
  }
  if(int rc = feAmgCheck(ctx)) return rc;
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if(ms_solve) *ms_solve = ms;
  if(stats)
    for(int p = 0; p < nprob; ++p) {
      stats[p].iterations = its[p];
      stats[p].relres_fp64 = bn[p] > 0. ? rn[p] / bn[p] : 0.;
      stats[p].converged = rn[p] <= target[p];
    }
  return FE_STATUS_OK;
}

}  // namespace

extern "C" int fe_hip_solve_batched(fe_hip_ctx *ctx, int nprob, const int *rows, const fe_hip_solver_opts *o,
                                    fe_hip_batch_stats *stats, double *ms_solve)
{
  if(!ctx || !rows || !o) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, o->ksp == FE_KSP_CG && (o->pc == FE_PC_JACOBI || o->pc == FE_PC_AMG),
             "the batched solver runs Jacobi-PCG or AMG-PCG");
  const bool fp32 = (o->precision == FE_PREC_FP32_IR);
  FE_REQUIRE(ctx, !fp32 || ctx->d_val32, "assemble with FE_PREC_FP32_IR to create the FP32 values");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  ctx->spmvTimer = FeKernelTimer{std::move(ctx->spmvTimer.ev)};  // profiling: this solve's launches only
  FeBatch *Bp = nullptr;
  if(int rc = setupBatch(ctx, nprob, rows, Bp)) return rc;
  FeBatch &B = *Bp;
  if(int rc = buildTiles(ctx, B, ctx->iterBlock)) return rc;
  if(o->pc == FE_PC_AMG)
    return ctx->batchAmgGlobal ? solveBatchedAmg(ctx, B, o, stats, ms_solve)
                               : solveBatchedAmgPerProblem(ctx, B, o, stats, ms_solve);
  const int n = ctx->n;
  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  hipLaunchKernelGGL(bDinv<double>, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_diag, ctx->d_val,
                     B.v64.dinv);
  if(fp32)
    hipLaunchKernelGGL(bDinv<float>, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_diag, ctx->d_val,
                       B.v32.dinv);
  FE_HIP_TRY(ctx, hipMemsetAsync(B.x64, 0, sizeof(double) * n, ctx->stream));
  std::vector<double> bn, rn;
  if(int rc = residuals(ctx, B, nullptr, bn)) return rc;  // r = b
  rn = bn;
  std::vector<double> target(nprob);
  for(int p = 0; p < nprob; ++p) target[p] = std::max(o->rtol * bn[p], o->atol);
  std::vector<int> its(nprob, 0), outerDone(nprob, 0), hDone(nprob), hIter(nprob), reason(nprob, 0);
  std::vector<double> tol2(nprob);
  bool inner32 = fp32;
  for(int outer = 0; outer < 60; ++outer) {
    bool any = false;
    for(int p = 0; p < nprob; ++p) {
      outerDone[p] = (rn[p] <= target[p] || its[p] >= o->max_iter || reason[p] < 0) ? 1 : 0;
      any |= !outerDone[p];
      const double t = inner32 ? o->inner_rtol * rn[p] : 0.5 * target[p];
      tol2[p] = std::max(t * t, 0.01 * target[p] * target[p]);
    }
    if(!any) break;
    FE_HIP_TRY(ctx, hipMemcpyAsync(B.outerDone, outerDone.data(), sizeof(int) * nprob, hipMemcpyHostToDevice, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(B.tol2, tol2.data(), sizeof(double) * nprob, hipMemcpyHostToDevice, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(B.bak64, B.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    int budget = 0;
    for(int p = 0; p < nprob; ++p)
      if(!outerDone[p]) budget = std::max(budget, o->max_iter - its[p]);
    if(inner32) {
      if(int rc = runInner<float>(ctx, B, ctx->d_val32, B.v32, budget, 1, hDone, hIter, outerDone)) return rc;
      hipLaunchKernelGGL(bAddCorrection<float>, dim3(1024), dim3(256), 0, ctx->stream, n, B.x64, B.v32.x);
    } else {
      if(int rc = runInner<double>(ctx, B, ctx->d_val, B.v64, budget, 0, hDone, hIter, outerDone)) return rc;
      hipLaunchKernelGGL(bAddCorrection<double>, dim3(1024), dim3(256), 0, ctx->stream, n, B.x64, B.v64.x);
    }
    std::vector<double> rnew;
    if(int rc = residuals(ctx, B, B.x64, rnew)) return rc;
    bool reject = false;
    for(int p = 0; p < nprob; ++p) {
      if(outerDone[p]) continue;
      its[p] += hIter[p];
      // a breakdown only ends this inner solve (the next outer iteration
      // restarts from the FP64 residual); NaN ends the problem
      if(hDone[p] == 4) reason[p] = -9;
      if(inner32 && !(rnew[p] < 0.9 * rn[p])) {
        // undo this problem's FP32 correction
        const size_t r0 = B.rows[p], cnt = B.rows[p + 1] - B.rows[p];
        FE_HIP_TRY(ctx, hipMemcpyAsync(B.x64 + r0, B.bak64 + r0, sizeof(double) * cnt, hipMemcpyDeviceToDevice,
                                       ctx->stream));
        reject = true;
        reason[p] = 0;
        continue;
      }
      rn[p] = rnew[p];
    }
    if(reject) {
      inner32 = false;
      if(int rc = residuals(ctx, B, B.x64, rnew)) return rc;
      for(int p = 0; p < nprob; ++p)
        if(!outerDone[p]) rn[p] = rnew[p];
    }
  }
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, B.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if(ms_solve) *ms_solve = ms;
  if(stats)
    for(int p = 0; p < nprob; ++p) {
      stats[p].iterations = its[p];
      stats[p].relres_fp64 = bn[p] > 0. ? rn[p] / bn[p] : 0.;
      stats[p].converged = rn[p] <= target[p];
    }
  return FE_STATUS_OK;
}

// Times kBatchGraphIters inner FP32 (or FP64) iterations of every launch shape
// and keeps the fastest (spmv_lanes, iter_block).
extern "C" int fe_hip_autotune_batched(fe_hip_ctx *ctx, int nprob, const int *rows, int precision, int *lanes,
                                       int *block, double *best_ms)
{
  if(!ctx || !rows) return FE_STATUS_POINTER_ERROR;
  const bool fp32 = (precision == FE_PREC_FP32_IR);
  FE_REQUIRE(ctx, !fp32 || ctx->d_val32, "assemble with FE_PREC_FP32_IR to create the FP32 values");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  FeBatch *Bp = nullptr;
  if(int rc = setupBatch(ctx, nprob, rows, Bp)) return rc;
  FeBatch &B = *Bp;
  const int n = ctx->n;
  std::vector<int> zeros(nprob, 0);
  std::vector<double> tiny(nprob, 0.);
  FE_HIP_TRY(ctx, hipMemcpy(B.outerDone, zeros.data(), sizeof(int) * nprob, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(B.tol2, tiny.data(), sizeof(double) * nprob, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(B.r64, ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice));
  hipLaunchKernelGGL(bDinv<double>, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_diag, ctx->d_val,
                     B.v64.dinv);
  hipLaunchKernelGGL(bDinv<float>, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, n, ctx->d_diag, ctx->d_val,
                     B.v32.dinv);
  const int saveL = ctx->spmvLanes, saveB = ctx->iterBlock;
  double bestMs = 1e300;
  int bestL = saveL, bestB = saveB;
  hipEvent_t e0, e1;
  FE_HIP_TRY(ctx, hipEventCreate(&e0));
  FE_HIP_TRY(ctx, hipEventCreate(&e1));
  for(int TB : {256, 512, 1024})
    for(int L : {2, 4, 8}) {
      ctx->spmvLanes = L;
      ctx->iterBlock = TB;
      if(int rc = buildTiles(ctx, B, TB)) return rc;
      if(int rc = setActive(ctx, B, zeros)) return rc;
      const int prec = fp32 ? 1 : 0;
      if(fp32 ? captureGraph<float>(ctx, B, ctx->d_val32, B.v32, 1 << 30, prec)
              : captureGraph<double>(ctx, B, ctx->d_val, B.v64, 1 << 30, prec))
        return FE_STATUS_ERROR;
      double ms2[2];
      for(int rep = 0; rep < 2; ++rep) {
        FE_HIP_TRY(ctx, hipMemsetAsync(B.itBase, 0, sizeof(int), ctx->stream));
        if(fp32) dispatch<float, InitL>(ctx, L, TB, B, ctx->d_val32, B.v32);
        else dispatch<double, InitL>(ctx, L, TB, B, ctx->d_val, B.v64);
        FE_HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
        FE_HIP_TRY(ctx, hipGraphLaunch(B.exec, ctx->stream));
        FE_HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
        FE_HIP_TRY(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, e0, e1));
        ms2[rep] = ms;
      }
      const double ms = ms2[1] / kBatchGraphIters;
      if(ms < bestMs) {
        bestMs = ms;
        bestL = L;
        bestB = TB;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->spmvLanes = bestL;
  ctx->iterBlock = bestB;
  ctx->lanesFixed = true;
  if(lanes) *lanes = bestL;
  if(block) *block = bestB;
  if(best_ms) *best_ms = bestMs;
  return FE_STATUS_OK;
}
