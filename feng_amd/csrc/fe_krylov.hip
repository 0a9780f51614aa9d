// Device Krylov layer: the solve of feLinearSystemPETSc::solve
// (feLinearSystemPETSc.cpp:900-1031) with the KSP/PC the configs select.
//
// Preconditioned CG is two fused kernels per iteration with every scalar kept
// on the device (no host round trip inside the iteration):
//   K1 cgSpmv   : beta = rz / rz_old, stop test on r.r, p_new = z + beta p_old
//                 (computed on the fly, also inside the gather), q = A p_new,
//                 partials of p_new . q
//   K2 cgUpdate : alpha = rz / (p.q), x += alpha p, r -= alpha q, z = D^-1 r,
//                 partials of r.z and r.r
// p is double-buffered (the gather of K1 reads p_old while rows write p_new).
// Dot products: every block stores its partial sum; the next kernel of the
// iteration folds the partials at its start (fixed order: deterministic), so
// no block ever issues a device-scope fence.  The host polls the iteration
// state every `check_every` iterations; kernels launched after the stop test
// fired return immediately.
//
// Convergence is judged on the FP64 true residual ||b - A x||_2 / ||b||_2:
// when the recursive residual meets the target, b - A x is recomputed in FP64
// and CG restarts from it if needed.  FE_PREC_FP32_IR runs the same CG in FP32
// (values, vectors; FP64 dot accumulation) as the inner solver of FP64
// iterative refinement x += A32^-1 (b - A64 x).
#include "fe_hip_internal.h"
#include "fe_tile.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

using namespace fe_tile;

// Iteration state mirrored to the host at every check.
struct CgState {
  double rr;      // ||r||^2 at the last checked iteration
  int iter;       // iterations completed when `done` was set
  int done;       // 0 running, 1 converged, 2 iteration budget, 3 breakdown, 4 NaN
  double hist[64];
};

// Iterations to launch before the next host check: the recursive residual's
// geometric decay over the last (up to 16) iterations predicts how many are
// left; launched iterations past convergence only exit early, but each still
// costs a dispatch.
int feNextBatch(const CgState *hs, int it, int last, int checkEvery, double tol2)
{
  const int m = std::min(last, 16);
  if(it <= m) return checkEvery;
  const double r1 = hs->hist[(it - 1) & 63], r0 = hs->hist[(it - 1 - m) & 63];
  if(!(r1 > 0.) || !(r0 > r1)) return checkEvery;
  const double rate = std::log(r1 / r0) / m;
  const double need = std::log(tol2 / r1) / rate;
  if(!(need > 0.)) return 1;
  return std::max(1, std::min(checkEvery, (int)std::ceil(need) + 1));
}

// Iteration index base, budget and stop threshold of a graph-replayed
// iteration batch: the captured kernels carry only their offset in the batch;
// the last node of the graph advances the base.
struct IterParams {
  int itBase;
  int maxit;
  double tol2;
};

__global__ void iterAdvance(IterParams *ip, int k)
{
  if(blockIdx.x == 0 && threadIdx.x == 0) ip->itBase += k;
}

constexpr int kTB = 256;       // block size of the one-off kernels
constexpr int kMaxGrid = 4096; // partial buffers hold this many blocks

// Blocks of a row-tile iteration kernel: one per tile up to iter_grid, rounded
// up to a multiple of 8 for the XCD-aware tile order (blocks past the last
// tile store zero partials).
int feIterGrid(const fe_hip_ctx *ctx, int nTiles)
{
  int G = std::max(1, std::min(ctx->iterGrid, nTiles));
  if(ctx->xcdTiles) G = std::min(kMaxGrid, (G + 7) & ~7);
  return G;
}
constexpr int kGrid = 1024;    // grid of the one-off reduction kernels

template <typename T>
struct CgVecs {
  T *x, *r, *z, *p[2], *q, *dinv;
};

// Partial-sum buffers: rz and rr double-buffered by iteration parity.
struct CgParts {
  double *rz[2], *rr[2], *pq;
};

// Folds up to three partial arrays (G entries each) in one pass; every
// thread gets the totals.
template <int TB, int NV>
__device__ __forceinline__ void feFold(const double *const (&parts)[NV], int G, double (&out)[NV])
{
  __shared__ double lds[NV][TB / 64 + 1];
  double s[NV];
#pragma unroll
  for(int k = 0; k < NV; ++k) s[k] = 0.;
  for(int b = threadIdx.x; b < G; b += TB)
#pragma unroll
    for(int k = 0; k < NV; ++k) s[k] += parts[k][b];
#pragma unroll
  for(int k = 0; k < NV; ++k) s[k] = feWaveSum(s[k]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if(lane == 0)
#pragma unroll
    for(int k = 0; k < NV; ++k) lds[k][wid] = s[k];
  __syncthreads();
  if(threadIdx.x == 0)
#pragma unroll
    for(int k = 0; k < NV; ++k) {
      double t = 0.;
      for(int i = 0; i < TB / 64; ++i) t += lds[k][i];
      lds[k][TB / 64] = t;
    }
  __syncthreads();
#pragma unroll
  for(int k = 0; k < NV; ++k) out[k] = lds[k][TB / 64];
}

// XCD-aware block order: hardware dispatches block b to XCD b % 8; map the 8
// XCDs onto 8 contiguous row ranges so each XCD's L2 sees neighbouring rows.
__device__ __forceinline__ int feXcdBlock()
{
  const int G = gridDim.x;
  if(G % 8) return blockIdx.x;
  return (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
}

template <typename T>
__global__ __launch_bounds__(kTB) void extractDinv(int n, const int *__restrict__ diag,
                                                   const double *__restrict__ val, T *__restrict__ dinv)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const int p = diag[i];
  const double d = p >= 0 ? val[p] : 0.;
  dinv[i] = (T)(d != 0. ? 1. / d : 1.);
}

// r = b (or keep r), z = D^-1 r, p_0 = 0 (and x = 0); partials of r.z, r.r
// into slot 0.
template <typename T, int TB>
__global__ __launch_bounds__(TB) void cgInit(int n, const double *__restrict__ b, CgVecs<T> v, CgParts P,
                                             CgState *st, int fromResidual)
{
  __shared__ double lds[TB / 64 + 1];
  double a0 = 0., a1 = 0.;
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T ri = fromResidual ? v.r[i] : (T)b[i];
    const T zi = v.dinv[i] * ri;
    v.r[i] = ri;
    v.z[i] = zi;
    v.p[0][i] = (T)0;
    if(!fromResidual) v.x[i] = (T)0;
    a0 += (double)ri * (double)zi;
    a1 += (double)ri * (double)ri;
  }
  feStoreBlockSum(a0, P.rz[0], lds);
  feStoreBlockSum(a1, P.rr[0], lds);
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->iter = 0;
  }
}

// Iteration `it`: beta = rz_it / rz_{it-1} (0 at it = 0), stop test on
// rr_it, then p_new = z + beta p_old (on the fly inside the gather too),
// q = A p_new, partials of p_new . q.  LPR lanes per row; UNR row groups per
// lane group in flight (independent load chains for latency hiding).
template <typename T, int LPR, int TB, int UNR>
__global__ __launch_bounds__(TB) void cgSpmv(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                             const T *__restrict__ val, const T *__restrict__ z,
                                             const T *__restrict__ pOld, T *__restrict__ pNew,
                                             T *__restrict__ q, CgParts P, CgState *st, int it, int maxit,
                                             double tol2)
{
  __shared__ double lds[TB / 64 + 1];
  if(st->done) return;
  const int s = it & 1;
  double f[3];
  const double *const parts[3] = {P.rr[s], P.rz[s], P.rz[s ^ 1]};
  feFold<TB, 3>(parts, gridDim.x, f);
  const double rr = f[0];
  const bool stop = !(rr > tol2) || it >= maxit;  // also stops on NaN
  if(stop) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->iter = it;
      st->done = (rr <= tol2) ? 1 : (rr == rr ? 2 : 4);
    }
    return;
  }
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->hist[it & 63] = rr;
  }
  const T beta = (T)(it > 0 ? f[1] / f[2] : 0.);
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double acc = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  const int stride = gridDim.x * UNR;
  for(int c0 = feXcdBlock() * UNR; c0 < nChunks; c0 += stride) {
    int row[UNR], kb[UNR], ke[UNR];
    T sum[UNR];
#pragma unroll
    for(int u = 0; u < UNR; ++u) {
      row[u] = (c0 + u) * RPB + threadIdx.x / LPR;
      kb[u] = 0;
      ke[u] = 0;
      if(row[u] < n) {
        kb[u] = ia[row[u]] + sub;
        ke[u] = ia[row[u] + 1];
      }
      sum[u] = (T)0;
    }
#pragma unroll
    for(int u = 0; u < UNR; ++u)
      sum[u] = feRowDot<FeKs<LPR>::value, LPR>(ja, val, kb[u], ke[u], [&](int j) { return z[j] + beta * pOld[j]; });
#pragma unroll
    for(int u = 0; u < UNR; ++u) {
#pragma unroll
      for(int o = LPR / 2; o > 0; o >>= 1) sum[u] += __shfl_xor(sum[u], o, LPR);
      if(sub == 0 && row[u] < n) {
        const T pn = z[row[u]] + beta * pOld[row[u]];
        pNew[row[u]] = pn;
        q[row[u]] = sum[u];
        acc += (double)pn * (double)sum[u];
      }
    }
  }
  feStoreBlockSum(acc, P.pq, lds);
}

// alpha = rz_it / (p.q); x += alpha p, r -= alpha q, z = D^-1 r; partials of
// r.z and r.r into slot it+1.
template <typename T, int TB>
__global__ __launch_bounds__(TB) void cgUpdate(int n, CgVecs<T> v, const T *__restrict__ p, CgParts P,
                                               CgState *st, int it)
{
  __shared__ double lds[TB / 64 + 1];
  if(st->done) return;
  const int s = it & 1;
  double f[2];
  const double *const parts[2] = {P.pq, P.rz[s]};
  feFold<TB, 2>(parts, gridDim.x, f);
  const double pq = f[0], rz = f[1];
  if(!(pq > 0.)) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->iter = it;
      st->done = 3;  // breakdown: p.Ap <= 0, the matrix is not SPD
    }
    return;
  }
  const T alpha = (T)(rz / pq);
  double a0 = 0., a1 = 0.;
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    v.x[i] += alpha * p[i];
    const T ri = v.r[i] - alpha * v.q[i];
    const T zi = v.dinv[i] * ri;
    v.r[i] = ri;
    v.z[i] = zi;
    a0 += (double)ri * (double)zi;
    a1 += (double)ri * (double)ri;
  }
  feStoreBlockSum(a0, P.rz[s ^ 1], lds);
  feStoreBlockSum(a1, P.rr[s ^ 1], lds);
}


// ---------------------------------------------------------------------------
// PCG with the split factored preconditioner M^-1 = G^T D^-1 G (G unit lower
// on the IC(0) pattern): three kernels per iteration, every vector update
// fused into one of them and every scalar folded from block partials.
//   cgSpmv  : beta, p_new = z + beta p_old, q = A p_new, partials of p.q
//   applyG  : alpha = rz / p.q, x += alpha p, r_new = r - alpha q,
//             y = G r_new (r_new gathered on the fly from r, q), partials r.r
//   applyGT : z = G^T D^-1 y, partials of r_new . z
template <typename T>
struct SplitVecs {
  T *x, *r[2], *p[2], *q, *y, *z;
  const T *gL, *gU, *dinv;
  const int *iaL, *jaL, *iaU, *jaU;
};

template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void splitInit(int n, const double *__restrict__ b, SplitVecs<T> v, CgParts P,
                                                CgState *st)
{
  __shared__ double lds[TB / 64 + 1];
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double rr = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    T sum = (T)0;
    if(row < n)
      sum = feRowDot<FeKs<LPR>::value, LPR>(v.jaL, v.gL, v.iaL[row] + sub, v.iaL[row + 1],
                                             [&](int j) { return (T)b[j]; });
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, LPR);
    if(sub == 0 && row < n) {
      const T bi = (T)b[row];
      v.r[0][row] = bi;
      v.x[row] = (T)0;
      v.p[0][row] = (T)0;
      v.p[1][row] = (T)0;
      v.y[row] = sum;
      rr += (double)bi * (double)bi;
    }
  }
  feStoreBlockSum(rr, P.rr[0], lds);
  if(threadIdx.x == 0) P.rz[1][blockIdx.x] = (blockIdx.x == 0) ? 1. : 0.;
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->iter = 0;
  }
}

template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void splitApplyG(int n, SplitVecs<T> v, CgParts P, CgState *st, int it)
{
  __shared__ double lds[TB / 64 + 1];
  if(st->done) return;
  const int c = it & 1, o = c ^ 1;
  double f[2];
  const double *const parts[2] = {P.pq, P.rz[c]};
  feFold<TB, 2>(parts, gridDim.x, f);
  if(!(f[0] > 0.)) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->iter = it;
      st->done = 3;
    }
    return;
  }
  const T alpha = (T)(f[1] / f[0]);
  const T *__restrict__ rc = v.r[c];
  const T *__restrict__ pn = v.p[o];  // p_new written by cgSpmv of this iteration
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double rr = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    T sum = (T)0;
    if(row < n)
      sum = feRowDot<FeKs<LPR>::value, LPR>(v.jaL, v.gL, v.iaL[row] + sub, v.iaL[row + 1],
                                             [&](int j) { return rc[j] - alpha * v.q[j]; });
#pragma unroll
    for(int q = LPR / 2; q > 0; q >>= 1) sum += __shfl_xor(sum, q, LPR);
    if(sub == 0 && row < n) {
      const T rn = rc[row] - alpha * v.q[row];
      v.x[row] += alpha * pn[row];
      v.r[o][row] = rn;
      v.y[row] = sum;
      rr += (double)rn * (double)rn;
    }
  }
  feStoreBlockSum(rr, P.rr[o], lds);
}

// z = G^T D^-1 y; partials of r[rs] . z into rz[slot].
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void splitApplyGT(int n, SplitVecs<T> v, CgParts P, CgState *st, int rs, int slot,
                                                   int checkDone)
{
  __shared__ double lds[TB / 64 + 1];
  if(checkDone && st->done) return;
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double rz = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    T sum = (T)0;
    if(row < n)
      sum = feRowDot<FeKs<LPR>::value, LPR>(v.jaU, v.gU, v.iaU[row] + sub, v.iaU[row + 1],
                                             [&](int j) { return v.dinv[j] * v.y[j]; });
#pragma unroll
    for(int q = LPR / 2; q > 0; q >>= 1) sum += __shfl_xor(sum, q, LPR);
    if(sub == 0 && row < n) {
      v.z[row] = sum;
      rz += (double)v.r[rs][row] * (double)sum;
    }
  }
  feStoreBlockSum(rz, P.rz[slot], lds);
}

// ---------------------------------------------------------------------------
// Pipelined Jacobi-PCG (Chronopoulos & Gear 1989): one fused kernel per
// iteration and one reduction phase (gamma = r.u, delta = w.u, rho = r.r with
// u = D^-1 r, w = A u).  Kernel i, with alpha_i / beta_i from the folded
// partials of iteration i and the scalars of iteration i-1:
//   p_i = u_i + beta_i p_{i-1},   s_i = w_i + beta_i s_{i-1}
//   x_{i+1} = x_i + alpha_i p_i,  r_{i+1} = r_i - alpha_i s_i
//   w_{i+1} = A u_{i+1}, u_{i+1} = D^-1 r_{i+1} recomputed on the fly for
//   every gathered column from r_i, w_i, s_{i-1} (read-only this kernel).
// r, w and s are double-buffered by iteration parity; x, p are own-row only.
// beta_i = gamma_i / gamma_{i-1}, alpha_i = gamma_i / (delta_i - beta_i
// gamma_i / alpha_{i-1}); the start writes gamma_{-1} = 1, alpha_{-1} = inf,
// s_{-1} = p_{-1} = 0, so every iteration runs the same code.
template <typename T>
struct PipeVecs {
  T *x, *p, *dinv, *r[2], *w[2], *s[2];
};

struct PipeParts {
  double *g[2], *d[2], *rho[2];  // partials of gamma, delta, rho by parity
};

struct PipeScalars {
  double gamma[2], alpha[2];  // by parity, written by block 0
};

// w_0 = A (D^-1 r_0); r_0 = b (or the given residual); x = 0 unless
// restarting; s_{-1} = p_{-1} = 0; partials of iteration 0.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void pipeInit(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const T *__restrict__ val, const double *__restrict__ b,
                                               PipeVecs<T> v, PipeParts P, PipeScalars *sc, CgState *st,
                                               int fromResidual)
{
  __shared__ double lds[TB / 64 + 1];
  // r_0 (slot 0) must be complete before the gather: when starting from b the
  // gather reads b directly; when restarting, r was written by an earlier
  // kernel.
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double g = 0., d = 0., rho = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  const T *r0 = v.r[0];
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    T sum = (T)0;
    if(row < n)
      sum = feRowDot<FeKs<LPR>::value, LPR>(ja, val, ia[row] + sub, ia[row + 1], [&](int j) {
        const T rj = fromResidual ? r0[j] : (T)b[j];
        return v.dinv[j] * rj;
      });
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, LPR);
    if(sub == 0 && row < n) {
      const T ri = fromResidual ? r0[row] : (T)b[row];
      const T ui = v.dinv[row] * ri;
      if(!fromResidual) {
        v.r[0][row] = ri;
        v.x[row] = (T)0;
      }
      v.w[0][row] = sum;
      v.s[1][row] = (T)0;
      v.p[row] = (T)0;
      g += (double)ri * (double)ui;
      d += (double)sum * (double)ui;
      rho += (double)ri * (double)ri;
    }
  }
  feStoreBlockSum(g, P.g[0], lds);
  feStoreBlockSum(d, P.d[0], lds);
  feStoreBlockSum(rho, P.rho[0], lds);
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[1] = 1.;
    sc->alpha[1] = __builtin_inf();
    st->done = 0;
    st->iter = 0;
  }
}

template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void pipeIter(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const T *__restrict__ val, PipeVecs<T> v, PipeParts P,
                                               PipeScalars *sc, CgState *st, int it, int maxit, double tol2)
{
  __shared__ double lds[TB / 64 + 1];
  if(st->done) return;
  const int c = it & 1, o = c ^ 1;
  double f[3];
  const double *const parts[3] = {P.g[c], P.d[c], P.rho[c]};
  feFold<TB, 3>(parts, gridDim.x, f);
  const double gamma = f[0], delta = f[1], rho = f[2];
  if(!(rho > tol2) || it >= maxit) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rho;
      st->iter = it;
      st->done = (rho <= tol2) ? 1 : (rho == rho ? 2 : 4);
    }
    return;
  }
  const double betaD = gamma / sc->gamma[o];
  const double den = delta - betaD * gamma / sc->alpha[o];
  if(!(den > 0.)) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->iter = it;
      st->done = 3;  // breakdown: the preconditioned operator is not SPD
    }
    return;
  }
  const double alphaD = gamma / den;
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[c] = gamma;
    sc->alpha[c] = alphaD;
    st->rr = rho;
    st->hist[it & 63] = rho;
  }
  const T alpha = (T)alphaD, beta = (T)betaD;
  const T *__restrict__ rC = v.r[c];
  const T *__restrict__ wC = v.w[c];
  const T *__restrict__ sO = v.s[o];
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR;
  double g = 0., d = 0., rh = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    T sum = (T)0;
    if(row < n)
      sum = feRowDot<FeKs<LPR>::value, LPR>(ja, val, ia[row] + sub, ia[row + 1], [&](int j) {
        return v.dinv[j] * (rC[j] - alpha * (wC[j] + beta * sO[j]));
      });
#pragma unroll
    for(int q = LPR / 2; q > 0; q >>= 1) sum += __shfl_xor(sum, q, LPR);
    if(sub == 0 && row < n) {
      const T di = v.dinv[row];
      const T ri = rC[row];
      const T pi = di * ri + beta * v.p[row];
      const T si = wC[row] + beta * sO[row];
      const T rn = ri - alpha * si;
      const T un = di * rn;
      v.x[row] += alpha * pi;
      v.p[row] = pi;
      v.s[c][row] = si;
      v.r[o][row] = rn;
      v.w[o][row] = sum;
      g += (double)rn * (double)un;
      d += (double)sum * (double)un;
      rh += (double)rn * (double)rn;
    }
  }
  feStoreBlockSum(g, P.g[o], lds);
  feStoreBlockSum(d, P.d[o], lds);
  feStoreBlockSum(rh, P.rho[o], lds);
}

// Tiled pipelined Jacobi-PCG (Ghysels & Vanroose, "Hiding global
// synchronization latency in the preconditioned conjugate gradient
// algorithm", Parallel Computing 40 (2014), Alg. 3, with M = D) in one kernel
// per iteration.  With a diagonal preconditioner the vectors u = D^-1 r,
// q = D^-1 s and m = D^-1 w follow from stored ones, so the iteration keeps
// x, r, w, s, z, p (own row only, updated in place) and m (the only gathered
// vector, double-buffered by parity):
//   beta = gamma_i / gamma_{i-1}, alpha = gamma_i / (delta_i - beta gamma_i / alpha_{i-1})
//   y = A m_i (gathered), z = y + beta z, s = w + beta s, p = D^-1 r + beta p,
//   x += alpha p, r -= alpha s, w -= alpha z, m_{i+1} = D^-1 w,
//   partials of gamma = (r, D^-1 r), delta = (w, D^-1 r), rho = (r, r).
// A block owns tiles of TB rows.  In the product phase each group of LPR
// lanes takes LPR rows of the tile, their gathers interleaved; the row sums
// go through LDS to the update phase, where thread t updates row t of the
// tile with coalesced loads and stores.  The row bounds and the own-row
// values of the first tile are loaded before the partials of the previous
// iteration are folded, so those loads overlap the fold.
// r_0 = b (or the residual already in r), x = 0 unless restarting,
// w_0 = A D^-1 r_0, m_0 = D^-1 w_0, s = z = p = 0; partials of iteration 0.
// When starting from b the product gathers D^-1 b directly; when restarting,
// r was written by an earlier kernel.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void tileInit(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const T *__restrict__ val, const double *__restrict__ b,
                                               TileVecs<T> v, PipeParts P, PipeScalars *sc, CgState *st,
                                               int fromResidual)
{
  __shared__ T ldsY[TB];
  const int nTiles = (n + TB - 1) / TB;
  double g = 0., d = 0., rho = 0.;
  for(int tile = tileStart(); tile < nTiles; tile += gridDim.x) {
    const int base = tile * TB;
    int kb[LPR], ke[LPR];
    tileBounds<LPR, TB>(n, base, ia, kb, ke);
    if(fromResidual) tileProduct<T, LPR, TB, true>(n, base, ia, ja, val, v.r, v.dinv, kb, ke, ldsY);
    else {
      // gather D^-1 b in T precision
      constexpr int RPB = TB / LPR;
      const int sub = threadIdx.x % LPR, gg = threadIdx.x / LPR;
#pragma unroll
      for(int u = 0; u < LPR; ++u) {
        T acc =
            feRowDot<FeKs<LPR>::value, LPR>(ja, val, kb[u], ke[u], [&](int j) { return v.dinv[j] * (T)b[j]; });
#pragma unroll
        for(int q = LPR / 2; q > 0; q >>= 1) acc += __shfl_xor(acc, q, LPR);
        if(sub == u) ldsY[u * RPB + gg] = acc;
      }
    }
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) {
      const T di = v.dinv[row];
      const T ri = fromResidual ? v.r[row] : (T)b[row];
      const T wi = ldsY[threadIdx.x];
      const T ui = di * ri;
      if(!fromResidual) {
        v.r[row] = ri;
        v.x[row] = (T)0;
      }
      v.w[row] = wi;
      v.m[0][row] = di * wi;
      v.s[row] = (T)0;
      v.z[row] = (T)0;
      v.p[row] = (T)0;
      g += (double)ri * (double)ui;
      d += (double)wi * (double)ui;
      rho += (double)ri * (double)ri;
    }
    __syncthreads();
  }
  tileStore3<TB>(g, d, rho, P.g[0], P.d[0], P.rho[0]);
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[1] = 1.;
    sc->alpha[1] = __builtin_inf();
    st->done = 0;
    st->iter = 0;
  }
}

template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void tileIter(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const T *__restrict__ val, TileVecs<T> v, PipeParts P,
                                               PipeScalars *sc, CgState *st, int it, int maxit, double tol2)
{
  __shared__ T ldsY[TB];
  const int c = it & 1, o = c ^ 1;
  const int nTiles = (n + TB - 1) / TB;
  int tile = tileStart();
  // first tile: row bounds and own-row values in flight during the fold
  int kb[LPR], ke[LPR];
  TileRow<T> me;
  tileBounds<LPR, TB>(n, tile * TB, ia, kb, ke);
  tileLoadRow(v, tile * TB + threadIdx.x, n, me);
  const int done0 = st->done;
  const double gPrev = sc->gamma[o], aPrev = sc->alpha[o];
  double f[3];
  const double *const parts[3] = {P.g[c], P.d[c], P.rho[c]};
  feFold<TB, 3>(parts, gridDim.x, f);
  const double gamma = f[0], delta = f[1], rho = f[2];
  if(done0) return;
  if(!(rho > tol2) || it >= maxit) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rho;
      st->iter = it;
      st->done = (rho <= tol2) ? 1 : (rho == rho ? 2 : 4);
    }
    return;
  }
  const double betaD = gamma / gPrev;
  const double den = delta - betaD * gamma / aPrev;
  if(!(den > 0.)) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->iter = it;
      st->done = 3;  // breakdown: the preconditioned operator is not SPD
    }
    return;
  }
  const double alphaD = gamma / den;
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[c] = gamma;
    sc->alpha[c] = alphaD;
    st->rr = rho;
    st->hist[it & 63] = rho;
  }
  const T alpha = (T)alphaD, beta = (T)betaD;
  const T *__restrict__ mC = v.m[c];
  T *__restrict__ mO = v.m[o];
  double g = 0., d = 0., rh = 0.;
  while(tile < nTiles) {
    const int base = tile * TB;
    tileProduct<T, LPR, TB, false>(n, base, ia, ja, val, mC, nullptr, kb, ke, ldsY);
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) {
      const T y = ldsY[threadIdx.x];
      const T z = y + beta * me.z;
      const T s = me.w + beta * me.s;
      const T p = me.d * me.r + beta * me.p;
      const T rn = me.r - alpha * s;
      const T wn = me.w - alpha * z;
      const T un = me.d * rn;
      v.x[row] = me.x + alpha * p;
      v.p[row] = p;
      v.s[row] = s;
      v.z[row] = z;
      v.r[row] = rn;
      v.w[row] = wn;
      mO[row] = me.d * wn;
      g += (double)rn * (double)un;
      d += (double)wn * (double)un;
      rh += (double)rn * (double)rn;
    }
    tile += gridDim.x;
    if(tile < nTiles) {
      tileBounds<LPR, TB>(n, tile * TB, ia, kb, ke);
      tileLoadRow(v, tile * TB + threadIdx.x, n, me);
    }
    __syncthreads();
  }
  tileStore3<TB>(g, d, rh, P.g[o], P.d[o], P.rho[o]);
}

// Tiled pipelined PCG with the split preconditioner M^-1 = G^T D^-1 G
// (Ghysels & Vanroose 2014, Alg. 3), three kernels per iteration and one
// reduction phase:
//   K1  t = D^-1 (G w)          gathers w (G: lower CSR with unit diagonal)
//   K2  m = G^T t               gathers t (G^T: upper CSR)
//   K3  fold gamma, delta, rho; alpha, beta; y = A m (gathers m);
//       z = y + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
//       x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z;
//       partials of gamma = (r, u), delta = (w, u), rho = (r, r).
// Every gathered vector is written by an earlier kernel and not modified by
// the kernel that gathers it, so no vector is double-buffered.
template <typename T>
struct SpVecs {
  T *x, *r, *u, *w, *z, *q, *s, *p, *m, *t;
  // FP32 shadows of w, t, m when the preconditioner is applied in FP32 inside
  // an FP64 solve (null otherwise)
  float *w32, *t32, *m32;
};

template <typename T>
struct SpRow {
  T r, u, w, z, q, s, p, x, m;
};

// y = scale .* (M x) over row tiles (scale may be null); skipped once the
// iteration has stopped when checkDone is set.
template <typename T, int LPR, int TB, typename TV = T>
__global__ __launch_bounds__(TB) void spSpmv(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                             const TV *__restrict__ val, const T *__restrict__ x,
                                             const TV *__restrict__ scale, T *__restrict__ y, const CgState *st,
                                             int checkDone)
{
  __shared__ T ldsY[TB];
  if(checkDone && st->done) return;
  const int nTiles = (n + TB - 1) / TB;
  for(int tile = tileStart(); tile < nTiles; tile += gridDim.x) {
    const int base = tile * TB;
    int kb[LPR], ke[LPR];
    tileBounds<LPR, TB>(n, base, ia, kb, ke);
    tileProductF<T, LPR, TB>(ja, val, kb, ke, ldsY, [&](int j) { return x[j]; });
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) y[row] = scale ? (T)scale[row] * ldsY[threadIdx.x] : ldsY[threadIdx.x];
    __syncthreads();
  }
}

// FP64 vectors, FP32 factor values.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void spSpmvF32(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                const float *__restrict__ val, const T *__restrict__ x,
                                                const float *__restrict__ scale, T *__restrict__ y,
                                                const CgState *st, int checkDone)
{
  __shared__ T ldsY[TB];
  if(checkDone && st->done) return;
  const int nTiles = (n + TB - 1) / TB;
  for(int tile = tileStart(); tile < nTiles; tile += gridDim.x) {
    const int base = tile * TB;
    int kb[LPR], ke[LPR];
    tileBounds<LPR, TB>(n, base, ia, kb, ke);
    tileProductF<T, LPR, TB>(ja, val, kb, ke, ldsY, [&](int j) { return x[j]; });
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) y[row] = scale ? (T)scale[row] * ldsY[threadIdx.x] : ldsY[threadIdx.x];
    __syncthreads();
  }
}

// r = b, x = 0 (FP64 b rounded to T)
template <typename T>
__global__ void spStart(int n, const double *__restrict__ b, SpVecs<T> v)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    v.r[i] = (T)b[i];
    v.x[i] = (T)0;
  }
}

// w = A u; z = q = s = p = 0; partials of iteration 0; scalar start values.
template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void spInitW(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                              const T *__restrict__ val, SpVecs<T> v, PipeParts P, PipeScalars *sc,
                                              CgState *st)
{
  __shared__ T ldsY[TB];
  const int nTiles = (n + TB - 1) / TB;
  double g = 0., d = 0., rho = 0.;
  for(int tile = tileStart(); tile < nTiles; tile += gridDim.x) {
    const int base = tile * TB;
    int kb[LPR], ke[LPR];
    tileBounds<LPR, TB>(n, base, ia, kb, ke);
    tileProduct<T, LPR, TB, false>(n, base, ia, ja, val, v.u, nullptr, kb, ke, ldsY);
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) {
      const T wi = ldsY[threadIdx.x], ri = v.r[row], ui = v.u[row];
      v.w[row] = wi;
      if(v.w32) v.w32[row] = (float)wi;
      v.z[row] = v.q[row] = v.s[row] = v.p[row] = (T)0;
      g += (double)ri * (double)ui;
      d += (double)wi * (double)ui;
      rho += (double)ri * (double)ri;
    }
    __syncthreads();
  }
  tileStore3<TB>(g, d, rho, P.g[0], P.d[0], P.rho[0]);
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[1] = 1.;
    sc->alpha[1] = __builtin_inf();
    st->done = 0;
    st->iter = 0;
  }
}

template <typename T>
__device__ __forceinline__ void spLoadRow(const SpVecs<T> &v, int row, int n, SpRow<T> &o)
{
  if(row < n) {
    o.r = v.r[row]; o.u = v.u[row]; o.w = v.w[row]; o.z = v.z[row]; o.q = v.q[row];
    o.s = v.s[row]; o.p = v.p[row]; o.x = v.x[row]; o.m = v.m32 ? (T)v.m32[row] : v.m[row];
  } else {
    o.r = o.u = o.w = o.z = o.q = o.s = o.p = o.x = o.m = (T)0;
  }
}

template <typename T, int LPR, int TB>
__global__ __launch_bounds__(TB) void spIter(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                             const T *__restrict__ val, SpVecs<T> v, PipeParts P, PipeScalars *sc,
                                             CgState *st, int it, int maxit, double tol2, const IterParams *ip)
{
  __shared__ T ldsY[TB];
  if(ip) {
    // graph replay: it is the offset in the captured batch
    it += ip->itBase;
    maxit = ip->maxit;
    tol2 = ip->tol2;
  }
  const int c = it & 1, o = c ^ 1;
  const int nTiles = (n + TB - 1) / TB;
  int tile = tileStart();
  int kb[LPR], ke[LPR];
  SpRow<T> me;
  tileBounds<LPR, TB>(n, tile * TB, ia, kb, ke);
  spLoadRow(v, tile * TB + threadIdx.x, n, me);
  const int done0 = st->done;
  const double gPrev = sc->gamma[o], aPrev = sc->alpha[o];
  double f[3];
  const double *const parts[3] = {P.g[c], P.d[c], P.rho[c]};
  feFold<TB, 3>(parts, gridDim.x, f);
  const double gamma = f[0], delta = f[1], rho = f[2];
  if(done0) return;
  if(!(rho > tol2) || it >= maxit) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rho;
      st->iter = it;
      st->done = (rho <= tol2) ? 1 : (rho == rho ? 2 : 4);
    }
    return;
  }
  const double betaD = gamma / gPrev;
  const double den = delta - betaD * gamma / aPrev;
  if(!(den > 0.)) {
    if(blockIdx.x == 0 && threadIdx.x == 0) {
      st->iter = it;
      st->done = 3;
    }
    return;
  }
  const double alphaD = gamma / den;
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gamma[c] = gamma;
    sc->alpha[c] = alphaD;
    st->rr = rho;
    st->hist[it & 63] = rho;
  }
  const T alpha = (T)alphaD, beta = (T)betaD;
  double g = 0., d = 0., rh = 0.;
  while(tile < nTiles) {
    const int base = tile * TB;
    if(v.m32) tileProductF<T, LPR, TB>(ja, val, kb, ke, ldsY, [&](int j) { return (T)v.m32[j]; });
    else tileProduct<T, LPR, TB, false>(n, base, ia, ja, val, v.m, nullptr, kb, ke, ldsY);
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) {
      const T y = ldsY[threadIdx.x];
      const T z = y + beta * me.z;
      const T q = me.m + beta * me.q;
      const T s = me.w + beta * me.s;
      const T p = me.u + beta * me.p;
      const T rn = me.r - alpha * s;
      const T un = me.u - alpha * q;
      const T wn = me.w - alpha * z;
      v.x[row] = me.x + alpha * p;
      v.z[row] = z;
      v.q[row] = q;
      v.s[row] = s;
      v.p[row] = p;
      v.r[row] = rn;
      v.u[row] = un;
      v.w[row] = wn;
      if(v.w32) v.w32[row] = (float)wn;
      g += (double)rn * (double)un;
      d += (double)wn * (double)un;
      rh += (double)rn * (double)rn;
    }
    tile += gridDim.x;
    if(tile < nTiles) {
      tileBounds<LPR, TB>(n, tile * TB, ia, kb, ke);
      spLoadRow(v, tile * TB + threadIdx.x, n, me);
    }
    __syncthreads();
  }
  tileStore3<TB>(g, d, rh, P.g[o], P.d[o], P.rho[o]);
}

// r = b - A x (FP64); partials of ||r||^2.
template <int LPR>
__global__ __launch_bounds__(kTB) void residual64(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                  const double *__restrict__ val, const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ r,
                                                  double *__restrict__ part)
{
  __shared__ double lds[kTB / 64 + 1];
  constexpr int RPB = kTB / LPR;
  const int sub = threadIdx.x % LPR;
  double acc = 0.;
  const int nChunks = (n + RPB - 1) / RPB;
  for(int chunk = feXcdBlock(); chunk < nChunks; chunk += gridDim.x) {
    const int row = chunk * RPB + threadIdx.x / LPR;
    double sum = 0.;
    if(row < n) sum = feRowDot<FeKs<LPR>::value, LPR>(ja, val, ia[row] + sub, ia[row + 1], [&](int j) { return x[j]; });
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, LPR);
    if(sub == 0 && row < n) {
      const double ri = b[row] - sum;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  feStoreBlockSum(acc, part, lds);
}

// One block folds `count` partial arrays of G entries into out[0..count).
__global__ __launch_bounds__(kTB) void finalizeSums(const double *__restrict__ part, int G, int count,
                                                    double *__restrict__ out)
{
  __shared__ double lds[kTB / 64 + 1];
  for(int k = 0; k < count; ++k) {
    const double s = feFoldPartials(part + (size_t)k * G, G, lds);
    if(threadIdx.x == 0) out[k] = s;
  }
}

template <typename T>
__global__ void addVec(int n, double *__restrict__ x, const T *__restrict__ d)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += (double)d[i];
}


__global__ __launch_bounds__(kTB) void sumSquares(int n, const double *__restrict__ a, double *__restrict__ part)
{
  __shared__ double lds[kTB / 64 + 1];
  double acc = 0.;
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += a[i] * a[i];
  feStoreBlockSum(acc, part, lds);
}

// Per-block maxima of |a|, |b|, |c| (the infinity norms of du, b, A du - b,
// feLinearSystemPETSc.cpp:1004-1015), folded by finalizeMax.
__global__ __launch_bounds__(kTB) void blockMaxAbs3(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                    const double *__restrict__ c, double *__restrict__ part)
{
  __shared__ double lds[3][kTB / 64];
  double m[3] = {0., 0., 0.};
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    m[0] = fmax(m[0], fabs(a[i]));
    m[1] = fmax(m[1], fabs(b[i]));
    m[2] = fmax(m[2], fabs(c[i]));
  }
  for(int k = 0; k < 3; ++k) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) m[k] = fmax(m[k], __shfl_xor(m[k], o, 64));
    if((threadIdx.x & 63) == 0) lds[k][threadIdx.x >> 6] = m[k];
  }
  __syncthreads();
  if(threadIdx.x == 0)
    for(int k = 0; k < 3; ++k) {
      double r = 0.;
      for(int w = 0; w < kTB / 64; ++w) r = fmax(r, lds[k][w]);
      part[(size_t)k * gridDim.x + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(64) void finalizeMax(const double *__restrict__ part, int G, int count,
                                                   double *__restrict__ out)
{
  for(int k = 0; k < count; ++k) {
    double r = 0.;
    for(int b = threadIdx.x; b < G; b += 64) r = fmax(r, part[(size_t)k * G + b]);
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    if(threadIdx.x == 0) out[k] = r;
  }
}

struct Workspace {
  CgVecs<double> v64;
  CgVecs<float> v32;
  PipeVecs<double> q64;
  PipeVecs<float> q32;
  TileVecs<double> t64;
  TileVecs<float> t32;
  SpVecs<double> sp64;
  SpVecs<float> sp32;
  float *spShadow[3];
  PipeParts pparts;
  PipeScalars *sc;
  SplitVecs<double> s64;
  SplitVecs<float> s32;
  CgParts parts;
  double *x64, *r64, *tmp64, *res, *part;
  CgState *st;
  IterParams *ip;
};

int ensureWorkspace(fe_hip_ctx *ctx, Workspace &w)
{
  const size_t n = ctx->n;
  const size_t need = n * (8 * 8 + 8 * 4 + 3 * 8 + 9 * 8 + 9 * 4 + 8 * 8 + 8 * 4 + 9 * 8 + 9 * 4 + 10 * 8 + 10 * 4 + 3 * 4) +
                      24 * kMaxGrid * 8 + 64 * 8 + 8192 + 64 * 1024;
  if(!ctx->d_work || ctx->workBytes < need) {
    if(ctx->d_work) feFree(ctx, ctx->d_work);
    char *p = nullptr;
    if(feAlloc(ctx, &p, need)) return FE_STATUS_OUT_OF_MEMORY;
    ctx->d_work = p;
    ctx->workBytes = need;
  }
  char *p = (char *)ctx->d_work;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  w.st = (CgState *)take(sizeof(CgState));
  w.ip = (IterParams *)take(sizeof(IterParams));
  w.res = (double *)take(64 * 8);
  w.parts.rz[0] = (double *)take(kMaxGrid * 8);
  w.parts.rz[1] = (double *)take(kMaxGrid * 8);
  w.parts.rr[0] = (double *)take(kMaxGrid * 8);
  w.parts.rr[1] = (double *)take(kMaxGrid * 8);
  w.parts.pq = (double *)take(kMaxGrid * 8);
  w.part = (double *)take(4 * kGrid * 8);
  w.sc = (PipeScalars *)take(sizeof(PipeScalars));
  for(int k = 0; k < 2; ++k) {
    w.pparts.g[k] = (double *)take(kMaxGrid * 8);
    w.pparts.d[k] = (double *)take(kMaxGrid * 8);
    w.pparts.rho[k] = (double *)take(kMaxGrid * 8);
  }
  w.q64.x = (double *)take(n * 8);
  w.q64.p = (double *)take(n * 8);
  w.q64.dinv = (double *)take(n * 8);
  w.q32.x = (float *)take(n * 4);
  w.q32.p = (float *)take(n * 4);
  w.q32.dinv = (float *)take(n * 4);
  for(int k = 0; k < 2; ++k) {
    w.q64.r[k] = (double *)take(n * 8);
    w.q64.w[k] = (double *)take(n * 8);
    w.q64.s[k] = (double *)take(n * 8);
    w.q32.r[k] = (float *)take(n * 4);
    w.q32.w[k] = (float *)take(n * 4);
    w.q32.s[k] = (float *)take(n * 4);
  }
  {
    double **a[] = {&w.t64.x, &w.t64.r, &w.t64.w, &w.t64.s, &w.t64.z, &w.t64.p, &w.t64.dinv, &w.t64.m[0], &w.t64.m[1]};
    for(auto pp : a) *pp = (double *)take(n * 8);
    float **b[] = {&w.t32.x, &w.t32.r, &w.t32.w, &w.t32.s, &w.t32.z, &w.t32.p, &w.t32.dinv, &w.t32.m[0], &w.t32.m[1]};
    for(auto pp : b) *pp = (float *)take(n * 4);
  }
  {
    double **a[] = {&w.sp64.x, &w.sp64.r, &w.sp64.u, &w.sp64.w, &w.sp64.z,
                     &w.sp64.q, &w.sp64.s, &w.sp64.p, &w.sp64.m, &w.sp64.t};
    for(auto pp : a) *pp = (double *)take(n * 8);
    w.spShadow[0] = (float *)take(n * 4);
    w.spShadow[1] = (float *)take(n * 4);
    w.spShadow[2] = (float *)take(n * 4);
    w.sp64.w32 = w.sp64.t32 = w.sp64.m32 = nullptr;
    w.sp32.w32 = w.sp32.t32 = w.sp32.m32 = nullptr;
    float **b[] = {&w.sp32.x, &w.sp32.r, &w.sp32.u, &w.sp32.w, &w.sp32.z,
                   &w.sp32.q, &w.sp32.s, &w.sp32.p, &w.sp32.m, &w.sp32.t};
    for(auto pp : b) *pp = (float *)take(n * 4);
  }
  double **d64[] = {&w.v64.x, &w.v64.r, &w.v64.z, &w.v64.p[0], &w.v64.p[1], &w.v64.q, &w.v64.dinv};
  for(auto pp : d64) *pp = (double *)take(n * 8);
  float **d32[] = {&w.v32.x, &w.v32.r, &w.v32.z, &w.v32.p[0], &w.v32.p[1], &w.v32.q, &w.v32.dinv};
  for(auto pp : d32) *pp = (float *)take(n * 4);
  {
    double **a[] = {&w.s64.x, &w.s64.r[0], &w.s64.r[1], &w.s64.p[0], &w.s64.p[1], &w.s64.q, &w.s64.y, &w.s64.z};
    for(auto pp : a) *pp = (double *)take(n * 8);
    float **b[] = {&w.s32.x, &w.s32.r[0], &w.s32.r[1], &w.s32.p[0], &w.s32.p[1], &w.s32.q, &w.s32.y, &w.s32.z};
    for(auto pp : b) *pp = (float *)take(n * 4);
  }
  w.x64 = (double *)take(n * 8);
  w.r64 = (double *)take(n * 8);
  w.tmp64 = (double *)take(n * 8);
  if(!ctx->h_state) FE_HIP_TRY(ctx, hipHostMalloc(&ctx->h_state, 4096));
  // second half of the pinned block: the iteration parameters of graph replays
  static_assert(sizeof(CgState) <= 2048, "pinned state block layout");
  ctx->h_iterParams = (char *)ctx->h_state + 2048;
  return FE_STATUS_OK;
}

template <typename T, int TB>
void launchSpmvTB(fe_hip_ctx *ctx, Workspace &w, const T *val, const T *z, const T *pOld, T *pNew, T *q, int it,
                  int maxit, double tol2)
{
  const int n = ctx->n, G = ctx->iterGrid;
#define FE_SPMV_CASE(L, U)                                                                                    \
  hipLaunchKernelGGL((cgSpmv<T, L, TB, U>), dim3(G), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, val, \
                     z, pOld, pNew, q, w.parts, w.st, it, maxit, tol2)
  const int L = ctx->spmvLanes, U = ctx->spmvUnroll;
  if(U == 2) {
    switch(L) {
      case 1: FE_SPMV_CASE(1, 2); break;
      case 2: FE_SPMV_CASE(2, 2); break;
      case 4: FE_SPMV_CASE(4, 2); break;
      case 8: FE_SPMV_CASE(8, 2); break;
      case 16: FE_SPMV_CASE(16, 2); break;
      default: FE_SPMV_CASE(32, 2); break;
    }
  } else {
    switch(L) {
      case 1: FE_SPMV_CASE(1, 1); break;
      case 2: FE_SPMV_CASE(2, 1); break;
      case 4: FE_SPMV_CASE(4, 1); break;
      case 8: FE_SPMV_CASE(8, 1); break;
      case 16: FE_SPMV_CASE(16, 1); break;
      default: FE_SPMV_CASE(32, 1); break;
    }
  }
#undef FE_SPMV_CASE
}

template <typename T>
void launchSpmv(fe_hip_ctx *ctx, Workspace &w, const T *val, const T *z, const T *pOld, T *pNew, T *q, int it,
                int maxit, double tol2)
{
  if(ctx->iterBlock == 1024) launchSpmvTB<T, 1024>(ctx, w, val, z, pOld, pNew, q, it, maxit, tol2);
  else if(ctx->iterBlock == 512) launchSpmvTB<T, 512>(ctx, w, val, z, pOld, pNew, q, it, maxit, tol2);
  else launchSpmvTB<T, 256>(ctx, w, val, z, pOld, pNew, q, it, maxit, tol2);
}

template <typename T>
void launchUpdate(fe_hip_ctx *ctx, Workspace &w, CgVecs<T> &v, const T *p, int it)
{
  const int G = ctx->iterGrid;
  if(ctx->iterBlock == 1024)
    hipLaunchKernelGGL((cgUpdate<T, 1024>), dim3(G), dim3(1024), 0, ctx->stream, ctx->n, v, p, w.parts, w.st, it);
  else if(ctx->iterBlock == 512)
    hipLaunchKernelGGL((cgUpdate<T, 512>), dim3(G), dim3(512), 0, ctx->stream, ctx->n, v, p, w.parts, w.st, it);
  else
    hipLaunchKernelGGL((cgUpdate<T, 256>), dim3(G), dim3(256), 0, ctx->stream, ctx->n, v, p, w.parts, w.st, it);
}

// r = b - A x in FP64; returns ||r||_2^2 (synchronous).
int residualNorm2(fe_hip_ctx *ctx, Workspace &w, const double *x, const double *b, double *r, double *out)
{
#define FE_RES_CASE(L)                                                                                       \
  case L:                                                                                                    \
    hipLaunchKernelGGL((residual64<L>), dim3(kGrid), dim3(kTB), 0, ctx->stream, ctx->n, ctx->d_ia, ctx->d_ja, \
                       ctx->d_val, x, b, r, w.part);                                                         \
    break;
  switch(ctx->spmvLanes) {
    FE_RES_CASE(1)
    FE_RES_CASE(2)
    FE_RES_CASE(4)
    FE_RES_CASE(8)
    FE_RES_CASE(16)
    FE_RES_CASE(32)
    default: FE_RES_CASE(64)
  }
#undef FE_RES_CASE
  hipLaunchKernelGGL(finalizeSums, dim3(1), dim3(kTB), 0, ctx->stream, w.part, kGrid, 1, w.res);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipMemcpyAsync(out, w.res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

// Launches CG iterations [0, maxit) in chunks of checkEvery, polling the state
// after each chunk.  *iters: iterations completed, *done: stop code.
template <typename T>
int runCg(fe_hip_ctx *ctx, Workspace &w, CgVecs<T> &v, const T *val, int checkEvery, int maxit, double tol2,
          int *iters, int *done)
{
  CgState *hs = (CgState *)ctx->h_state;
  FeKernelTimer &tm = ctx->spmvTimer;
  tm.used = 0;
  int it = 0;
  for(;;) {
    for(int k = 0; k < checkEvery; ++k, ++it) {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if(ctx->profiling) {
        if(tm.used + 2 > tm.ev.size())
          for(int a = 0; a < 64; ++a) {
            hipEvent_t e;
            FE_HIP_TRY(ctx, hipEventCreate(&e));
            tm.ev.push_back(e);
          }
        e0 = tm.ev[tm.used++];
        e1 = tm.ev[tm.used++];
        FE_HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
      }
      const int par = it & 1;
      launchSpmv<T>(ctx, w, val, v.z, v.p[par], v.p[par ^ 1], v.q, it, maxit, tol2);
      if(ctx->profiling) FE_HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
      launchUpdate<T>(ctx, w, v, v.p[par ^ 1], it);
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(hs->done) break;
  }
  if(ctx->profiling) {
    // launches after the stop did no work: count the first hs->iter ones
    const size_t valid = std::min<size_t>(tm.used / 2, (size_t)std::max(0, hs->iter));
    for(size_t pi = 0; pi < valid; ++pi) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, tm.ev[2 * pi], tm.ev[2 * pi + 1]));
      tm.total_ms += ms;
      tm.launches++;
    }
    tm.used = 0;
  }
  *iters = hs->iter;
  *done = hs->done;
  return FE_STATUS_OK;
}

template <typename T>
int cgStart(fe_hip_ctx *ctx, Workspace &w, CgVecs<T> &v, const double *b, int fromResidual)
{
  const int G = ctx->iterGrid;
  if(ctx->iterBlock == 1024)
    hipLaunchKernelGGL((cgInit<T, 1024>), dim3(G), dim3(1024), 0, ctx->stream, ctx->n, b, v, w.parts, w.st, fromResidual);
  else if(ctx->iterBlock == 512)
    hipLaunchKernelGGL((cgInit<T, 512>), dim3(G), dim3(512), 0, ctx->stream, ctx->n, b, v, w.parts, w.st, fromResidual);
  else
    hipLaunchKernelGGL((cgInit<T, 256>), dim3(G), dim3(256), 0, ctx->stream, ctx->n, b, v, w.parts, w.st, fromResidual);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}


template <typename T, int TB>
void launchPipeTB(fe_hip_ctx *ctx, Workspace &w, PipeVecs<T> &v, const T *val, int it, int maxit, double tol2,
                  bool init, const double *b, int fromResidual)
{
  const int n = ctx->n, G = ctx->iterGrid;
#define FE_PIPE_CASE(L)                                                                                           \
  if(init)                                                                                                        \
    hipLaunchKernelGGL((pipeInit<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, val, b, v, \
                       w.pparts, w.sc, w.st, fromResidual);                                                       \
  else                                                                                                            \
    hipLaunchKernelGGL((pipeIter<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, val, v,    \
                       w.pparts, w.sc, w.st, it, maxit, tol2);
  switch(ctx->spmvLanes) {
    case 1: FE_PIPE_CASE(1) break;
    case 2: FE_PIPE_CASE(2) break;
    case 4: FE_PIPE_CASE(4) break;
    case 8: FE_PIPE_CASE(8) break;
    case 16: FE_PIPE_CASE(16) break;
    default: FE_PIPE_CASE(32) break;
  }
#undef FE_PIPE_CASE
}

template <typename T>
void launchPipe(fe_hip_ctx *ctx, Workspace &w, PipeVecs<T> &v, const T *val, int it, int maxit, double tol2,
                bool init = false, const double *b = nullptr, int fromResidual = 0)
{
  if(ctx->iterBlock == 1024) launchPipeTB<T, 1024>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
  else if(ctx->iterBlock == 512) launchPipeTB<T, 512>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
  else launchPipeTB<T, 256>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
}

// Pipelined Jacobi-PCG iterations until the stop test fires; profiling times
// every iteration kernel (the dominant kernel of the solve).
template <typename T>
int runPipe(fe_hip_ctx *ctx, Workspace &w, PipeVecs<T> &v, const T *val, int checkEvery, int maxit, double tol2,
            int *iters, int *done)
{
  CgState *hs = (CgState *)ctx->h_state;
  FeKernelTimer &tm = ctx->spmvTimer;
  tm.used = 0;
  int it = 0, batch = checkEvery;
  for(;;) {
    for(int k = 0; k < batch; ++k, ++it) {
      if(ctx->profiling) {
        if(tm.used + 2 > tm.ev.size())
          for(int a = 0; a < 64; ++a) {
            hipEvent_t e;
            FE_HIP_TRY(ctx, hipEventCreate(&e));
            tm.ev.push_back(e);
          }
        FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
      }
      launchPipe<T>(ctx, w, v, val, it, maxit, tol2);
      if(ctx->profiling) FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(hs->done) break;
    batch = feNextBatch(hs, it, batch, checkEvery, tol2);
  }
  if(ctx->profiling) {
    const size_t valid = std::min<size_t>(tm.used / 2, (size_t)std::max(0, hs->iter));
    for(size_t pi = 0; pi < valid; ++pi) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, tm.ev[2 * pi], tm.ev[2 * pi + 1]));
      tm.total_ms += ms;
      tm.launches++;
    }
    tm.used = 0;
  }
  *iters = hs->iter;
  *done = hs->done;
  return FE_STATUS_OK;
}


template <typename T, int TB>
void launchTileTB(fe_hip_ctx *ctx, Workspace &w, TileVecs<T> &v, const T *val, int it, int maxit, double tol2,
                  bool init, const double *b, int fromResidual)
{
  const int n = ctx->n;
  const int nTiles = (n + TB - 1) / TB;
  const int G = feIterGrid(ctx, nTiles);
#define FE_TILE_CASE(L)                                                                                            \
  if(init)                                                                                                         \
    hipLaunchKernelGGL((tileInit<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, val, b, v, \
                       w.pparts, w.sc, w.st, fromResidual);                                                        \
  else                                                                                                             \
    hipLaunchKernelGGL((tileIter<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, val, v,    \
                       w.pparts, w.sc, w.st, it, maxit, tol2);
  switch(ctx->spmvLanes) {
    case 1: FE_TILE_CASE(1) break;
    case 2: FE_TILE_CASE(2) break;
    case 4: FE_TILE_CASE(4) break;
    case 8: FE_TILE_CASE(8) break;
    default: FE_TILE_CASE(16) break;
  }
#undef FE_TILE_CASE
}

template <typename T>
void launchTile(fe_hip_ctx *ctx, Workspace &w, TileVecs<T> &v, const T *val, int it, int maxit, double tol2,
                bool init = false, const double *b = nullptr, int fromResidual = 0)
{
  if(ctx->iterBlock == 1024) launchTileTB<T, 1024>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
  else if(ctx->iterBlock == 512) launchTileTB<T, 512>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
  else launchTileTB<T, 256>(ctx, w, v, val, it, maxit, tol2, init, b, fromResidual);
}

// Pipelined split-preconditioned PCG from x = 0 on A x = b (b in FP64) with the
// factor G (lower CSR gL, upper CSR gU of G^T) and dinv; every iteration is
// K1, K2, K3 (the A product K3 is the timed kernel).
template <typename T>
int runSplitPipe(fe_hip_ctx *ctx, Workspace &w, SpVecs<T> &v, const T *val, const T *gL, const T *gU,
                 const T *dinv, const double *b, int checkEvery, int maxit, double tol2, int *iters, int *done,
                 int firstBatch = 0)
{
  const int n = ctx->n;
  const FePrecond &P = ctx->pc;
  const int TB = ctx->iterBlock;
  const int G = feIterGrid(ctx, (n + TB - 1) / TB);
  const int L = ctx->spmvLanes, LF = std::max(1, L / 2);
  CgState *hs = (CgState *)ctx->h_state;
  FeKernelTimer &tm = ctx->spmvTimer;
  tm.used = 0;
#define FE_SP_LAUNCH_T(KER, TT, LL, ...)                                                                                 \
  do {                                                                                                             \
    if(TB == 1024) {                                                                                               \
      if(LL == 1) hipLaunchKernelGGL((KER<TT, 1, 1024>), dim3(G), dim3(1024), 0, ctx->stream, __VA_ARGS__);          \
      else if(LL == 2) hipLaunchKernelGGL((KER<TT, 2, 1024>), dim3(G), dim3(1024), 0, ctx->stream, __VA_ARGS__);     \
      else if(LL == 4) hipLaunchKernelGGL((KER<TT, 4, 1024>), dim3(G), dim3(1024), 0, ctx->stream, __VA_ARGS__);     \
      else hipLaunchKernelGGL((KER<TT, 8, 1024>), dim3(G), dim3(1024), 0, ctx->stream, __VA_ARGS__);                 \
    } else if(TB == 512) {                                                                                         \
      if(LL == 1) hipLaunchKernelGGL((KER<TT, 1, 512>), dim3(G), dim3(512), 0, ctx->stream, __VA_ARGS__);            \
      else if(LL == 2) hipLaunchKernelGGL((KER<TT, 2, 512>), dim3(G), dim3(512), 0, ctx->stream, __VA_ARGS__);       \
      else if(LL == 4) hipLaunchKernelGGL((KER<TT, 4, 512>), dim3(G), dim3(512), 0, ctx->stream, __VA_ARGS__);       \
      else hipLaunchKernelGGL((KER<TT, 8, 512>), dim3(G), dim3(512), 0, ctx->stream, __VA_ARGS__);                   \
    } else {                                                                                                       \
      if(LL == 1) hipLaunchKernelGGL((KER<TT, 1, 256>), dim3(G), dim3(256), 0, ctx->stream, __VA_ARGS__);            \
      else if(LL == 2) hipLaunchKernelGGL((KER<TT, 2, 256>), dim3(G), dim3(256), 0, ctx->stream, __VA_ARGS__);       \
      else if(LL == 4) hipLaunchKernelGGL((KER<TT, 4, 256>), dim3(G), dim3(256), 0, ctx->stream, __VA_ARGS__);       \
      else hipLaunchKernelGGL((KER<TT, 8, 256>), dim3(G), dim3(256), 0, ctx->stream, __VA_ARGS__);                   \
    }                                                                                                              \
  } while(0)
#define FE_SP_LAUNCH(KER, LL, ...) FE_SP_LAUNCH_T(KER, T, LL, __VA_ARGS__)
  hipLaunchKernelGGL(spStart<T>, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, b, v);
  const bool amg = ctx->amgActive;
  const int *dDone = &((const CgState *)w.st)->done;
  // u_0 = M^-1 r_0, w_0 = A u_0
  if(amg) {
    if(int rc = feAmgApply(ctx, (const T *)v.r, v.u, nullptr, ctx->pcValues32 != 0)) return rc;
  } else if(sizeof(T) == 8 && ctx->pcValues32 && !v.m32) {
    FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaL, P.d_jaL, (const float *)P.d_gL32, (const T *)v.r,
                 (const float *)P.d_dinv32, v.t, (const CgState *)w.st, 0);
    FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaU, P.d_jaU, (const float *)P.d_gU32, (const T *)v.t, (const float *)nullptr,
                 v.u, (const CgState *)w.st, 0);
  } else {
    FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaL, P.d_jaL, gL, (const T *)v.r, dinv, v.t, (const CgState *)w.st, 0);
    FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaU, P.d_jaU, gU, (const T *)v.t, (const T *)nullptr, v.u, (const CgState *)w.st,
                 0);
  }
  FE_SP_LAUNCH(spInitW, L, n, ctx->d_ia, ctx->d_ja, val, v, w.pparts, w.sc, w.st);
  // m = M^-1 w, skipped once the iteration has stopped
  auto applyPc = [&]() -> int {
    if(amg) return feAmgApply(ctx, (const T *)v.w, v.m, dDone, ctx->pcValues32 != 0);
    if(v.m32) {
      FE_SP_LAUNCH_T(spSpmv, float, LF, n, P.d_iaL, P.d_jaL, (const float *)P.d_gL32, (const float *)v.w32,
                     (const float *)P.d_dinv32, v.t32, (const CgState *)w.st, 1);
      FE_SP_LAUNCH_T(spSpmv, float, LF, n, P.d_iaU, P.d_jaU, (const float *)P.d_gU32, (const float *)v.t32,
                     (const float *)nullptr, v.m32, (const CgState *)w.st, 1);
    } else if(sizeof(T) == 8 && ctx->pcValues32) {
      FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaL, P.d_jaL, (const float *)P.d_gL32, (const T *)v.w,
                   (const float *)P.d_dinv32, v.t, (const CgState *)w.st, 1);
      FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaU, P.d_jaU, (const float *)P.d_gU32, (const T *)v.t, (const float *)nullptr,
                   v.m, (const CgState *)w.st, 1);
    } else {
      FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaL, P.d_jaL, gL, (const T *)v.w, dinv, v.t, (const CgState *)w.st, 1);
      FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaU, P.d_jaU, gU, (const T *)v.t, (const T *)nullptr, v.m,
                   (const CgState *)w.st, 1);
    }
    return FE_STATUS_OK;
  };
  // Graph replay of K iterations (preconditioner + A product each) when not
  // profiling: one hipGraphLaunch instead of 3 to 8 launches per iteration,
  // so short kernels are not starved by the host's launch rate.
  const int K = ctx->profiling ? 0 : ctx->graphIters;
  hipGraphExec_t gx = nullptr;
  if(K > 0) {
    IterParams *hp = (IterParams *)ctx->h_iterParams;
    hp->itBase = 0;
    hp->maxit = maxit;
    hp->tol2 = tol2;
    FE_HIP_TRY(ctx, hipMemcpyAsync(w.ip, hp, sizeof(IterParams), hipMemcpyHostToDevice, ctx->stream));
    char key[256];
    snprintf(key, sizeof(key), "sp%zu a%d v%d m%d L%d TB%d G%d K%d p%d s%p x%p", sizeof(T), (int)amg, ctx->pcValues32,
             v.m32 ? 1 : 0, L, TB, G, K, amg ? feAmgPlanFrom(ctx, sizeof(T) == 8 ? (ctx->pcValues32 ? 1 : 0) : 2) : -1,
             (void *)w.st, (void *)v.x);
    auto f = ctx->graphs.find(key);
    if(f != ctx->graphs.end()) {
      gx = (hipGraphExec_t)f->second;
    } else {
      hipGraph_t g = nullptr;
      FE_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
      int rc = FE_STATUS_OK;
      for(int k = 0; k < K && rc == FE_STATUS_OK; ++k) {
        rc = applyPc();
        FE_SP_LAUNCH(spIter, L, n, ctx->d_ia, ctx->d_ja, val, v, w.pparts, w.sc, w.st, k, maxit, tol2,
                     (const IterParams *)w.ip);
      }
      hipLaunchKernelGGL(iterAdvance, dim3(1), dim3(64), 0, ctx->stream, w.ip, K);
      const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
      if(rc != FE_STATUS_OK) return rc;
      FE_HIP_TRY(ctx, ec);
      FE_HIP_TRY(ctx, hipGraphInstantiate(&gx, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      ctx->graphs[key] = (void *)gx;
    }
  }
  int it = 0, batch = firstBatch > 0 ? std::min(firstBatch, checkEvery) : checkEvery;
  for(;;) {
    if(gx) {
      const int reps = (batch + K - 1) / K;
      for(int r = 0; r < reps; ++r) FE_HIP_TRY(ctx, hipGraphLaunch(gx, ctx->stream));
      it += reps * K;
    }
    for(int k = 0; !gx && k < batch; ++k, ++it) {
      if(amg) {
        if(int rc = feAmgApply(ctx, (const T *)v.w, v.m, dDone, ctx->pcValues32 != 0)) return rc;
      } else if(v.m32) {
        // preconditioner applied in FP32 (FP32 factor, FP32 shadows of w, t, m)
        FE_SP_LAUNCH_T(spSpmv, float, LF, n, P.d_iaL, P.d_jaL, (const float *)P.d_gL32, (const float *)v.w32,
                       (const float *)P.d_dinv32, v.t32, (const CgState *)w.st, 1);
        FE_SP_LAUNCH_T(spSpmv, float, LF, n, P.d_iaU, P.d_jaU, (const float *)P.d_gU32, (const float *)v.t32,
                       (const float *)nullptr, v.m32, (const CgState *)w.st, 1);
      } else if(sizeof(T) == 8 && ctx->pcValues32) {
        // FP64 vectors, FP32 factor values: an exactly applied SPD operator
        FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaL, P.d_jaL, (const float *)P.d_gL32, (const T *)v.w,
                     (const float *)P.d_dinv32, v.t, (const CgState *)w.st, 1);
        FE_SP_LAUNCH(spSpmvF32, LF, n, P.d_iaU, P.d_jaU, (const float *)P.d_gU32, (const T *)v.t,
                     (const float *)nullptr, v.m, (const CgState *)w.st, 1);
      } else {
        FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaL, P.d_jaL, gL, (const T *)v.w, dinv, v.t, (const CgState *)w.st, 1);
        FE_SP_LAUNCH(spSpmv, LF, n, P.d_iaU, P.d_jaU, gU, (const T *)v.t, (const T *)nullptr, v.m,
                     (const CgState *)w.st, 1);
      }
      if(ctx->profiling) {
        if(tm.used + 2 > tm.ev.size())
          for(int a = 0; a < 64; ++a) {
            hipEvent_t e;
            FE_HIP_TRY(ctx, hipEventCreate(&e));
            tm.ev.push_back(e);
          }
        FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
      }
      FE_SP_LAUNCH(spIter, L, n, ctx->d_ia, ctx->d_ja, val, v, w.pparts, w.sc, w.st, it, maxit, tol2,
                   (const IterParams *)nullptr);
      if(ctx->profiling) FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(hs->done) break;
    batch = feNextBatch(hs, it, batch, checkEvery, tol2);
  }
#undef FE_SP_LAUNCH
#undef FE_SP_LAUNCH_T
  if(ctx->profiling) {
    const size_t valid = std::min<size_t>(tm.used / 2, (size_t)std::max(0, hs->iter));
    for(size_t pi = 0; pi < valid; ++pi) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, tm.ev[2 * pi], tm.ev[2 * pi + 1]));
      tm.total_ms += ms;
      tm.launches++;
    }
    tm.used = 0;
  }
  *iters = hs->iter;
  *done = hs->done;
  return FE_STATUS_OK;
}

// Jacobi-PCG iterations (pipelined kernel chosen by the pipe_kernel tuning
// key) until the stop test fires.
template <typename T>
int runJacobi(fe_hip_ctx *ctx, bool tiled, Workspace &w, PipeVecs<T> &pv, TileVecs<T> &tv, const T *val,
              int checkEvery, int maxit, double tol2, int *iters, int *done, int firstBatch = 0)
{
  if(!tiled) return runPipe<T>(ctx, w, pv, val, checkEvery, maxit, tol2, iters, done);
  CgState *hs = (CgState *)ctx->h_state;
  FeKernelTimer &tm = ctx->spmvTimer;
  tm.used = 0;
  int it = 0, batch = firstBatch > 0 ? std::min(firstBatch, checkEvery) : checkEvery;
  for(;;) {
    for(int k = 0; k < batch; ++k, ++it) {
      if(ctx->profiling) {
        if(tm.used + 2 > tm.ev.size())
          for(int a = 0; a < 64; ++a) {
            hipEvent_t e;
            FE_HIP_TRY(ctx, hipEventCreate(&e));
            tm.ev.push_back(e);
          }
        FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
      }
      launchTile<T>(ctx, w, tv, val, it, maxit, tol2);
      if(ctx->profiling) FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(hs->done) break;
    batch = feNextBatch(hs, it, batch, checkEvery, tol2);
  }
  if(ctx->profiling) {
    const size_t valid = std::min<size_t>(tm.used / 2, (size_t)std::max(0, hs->iter));
    for(size_t pi = 0; pi < valid; ++pi) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, tm.ev[2 * pi], tm.ev[2 * pi + 1]));
      tm.total_ms += ms;
      tm.launches++;
    }
    tm.used = 0;
  }
  *iters = hs->iter;
  *done = hs->done;
  return FE_STATUS_OK;
}

template <typename T>
void startJacobi(fe_hip_ctx *ctx, bool tiled, Workspace &w, PipeVecs<T> &pv, TileVecs<T> &tv, const T *val,
                 const double *b, int fromResidual)
{
  if(tiled) launchTile<T>(ctx, w, tv, val, 0, 0, 0., true, b, fromResidual);
  else launchPipe<T>(ctx, w, pv, val, 0, 0, 0., true, b, fromResidual);
}

template <typename T, int TB>
void launchSplitTB(fe_hip_ctx *ctx, Workspace &w, SplitVecs<T> &v, int kind, int it, const double *b, int rs, int slot,
                   int checkDone)
{
  const int n = ctx->n, G = ctx->iterGrid;
#define FE_SPLIT_CASE(L)                                                                                          \
  if(kind == 0)                                                                                                   \
    hipLaunchKernelGGL((splitInit<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, b, v, w.parts, w.st);          \
  else if(kind == 1)                                                                                              \
    hipLaunchKernelGGL((splitApplyG<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, v, w.parts, w.st, it);       \
  else                                                                                                            \
    hipLaunchKernelGGL((splitApplyGT<T, L, TB>), dim3(G), dim3(TB), 0, ctx->stream, n, v, w.parts, w.st, rs, slot, \
                       checkDone);
  // the factors have about half the row length of A
  const int L = std::max(1, ctx->spmvLanes / 2);
  switch(L) {
    case 1: FE_SPLIT_CASE(1) break;
    case 2: FE_SPLIT_CASE(2) break;
    case 4: FE_SPLIT_CASE(4) break;
    case 8: FE_SPLIT_CASE(8) break;
    default: FE_SPLIT_CASE(16) break;
  }
#undef FE_SPLIT_CASE
}

template <typename T>
void launchSplit(fe_hip_ctx *ctx, Workspace &w, SplitVecs<T> &v, int kind, int it, const double *b = nullptr,
                 int rs = 0, int slot = 0, int checkDone = 1)
{
  if(ctx->iterBlock == 1024) launchSplitTB<T, 1024>(ctx, w, v, kind, it, b, rs, slot, checkDone);
  else if(ctx->iterBlock == 512) launchSplitTB<T, 512>(ctx, w, v, kind, it, b, rs, slot, checkDone);
  else launchSplitTB<T, 256>(ctx, w, v, kind, it, b, rs, slot, checkDone);
}

// Split-preconditioned PCG from x = 0 on A x = b (b in FP64).
template <typename T>
int runSplit(fe_hip_ctx *ctx, Workspace &w, SplitVecs<T> &v, const T *val, const double *b, int checkEvery,
             int maxit, double tol2, int *iters, int *done)
{
  CgState *hs = (CgState *)ctx->h_state;
  FeKernelTimer &tm = ctx->spmvTimer;
  tm.used = 0;
  launchSplit<T>(ctx, w, v, 0, 0, b);
  launchSplit<T>(ctx, w, v, 2, 0, nullptr, 0, 0, 0);
  int it = 0;
  for(;;) {
    for(int k = 0; k < checkEvery; ++k, ++it) {
      const int c = it & 1;
      if(ctx->profiling) {
        if(tm.used + 2 > tm.ev.size())
          for(int a = 0; a < 64; ++a) {
            hipEvent_t e;
            FE_HIP_TRY(ctx, hipEventCreate(&e));
            tm.ev.push_back(e);
          }
        FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
      }
      launchSpmv<T>(ctx, w, val, v.z, v.p[c], v.p[c ^ 1], v.q, it, maxit, tol2);
      if(ctx->profiling) FE_HIP_TRY(ctx, hipEventRecord(tm.ev[tm.used++], ctx->stream));
      launchSplit<T>(ctx, w, v, 1, it);
      launchSplit<T>(ctx, w, v, 2, it, nullptr, c ^ 1, c ^ 1, 1);
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, w.st, sizeof(CgState), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if(hs->done) break;
  }
  if(ctx->profiling) {
    const size_t valid = std::min<size_t>(tm.used / 2, (size_t)std::max(0, hs->iter));
    for(size_t pi = 0; pi < valid; ++pi) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, tm.ev[2 * pi], tm.ev[2 * pi + 1]));
      tm.total_ms += ms;
      tm.launches++;
    }
    tm.used = 0;
  }
  *iters = hs->iter;
  *done = hs->done;
  return FE_STATUS_OK;
}

template <typename T>
__global__ void castResidual(int n, const double *__restrict__ a, T *__restrict__ b)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (T)a[i];
}

}  // namespace

int feResidualNorm2(fe_hip_ctx *ctx, const double *x, double *r, double *out)
{
  Workspace w;
  if(int rc = ensureWorkspace(ctx, w)) return rc;
  return residualNorm2(ctx, w, x, ctx->d_rhs, r, out);
}

int feSolve(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double *du, fe_hip_solve_stats *stats)
{
  if(o->ksp == FE_KSP_GMRES) return feSolveGmres(ctx, o, du, stats);
  if(o->ksp == FE_KSP_DIRECT) return feSolveDirect(ctx, o, du, stats);
  FE_REQUIRE(ctx, o->ksp == FE_KSP_CG, "unknown Krylov method");
  FE_REQUIRE(ctx, o->pc == FE_PC_JACOBI || o->pc == FE_PC_IC0 || o->pc == FE_PC_LEARNED || o->pc == FE_PC_AMG,
             "CG takes the JACOBI, IC0, LEARNED or AMG preconditioner");
  const bool split = (o->pc != FE_PC_JACOBI);
  const bool amg = (o->pc == FE_PC_AMG);
  ctx->amgActive = amg;
  // pipelined form per precision: the tiled Ghysels-Vanroose iteration (one
  // gathered vector) for FP64 solves, Chronopoulos-Gear (w = A u every
  // iteration, fewer refinement restarts) for FP32 inner solves
  const bool tiled64 = ctx->pipeKernel == 2 ? true : ctx->pipeKernel == 1;
  const bool tiled32 = ctx->pipeKernel == 2 ? false : ctx->pipeKernel == 1;
  Workspace w;
  if(int rc = ensureWorkspace(ctx, w)) return rc;
  const int n = ctx->n;
  const int checkEvery = std::max(1, o->check_every);
  hipEvent_t t0, t1;
  FE_HIP_TRY(ctx, hipEventCreate(&t0));
  FE_HIP_TRY(ctx, hipEventCreate(&t1));
  FE_HIP_TRY(ctx, hipEventRecord(t0, ctx->stream));
  ctx->spmvTimer.total_ms = 0.;
  ctx->spmvTimer.launches = 0;
  ctx->iterTimer.total_ms = 0.;
  ctx->iterTimer.launches = 0;

  hipLaunchKernelGGL(sumSquares, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, ctx->d_rhs, w.part);
  hipLaunchKernelGGL(finalizeSums, dim3(1), dim3(kTB), 0, ctx->stream, w.part, kGrid, 1, w.res + 8);
  FE_HIP_TRY(ctx, hipGetLastError());
  double bb = 0.;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&bb, w.res + 8, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const double bn = std::sqrt(bb);
  const double target = std::max(o->rtol * bn, o->atol);
  int iters = 0, outer = 0, converged = 0, reason = 0, done = 0, breakdowns = 0;
  double relres = bn > 0. ? 1. : 0.;

  double setupMs = 0.;
  if(split && bn != 0.) {
    hipEvent_t s0, s1;
    FE_HIP_TRY(ctx, hipEventCreate(&s0));
    FE_HIP_TRY(ctx, hipEventCreate(&s1));
    FE_HIP_TRY(ctx, hipEventRecord(s0, ctx->stream));
    if(amg) {
      if(int rc = feAmgSetup(ctx)) return rc;
      // the nested-ordered PCG runs its own numeric set-up inside the solve
      if(!feAmgNestedUsable(ctx))
        if(int rc = feAmgNumeric(ctx)) return rc;
    } else {
      if(int rc = fePrecondFactor(ctx, o->pc == FE_PC_LEARNED ? kLearnedIcSweeps : ctx->icSweeps)) return rc;
      if(o->pc == FE_PC_LEARNED)
        if(int rc = feLearnedCorrect(ctx)) return rc;
    }
    FE_HIP_TRY(ctx, hipEventRecord(s1, ctx->stream));
    FE_HIP_TRY(ctx, hipEventSynchronize(s1));
    float sm = 0.f;
    FE_HIP_TRY(ctx, hipEventElapsedTime(&sm, s0, s1));
    setupMs = sm;
    (void)hipEventDestroy(s0);
    (void)hipEventDestroy(s1);
    FePrecond &P = ctx->pc;
    w.s64.gL = P.d_gL64; w.s64.gU = P.d_gU64; w.s64.dinv = P.d_dinv64;
    w.s32.gL = P.d_gL32; w.s32.gU = P.d_gU32; w.s32.dinv = P.d_dinv32;
    w.s64.iaL = w.s32.iaL = P.d_iaL; w.s64.jaL = w.s32.jaL = P.d_jaL;
    w.s64.iaU = w.s32.iaU = P.d_iaU; w.s64.jaU = w.s32.jaU = P.d_jaU;
  }

  const bool nested = amg && feAmgNestedUsable(ctx);
  if(bn == 0.) {
    FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_du, 0, sizeof(double) * n, ctx->stream));
    converged = 1;
    reason = 2;
    relres = 0.;
  } else if(nested) {
    // multigrid PCG over the nested ordering, FP32 or FP64 inner solves under
    // FP64 refinement (fe_amg_pcg.hip)
    FE_HIP_TRY(ctx, hipMemsetAsync(w.x64, 0, sizeof(double) * n, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(w.r64, ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    FeAmgPcgResult R;
    if(int rc = feAmgPcgSolve(ctx, o, bn, target, w.x64, w.r64, &R)) return rc;
    iters = R.iters;
    outer = R.outer;
    converged = R.converged;
    reason = R.reason;
    relres = R.relres;
    FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, w.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  } else if(split) {
    // split-preconditioned PCG: FP64 directly, or FP32 inner + FP64 refinement
    const bool fp32 = (o->precision == FE_PREC_FP32_IR);
    FE_REQUIRE(ctx, !fp32 || ctx->d_val32, "assemble with FE_PREC_FP32_IR to create the FP32 values");
    FE_HIP_TRY(ctx, hipMemsetAsync(w.x64, 0, sizeof(double) * n, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(w.r64, ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    double rn = bn;
    bool inner32 = fp32;
    int lastInner = 0;  // iterations of the previous inner solve: first check batch of the next
    while(outer < 60) {
      relres = rn / bn;
      if(rn <= target) { converged = 1; reason = 2; break; }
      if(iters >= o->max_iter) { reason = -3; break; }
      double tol2 = inner32 ? (o->inner_rtol * rn) * (o->inner_rtol * rn) : target * target * 0.25;
      tol2 = std::max(tol2, target * target * 0.01);
      int it0 = 0;
      FE_HIP_TRY(ctx, hipMemcpyAsync(w.tmp64, w.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
      const FePrecond &Pc = ctx->pc;
      if(inner32) {
        if(tiled32 || amg) {
          if(int rc = runSplitPipe<float>(ctx, w, w.sp32, ctx->d_val32, Pc.d_gL32, Pc.d_gU32, Pc.d_dinv32, w.r64,
                                          checkEvery, o->max_iter - iters, tol2, &it0, &done, lastInner > 0 ? lastInner + 2 : 0))
            return rc;
        } else if(int rc = runSplit<float>(ctx, w, w.s32, ctx->d_val32, w.r64, checkEvery, o->max_iter - iters, tol2,
                                           &it0, &done)) {
          return rc;
        }
        hipLaunchKernelGGL(addVec<float>, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, w.x64,
                           (tiled32 || amg) ? w.sp32.x : w.s32.x);
      } else {
        if(tiled64 || amg) {
          const bool mixed = ctx->pcFp32 != 0 && !amg;
          w.sp64.w32 = mixed ? w.spShadow[0] : nullptr;
          w.sp64.t32 = mixed ? w.spShadow[1] : nullptr;
          w.sp64.m32 = mixed ? w.spShadow[2] : nullptr;
          if(int rc = runSplitPipe<double>(ctx, w, w.sp64, ctx->d_val, Pc.d_gL64, Pc.d_gU64, Pc.d_dinv64, w.r64,
                                           checkEvery, o->max_iter - iters, tol2, &it0, &done, lastInner > 0 ? lastInner + 2 : 0))
            return rc;
        } else if(int rc = runSplit<double>(ctx, w, w.s64, ctx->d_val, w.r64, checkEvery, o->max_iter - iters, tol2,
                                            &it0, &done)) {
          return rc;
        }
        hipLaunchKernelGGL(addVec<double>, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, w.x64,
                           (tiled64 || amg) ? w.sp64.x : w.s64.x);
      }
      iters += it0;
      lastInner = it0;
      double rr = 0.;
      if(int rc = residualNorm2(ctx, w, w.x64, ctx->d_rhs, w.r64, &rr)) return rc;
      outer++;
      if(inner32 && !(std::sqrt(rr) < 0.9 * rn)) {
        // the FP32 correction did not reduce the FP64 residual: undo it and
        // finish with FP64 inner solves
        FE_HIP_TRY(ctx, hipMemcpyAsync(w.x64, w.tmp64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        if(int rc = residualNorm2(ctx, w, w.x64, ctx->d_rhs, w.r64, &rr)) return rc;
        inner32 = false;
        continue;
      }
      rn = std::sqrt(rr);
      // a breakdown of the recurrences ends the inner solve; the next outer
      // iteration restarts it from the FP64 residual
      if(done == 3 && ++breakdowns > 20) { reason = -3; break; }
      if(done == 4 || !(rn == rn)) { reason = -9; break; }
      if(rn / bn > o->dtol) { reason = -4; break; }
    }
    relres = rn / bn;
    if(!converged && rn <= target) { converged = 1; reason = 2; }
    if(amg)
      if(int rc = feAmgCheck(ctx)) return rc;
    FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, w.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  } else if(o->precision == FE_PREC_FP64) {
    double *dinv64 = tiled64 ? w.t64.dinv : w.q64.dinv;
    double *x64 = tiled64 ? w.t64.x : w.q64.x;
    double *r64 = tiled64 ? w.t64.r : w.q64.r[0];
    hipLaunchKernelGGL(extractDinv<double>, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, ctx->stream, n, ctx->d_diag,
                       ctx->d_val, dinv64);
    // recursive-residual target; the FP64 true residual is checked after
    double tol2 = target * target * 0.25;
    startJacobi<double>(ctx, tiled64, w, w.q64, w.t64, ctx->d_val, ctx->d_rhs, 0);
    for(;;) {
      int it = 0;
      if(int rc = runJacobi<double>(ctx, tiled64, w, w.q64, w.t64, ctx->d_val, checkEvery, o->max_iter - iters, tol2, &it,
                                    &done))
        return rc;
      iters += it;
      double rr = 0.;
      if(int rc = residualNorm2(ctx, w, x64, ctx->d_rhs, r64, &rr)) return rc;
      relres = std::sqrt(rr) / bn;
      outer++;
      if(std::sqrt(rr) <= target) { converged = 1; reason = 2; break; }
      if(done == 4 || !(rr == rr)) { reason = -9; break; }
      if(iters >= o->max_iter) { reason = -3; break; }
      if(relres > o->dtol) { reason = -4; break; }
      // restart from the true residual (r holds b - A x): after a breakdown of
      // the recurrences, or when the recursive residual met a target the true
      // one has not (then aim lower)
      if(done == 3) {
        if(++breakdowns > 20) { reason = -3; break; }
      } else {
        tol2 *= 0.01;
      }
      startJacobi<double>(ctx, tiled64, w, w.q64, w.t64, ctx->d_val, ctx->d_rhs, 1);
    }
    FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    FE_REQUIRE(ctx, ctx->d_val32, "assemble with FE_PREC_FP32_IR to create the FP32 values");
    hipLaunchKernelGGL(extractDinv<float>, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, ctx->stream, n, ctx->d_diag,
                       ctx->d_val, tiled32 ? w.t32.dinv : w.q32.dinv);
    FE_HIP_TRY(ctx, hipMemsetAsync(w.x64, 0, sizeof(double) * n, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(w.r64, ctx->d_rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    double rn = bn;
    const int maxOuter = 60;
    bool inner32 = true;
    while(outer < maxOuter) {
      relres = rn / bn;
      if(rn <= target) { converged = 1; reason = 2; break; }
      if(iters >= o->max_iter) { reason = -3; break; }
      double tol2 = inner32 ? (o->inner_rtol * rn) * (o->inner_rtol * rn) : target * target * 0.25;
      tol2 = std::max(tol2, target * target * 0.01);
      int it0 = 0;
      const float *d32 = tiled32 ? w.t32.x : w.q32.x;
      FE_HIP_TRY(ctx, hipMemcpyAsync(w.tmp64, w.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
      if(inner32) {
        startJacobi<float>(ctx, tiled32, w, w.q32, w.t32, ctx->d_val32, w.r64, 0);
        if(int rc = runJacobi<float>(ctx, tiled32, w, w.q32, w.t32, ctx->d_val32, checkEvery, o->max_iter - iters, tol2, &it0,
                                     &done))
          return rc;
        hipLaunchKernelGGL(addVec<float>, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, w.x64, d32);
      } else {
        // FP64 inner solve of A d = r (after an FP32 correction was rejected)
        hipLaunchKernelGGL(extractDinv<double>, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, ctx->stream, n,
                           ctx->d_diag, ctx->d_val, tiled64 ? w.t64.dinv : w.q64.dinv);
        startJacobi<double>(ctx, tiled64, w, w.q64, w.t64, ctx->d_val, w.r64, 0);
        if(int rc = runJacobi<double>(ctx, tiled64, w, w.q64, w.t64, ctx->d_val, checkEvery, o->max_iter - iters, tol2, &it0,
                                      &done))
          return rc;
        hipLaunchKernelGGL(addVec<double>, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, w.x64,
                           tiled64 ? w.t64.x : w.q64.x);
      }
      iters += it0;
      double rr = 0.;
      if(int rc = residualNorm2(ctx, w, w.x64, ctx->d_rhs, w.r64, &rr)) return rc;
      outer++;
      if(inner32 && !(std::sqrt(rr) < 0.9 * rn)) {
        // the FP32 correction did not reduce the FP64 residual: undo it and
        // finish with FP64 inner solves
        FE_HIP_TRY(ctx, hipMemcpyAsync(w.x64, w.tmp64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        if(int rc = residualNorm2(ctx, w, w.x64, ctx->d_rhs, w.r64, &rr)) return rc;
        inner32 = false;
        continue;
      }
      rn = std::sqrt(rr);
      // a breakdown of the recurrences ends the inner solve; the next outer
      // iteration restarts it from the FP64 residual
      if(done == 3 && ++breakdowns > 20) { reason = -3; break; }
      if(done == 4 || !(rn == rn)) { reason = -9; break; }
      if(rn / bn > o->dtol) { reason = -4; break; }
    }
    relres = rn / bn;
    if(!converged && rn <= target) { converged = 1; reason = 2; }
    FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_du, w.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  }
  FE_HIP_TRY(ctx, hipEventRecord(t1, ctx->stream));
  FE_HIP_TRY(ctx, hipEventSynchronize(t1));
  float ms = 0.f;
  FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);

  if(stats) {
    memset(stats, 0, sizeof(*stats));
    double dummy = 0.;
    if(int rc = residualNorm2(ctx, w, ctx->d_du, ctx->d_rhs, w.tmp64, &dummy)) return rc;
    hipLaunchKernelGGL(blockMaxAbs3, dim3(kGrid), dim3(kTB), 0, ctx->stream, n, ctx->d_du, ctx->d_rhs, w.tmp64,
                       w.part);
    hipLaunchKernelGGL(finalizeMax, dim3(1), dim3(64), 0, ctx->stream, w.part, kGrid, 3, w.res + 16);
    FE_HIP_TRY(ctx, hipGetLastError());
    double inf[3];
    FE_HIP_TRY(ctx, hipMemcpyAsync(inf, w.res + 16, sizeof(inf), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    stats->iterations = iters;
    stats->outer_iterations = outer;
    stats->converged = converged;
    stats->reason = reason;
    stats->relres_fp64 = relres;
    stats->norm_dx_inf = inf[0];
    stats->norm_rhs_inf = inf[1];
    stats->norm_axb_inf = inf[2];
    stats->ms_solve = ms;
    stats->ms_setup = setupMs;
  }
  if(du) {
    FE_HIP_TRY(ctx, hipMemcpyAsync(du, ctx->d_du, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return FE_STATUS_OK;
}
