// Multigrid-preconditioned CG with the aggregation hierarchy of FE_PC_AMG
// stored in a nested order, the role PETSc's `-ksp_type cg -pc_type gamg`
// plays behind the reference's feLinearSystemPETSc::solve
// (feLinearSystemPETSc.cpp:900-1031; only the role is alike).
//
// Ordering.  The aggregates of every level are the ones feAmgSetup builds
// (the hierarchy the oracle restates); only the numbering changes.  The
// coarsest level keeps its order; the rows of level l are ordered by the
// position of their aggregate on level l + 1, then by row.  So on every level
// each aggregate is a contiguous range of rows, aggregates follow the order
// of the coarse rows, and a restriction is a segmented sum inside a tile of
// whole aggregates: no member lists, no index indirection.  Level 0 (the
// system) gets a permuted FP32/FP64 copy of its values per solve; the FP64
// residual of the refinement stays on the original CSR (the reference's
// pattern, bit-exact).
//
// Iteration (standard PCG, x0 = 0, M = the symmetric V-cycle of fe_amg.hip:
// l1-Jacobi smoothing c D_l1^-1 before and after, over-corrected coarse
// correction), one kernel per step and every reduction folded into the next
// kernel (each block folds the previous kernel's block partials in the same
// fixed order, so every block holds the same scalar):
//   product : rho_k = (r_k, z_k); p_k = z_k + (rho_k / rho_k-1) p_k-1 (own rows
//             stored, neighbours recomputed); q_k = A p_k; partials of (p, q)
//   down0   : alpha = rho_k / (p, q); x += alpha p; r_k+1 = r_k - alpha q;
//             u = S r_k+1; b_1 = P^T (r_k+1 - A u) (tile segmented sums);
//             partials of (r, r)
//   down_l  : u = S b_l; b_l+1 = P^T (b_l - A u)          (levels 1 .. B-1)
//   block   : levels B .. coarsest in one workgroup (LDS), dense coarsest
//   up_l    : y = u + oc P e_l+1; e_l = y + S (b_l - A y)  (levels B-1 .. 1)
//   up0     : z_k+1 = y + S (r_k+1 - A y); partials of (r, z)
// The product kernel's first block also checks the residual against the
// inner tolerance and breakdowns (rho <= 0, (p, q) <= 0) and raises a flag
// the later kernels read first, so a batch of iterations stops where it
// converged.  FP32 inner solves run under FP64 iterative refinement on the
// original system.
#include "fe_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace {

constexpr int kPTB = 256;             // threads of the level kernels
constexpr int kPLPR = 4;              // lanes per row
constexpr int kPRPB = kPTB / kPLPR;   // rows per block: the tile size limit
constexpr int kPBlockRows = 1024;     // levels of at most this many rows run in the single-workgroup kernel
constexpr int kPBlockTB = 1024;
constexpr int kPBlockLds = 150 * 1024;
constexpr int kPDense = 64;           // the coarsest level is inverted densely up to this size
constexpr int kPBlockLev = 8;

template <typename T, typename TV>
struct PLev {
  int n, nc, nTiles;
  const int *ia, *ja;
  const TV *val, *l1;
  const int *agg;       // row -> aggregate on the next level (l < coarsest)
  const int *aggStart;  // nc + 1: first row of each aggregate
  const int *tileStart, *tileAgg;  // nTiles + 1: tile rows and first aggregate
  T *b, *u, *e;         // levels >= 1
};

// LDS plan of the single-workgroup kernel: levels from .. coarsest - 1 with
// their CSR (16-bit columns), values, l1, aggregates, and vectors b, u, y, e;
// the coarsest right-hand side and solution.
struct PBlockPlan {
  int from = 0, coarsest = 0, bytes = 0;
  int n[kPBlockLev + 1], nc[kPBlockLev], nnz[kPBlockLev];
  int oIa[kPBlockLev], oJa[kPBlockLev], oVal[kPBlockLev], oL1[kPBlockLev], oAgg[kPBlockLev], oAs[kPBlockLev];
  int oB[kPBlockLev + 1], oU[kPBlockLev], oY[kPBlockLev], oE[kPBlockLev];
  int oCx = 0;
};

// ---------------------------------------------------------------------------
// Device helpers

// Sum of G block partials in a fixed order (every block gets the same value).
template <int TB>
__device__ __forceinline__ double foldPartials(const double *__restrict__ part, int G, double *red)
{
  double s = 0.;
  for(int g = threadIdx.x; g < G; g += TB) s += part[g];
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0.;
#pragma unroll
  for(int w = 0; w < TB / 64; ++w) r += red[w];
  __syncthreads();
  return r;
}

// Block sum of one value per thread into part[blockIdx.x] (fixed order).
template <int TB>
__device__ __forceinline__ void blockPartial(double v, double *red, double *__restrict__ part)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if(threadIdx.x == 0) {
    double r = 0.;
#pragma unroll
    for(int w = 0; w < TB / 64; ++w) r += red[w];
    part[blockIdx.x] = r;
  }
}

// sum_k val[k] f(ja[k]) over [kb, ke), kPLPR lanes per row (lanes sub of a
// group of kPLPR consecutive lanes); every lane of the group gets the sum.
template <typename T, typename TV, typename F>
__device__ __forceinline__ T rowSum(const int *__restrict__ ja, const TV *__restrict__ val, int kb, int ke, int sub,
                                    F f)
{
  T acc = (T)0;
  for(int k = kb + sub; k < ke; k += kPLPR) acc += (T)val[k] * f(ja[k]);
#pragma unroll
  for(int o = 1; o < kPLPR; o <<= 1) acc += __shfl_xor(acc, o, 64);
  return acc;
}

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c)
{
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmaT<float>(float a, float b, float c)
{
  return __builtin_fmaf(a, b, c);
}

// ---------------------------------------------------------------------------
// Kernels of one PCG iteration

// rho = (r, z) from the previous up0 partials; p = z + beta p_old; q = A p.
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgProduct(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                   const TV *__restrict__ val, const T *__restrict__ z,
                                                   const T *__restrict__ pOld, T *__restrict__ pNew,
                                                   T *__restrict__ q, const double *__restrict__ rzPart, int Grz,
                                                   const double *__restrict__ rrPart, int Grr, double *scal, int slot,
                                                   int first, double tol2, int *done, int *stat, int k,
                                                   double *__restrict__ pqPart)
{
  __shared__ double red[kPTB / 64];
  const int dn = *done;
  const int row = blockIdx.x * kPRPB + (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  int kb = 0, ke = 0;
  if(row < n) {
    kb = ia[row];
    ke = ia[row + 1];
  }
  if(dn) return;
  const double rho = foldPartials<kPTB>(rzPart, Grz, red);
  double beta = 0.;
  if(!first) beta = rho / scal[slot ^ 1];
  if(blockIdx.x == 0) {
    const double rr = foldPartials<kPTB>(rrPart, Grr, red);
    if(threadIdx.x == 0) {
      scal[slot] = rho;
      scal[2] = rr;
      if(!(rho > 0.) || !(rr == rr)) {
        *done = 3;
        stat[1] = k;
      } else if(!first && rr <= tol2) {
        *done = 1;
        stat[1] = k;
      }
    }
  }
  const T bt = (T)beta;
  // first: p = z (the previous direction is not read)
  const T acc = rowSum<T, TV>(ja, val, kb, ke, sub, [&](int j) { return first ? z[j] : fmaT<T>(bt, pOld[j], z[j]); });
  double pq = 0.;
  if(row < n && sub == 0) {
    const T p = first ? z[row] : fmaT<T>(bt, pOld[row], z[row]);
    pNew[row] = p;
    q[row] = acc;
    pq = (double)p * (double)acc;
  }
  blockPartial<kPTB>(pq, red, pqPart);
}

// alpha = rho / (p, q); x += alpha p; r_new = r_old - alpha q; the first
// smoothing step and the restriction to level 1 on a tile of whole
// aggregates; partials of (r_new, r_new).  first: alpha = 0, x untouched.
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgDown0(PLev<T, TV> L, T *__restrict__ x, const T *__restrict__ p,
                                                 const T *__restrict__ q, const T *__restrict__ rOld,
                                                 T *__restrict__ rNew, T *__restrict__ u, T *__restrict__ bNext,
                                                 const double *__restrict__ pqPart, int Gpq, const double *scal,
                                                 int slot, int first, int *done, int *stat, int k,
                                                 double *__restrict__ rrPart)
{
  __shared__ double red[kPTB / 64];
  __shared__ T res[kPRPB];
  const int dn = *done;
  const int t = blockIdx.x;
  const int r0 = L.tileStart[t], r1 = L.tileStart[t + 1];
  const int rl = (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  const int row = r0 + rl;
  int kb = 0, ke = 0;
  if(row < r1) {
    kb = L.ia[row];
    ke = L.ia[row + 1];
  }
  if(dn) return;
  double alpha = 0.;
  if(!first) {
    const double pq = foldPartials<kPTB>(pqPart, Gpq, red);
    if(pq > 0.) alpha = scal[slot] / pq;
    else if(t == 0 && threadIdx.x == 0) {
      *done = 3;  // (p, A p) <= 0: no update, the host restarts from the FP64 residual
      stat[1] = k;
    }
  }
  const T a = (T)alpha;
  // first: r_new = r_old (q and p are not read)
  const T acc = rowSum<T, TV>(L.ja, L.val, kb, ke, sub,
                              [&](int j) { return (T)L.l1[j] * (first ? rOld[j] : fmaT<T>(-a, q[j], rOld[j])); });
  double rr = 0.;
  if(row < r1 && sub == 0) {
    const T rn = first ? rOld[row] : fmaT<T>(-a, q[row], rOld[row]);
    rNew[row] = rn;
    if(!first) x[row] = fmaT<T>(a, p[row], x[row]);
    u[row] = (T)L.l1[row] * rn;
    res[rl] = rn - acc;
    rr = (double)rn * (double)rn;
  }
  __syncthreads();
  const int I1 = L.tileAgg[t + 1];
  for(int I = L.tileAgg[t] + (int)threadIdx.x; I < I1; I += kPTB) {
    const int m1 = L.aggStart[I + 1] - r0;
    T s = (T)0;
    for(int m = L.aggStart[I] - r0; m < m1; ++m) s += res[m];
    bNext[I] = s;
  }
  blockPartial<kPTB>(rr, red, rrPart);
}

// Level l >= 1: u = S b; b_next = P^T (b - A u) over tiles of whole aggregates.
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgDownL(PLev<T, TV> L, T *__restrict__ bNext, const int *done)
{
  __shared__ T res[kPRPB];
  const int dn = *done;
  const int t = blockIdx.x;
  const int r0 = L.tileStart[t], r1 = L.tileStart[t + 1];
  const int rl = (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  const int row = r0 + rl;
  int kb = 0, ke = 0;
  if(row < r1) {
    kb = L.ia[row];
    ke = L.ia[row + 1];
  }
  if(dn) return;
  const T acc = rowSum<T, TV>(L.ja, L.val, kb, ke, sub, [&](int j) { return (T)L.l1[j] * L.b[j]; });
  if(row < r1 && sub == 0) {
    const T bi = L.b[row];
    L.u[row] = (T)L.l1[row] * bi;
    res[rl] = bi - acc;
  }
  __syncthreads();
  const int I1 = L.tileAgg[t + 1];
  for(int I = L.tileAgg[t] + (int)threadIdx.x; I < I1; I += kPTB) {
    const int m1 = L.aggStart[I + 1] - r0;
    T s = (T)0;
    for(int m = L.aggStart[I] - r0; m < m1; ++m) s += res[m];
    bNext[I] = s;
  }
}

// Level l >= 1: y = u + oc e_next[agg]; e = y + S (b - A y).
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgUpL(PLev<T, TV> L, const T *__restrict__ eNext, T oc, const int *done)
{
  const int dn = *done;
  const int row = blockIdx.x * kPRPB + (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  int kb = 0, ke = 0;
  if(row < L.n) {
    kb = L.ia[row];
    ke = L.ia[row + 1];
  }
  if(dn) return;
  const T acc = rowSum<T, TV>(L.ja, L.val, kb, ke, sub, [&](int j) { return fmaT<T>(oc, eNext[L.agg[j]], L.u[j]); });
  if(row < L.n && sub == 0) {
    const T y = fmaT<T>(oc, eNext[L.agg[row]], L.u[row]);
    L.e[row] = fmaT<T>((T)L.l1[row], L.b[row] - acc, y);
  }
}

// Level 0: z = y + S (r - A y), y = u + oc e_1[agg]; partials of (r, z).
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgUp0(PLev<T, TV> L, const T *__restrict__ u, const T *__restrict__ eNext,
                                               const T *__restrict__ r, T *__restrict__ z, T oc, const int *done,
                                               double *__restrict__ rzPart)
{
  __shared__ double red[kPTB / 64];
  const int dn = *done;
  const int row = blockIdx.x * kPRPB + (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  int kb = 0, ke = 0;
  if(row < L.n) {
    kb = L.ia[row];
    ke = L.ia[row + 1];
  }
  if(dn) return;
  const T acc = rowSum<T, TV>(L.ja, L.val, kb, ke, sub, [&](int j) { return fmaT<T>(oc, eNext[L.agg[j]], u[j]); });
  double rz = 0.;
  if(row < L.n && sub == 0) {
    const T y = fmaT<T>(oc, eNext[L.agg[row]], u[row]);
    const T ri = r[row];
    const T zi = fmaT<T>((T)L.l1[row], ri - acc, y);
    z[row] = zi;
    rz = (double)ri * (double)zi;
  }
  blockPartial<kPTB>(rz, red, rzPart);
}

// Levels from .. coarsest in one workgroup, staged in LDS: b of level from is
// read from global memory, its correction e written back.
template <typename T, typename TV>
__global__ __launch_bounds__(kPBlockTB) void pcgBlock(const PLev<T, TV> *__restrict__ lv, PBlockPlan pl,
                                                      const T *__restrict__ cinv, T oc, const int *done)
{
  __shared__ __align__(16) unsigned char lds[kPBlockLds];
  if(*done) return;
  const int t = threadIdx.x, nt = blockDim.x;
  const int nLev = pl.coarsest - pl.from;
  for(int s = 0; s < nLev; ++s) {
    const PLev<T, TV> L = lv[pl.from + s];
    int *ia = (int *)(lds + pl.oIa[s]);
    unsigned short *ja = (unsigned short *)(lds + pl.oJa[s]);
    TV *val = (TV *)(lds + pl.oVal[s]), *l1 = (TV *)(lds + pl.oL1[s]);
    unsigned short *agg = (unsigned short *)(lds + pl.oAgg[s]), *as = (unsigned short *)(lds + pl.oAs[s]);
    for(int i = t; i <= pl.n[s]; i += nt) ia[i] = L.ia[i];
    for(int k = t; k < pl.nnz[s]; k += nt) {
      ja[k] = (unsigned short)L.ja[k];
      val[k] = L.val[k];
    }
    for(int i = t; i < pl.n[s]; i += nt) {
      l1[i] = L.l1[i];
      agg[i] = (unsigned short)L.agg[i];
    }
    for(int I = t; I <= pl.nc[s]; I += nt) as[I] = (unsigned short)L.aggStart[I];
  }
  {
    T *b = (T *)(lds + pl.oB[0]);
    const T *src = lv[pl.from].b;
    for(int i = t; i < pl.n[0]; i += nt) b[i] = src[i];
  }
  __syncthreads();
  for(int s = 0; s < nLev; ++s) {
    const int n = pl.n[s], nc = pl.nc[s];
    const int *ia = (const int *)(lds + pl.oIa[s]);
    const unsigned short *ja = (const unsigned short *)(lds + pl.oJa[s]);
    const TV *val = (const TV *)(lds + pl.oVal[s]), *l1 = (const TV *)(lds + pl.oL1[s]);
    const unsigned short *as = (const unsigned short *)(lds + pl.oAs[s]);
    const T *b = (const T *)(lds + pl.oB[s]);
    T *u = (T *)(lds + pl.oU[s]), *r = (T *)(lds + pl.oY[s]), *bn = (T *)(lds + pl.oB[s + 1]);
    for(int i = t; i < n; i += nt) {
      T acc = (T)0;
      for(int k = ia[i]; k < ia[i + 1]; ++k) acc += (T)val[k] * ((T)l1[ja[k]] * b[ja[k]]);
      u[i] = (T)l1[i] * b[i];
      r[i] = b[i] - acc;
    }
    __syncthreads();
    for(int I = t; I < nc; I += nt) {
      T acc = (T)0;
      for(int m = as[I]; m < as[I + 1]; ++m) acc += r[m];
      bn[I] = acc;
    }
    __syncthreads();
  }
  {
    const int nC = pl.n[nLev];
    const T *b = (const T *)(lds + pl.oB[nLev]);
    T *x = nLev == 0 ? lv[pl.from].e : (T *)(lds + pl.oCx);
    for(int i = t; i < nC; i += nt) {
      T acc = (T)0;
      for(int j = 0; j < nC; ++j) acc += cinv[i * nC + j] * b[j];
      x[i] = acc;
    }
    __syncthreads();
  }
  for(int s = nLev - 1; s >= 0; --s) {
    const int n = pl.n[s];
    const int *ia = (const int *)(lds + pl.oIa[s]);
    const unsigned short *ja = (const unsigned short *)(lds + pl.oJa[s]);
    const TV *val = (const TV *)(lds + pl.oVal[s]), *l1 = (const TV *)(lds + pl.oL1[s]);
    const unsigned short *agg = (const unsigned short *)(lds + pl.oAgg[s]);
    const T *b = (const T *)(lds + pl.oB[s]), *u = (const T *)(lds + pl.oU[s]);
    T *y = (T *)(lds + pl.oY[s]);
    const T *en = (s + 1 == nLev) ? (const T *)(lds + pl.oCx) : (const T *)(lds + pl.oE[s + 1]);
    T *e = s == 0 ? lv[pl.from].e : (T *)(lds + pl.oE[s]);
    for(int i = t; i < n; i += nt) y[i] = fmaT<T>(oc, en[agg[i]], u[i]);
    __syncthreads();
    for(int i = t; i < n; i += nt) {
      T acc = (T)0;
      for(int k = ia[i]; k < ia[i + 1]; ++k) acc += (T)val[k] * y[ja[k]];
      e[i] = fmaT<T>((T)l1[i], b[i] - acc, y[i]);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Numeric set-up

// Level 0: permuted values (the original CSR entry of every permuted entry)
// and l1 = c / sum |a_ij|, 4 lanes per row.
template <typename TV>
__global__ __launch_bounds__(kPTB) void pcgSetup0(int n, const int *__restrict__ iaP, const int *__restrict__ map,
                                                  const double *__restrict__ val, double c, TV *__restrict__ valP,
                                                  TV *__restrict__ l1)
{
  const int row = blockIdx.x * kPRPB + (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  if(row >= n) return;
  double s = 0.;
  for(int k = iaP[row] + sub; k < iaP[row + 1]; k += kPLPR) {
    const double v = val[map[k]];
    valP[k] = (TV)v;
    s += fabs(v);
  }
#pragma unroll
  for(int o = 1; o < kPLPR; o <<= 1) s += __shfl_xor(s, o, 64);
  if(sub == 0) l1[row] = (TV)(c / s);
}

// Level l >= 1: Galerkin entries as fixed-order sums of the previous level's
// FP64 entries, their TV copy, and l1; one row per 4 lanes.
template <typename TV>
__global__ __launch_bounds__(kPTB) void pcgSetupL(int n, const int *__restrict__ ia, const int *__restrict__ cptr,
                                                  const int *__restrict__ clist, const double *__restrict__ src,
                                                  double c, double *__restrict__ val64, TV *__restrict__ valT,
                                                  TV *__restrict__ l1)
{
  const int row = blockIdx.x * kPRPB + (int)threadIdx.x / kPLPR, sub = threadIdx.x % kPLPR;
  if(row >= n) return;
  double s = 0.;
  for(int k = ia[row] + sub; k < ia[row + 1]; k += kPLPR) {
    double v = 0.;
    for(int m = cptr[k]; m < cptr[k + 1]; ++m) v += src[clist[m]];
    val64[k] = v;
    valT[k] = (TV)v;
    s += fabs(v);
  }
#pragma unroll
  for(int o = 1; o < kPLPR; o <<= 1) s += __shfl_xor(s, o, 64);
  if(sub == 0) l1[row] = (TV)(c / s);
}

// Dense inverse of the coarsest level (n <= kPDense) by Gauss-Jordan in LDS
// (FP64, SPD: no pivoting); flag = 1 on a non-positive pivot.
__global__ __launch_bounds__(1024) void pcgDenseInverse(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                        const double *__restrict__ val, double *__restrict__ inv64,
                                                        float *__restrict__ inv32, int *flag)
{
  __shared__ double M[kPDense][2 * kPDense + 1];
  __shared__ double f[kPDense];
  const int t = threadIdx.x, nt = blockDim.x, w = 2 * n;
  for(int q = t; q < n * w; q += nt) M[q / w][q % w] = (q % w) == n + q / w ? 1. : 0.;
  __syncthreads();
  for(int i = t; i < n; i += nt)
    for(int k = ia[i]; k < ia[i + 1]; ++k) M[i][ja[k]] = val[k];
  __syncthreads();
  for(int p = 0; p < n; ++p) {
    const double piv = M[p][p];
    if(!(piv > 0.)) {
      if(t == 0) *flag = 1;
      return;
    }
    __syncthreads();
    for(int c = t; c < w; c += nt) M[p][c] /= piv;
    for(int i = t; i < n; i += nt) f[i] = M[i][p];  // multipliers, read before any row changes
    __syncthreads();
    for(int q = t; q < n * w; q += nt) {
      const int i = q / w, c = q % w;
      if(i != p) M[i][c] -= f[i] * M[p][c];
    }
    __syncthreads();
  }
  for(int q = t; q < n * n; q += nt) {
    const double v = M[q / n][n + q % n];
    inv64[q] = v;
    inv32[q] = (float)v;
  }
}

// ---------------------------------------------------------------------------
// Outer-loop helpers

template <typename T>
__global__ void pcgPermuteRhs(int n, const int *__restrict__ perm, const double *__restrict__ r64, T *__restrict__ r)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r[i] = (T)r64[perm[i]];
}

template <typename T>
__global__ void pcgAddBack(int n, const int *__restrict__ perm, const T *__restrict__ x, double *__restrict__ x64)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x64[perm[i]] += (double)x[i];
}

template <typename T>
__global__ void pcgScatterBack(int n, const int *__restrict__ perm, const T *__restrict__ x, double *__restrict__ y)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[perm[i]] = (double)x[i];
}

template <typename T>
__global__ void pcgZero(int n, T *__restrict__ a, int *done, int *stat)
{
  for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = (T)0;
  if(blockIdx.x == 0 && threadIdx.x == 0) {
    *done = 0;
    stat[0] = 0;
    stat[1] = -1;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host state

struct FeAmgPcgLevel {
  int n = 0, nc = 0, nTiles = 0;
  int64_t nnz = 0;
  int *ia = nullptr, *ja = nullptr, *agg = nullptr, *aggStart = nullptr, *tileStart = nullptr, *tileAgg = nullptr;
  int *cptr = nullptr, *clist = nullptr;  // levels >= 1
  double *val64 = nullptr;                // levels >= 1 (level 0: the values are permuted from the system's)
  float *val32 = nullptr, *l1_32 = nullptr;
  double *valD = nullptr, *l1_64 = nullptr;  // FP64 value copy (levels >= 1: val64), l1 in FP64
  float *b32 = nullptr, *u32 = nullptr, *e32 = nullptr;
  double *b64 = nullptr, *u64 = nullptr, *e64 = nullptr;
};

struct FeAmgPcg {
  std::vector<FeAmgPcgLevel> lv;
  int *perm = nullptr;  // level-0 permuted row -> original row
  int *map = nullptr;   // level-0 permuted entry -> original entry
  float *cinv32 = nullptr;
  double *cinv64 = nullptr;
  int *flag = nullptr;  // non-positive coarsest pivot
  PBlockPlan plan[3];   // (vectors, values): 0 float/float, 1 double/float, 2 double/double
  void *desc[3] = {nullptr, nullptr, nullptr};
  // level-0 solver vectors (permuted), one set per vector type
  float *x32 = nullptr, *r32[2] = {}, *p32[2] = {}, *q32 = nullptr, *z32 = nullptr, *rhs32 = nullptr;
  double *x64v = nullptr, *r64v[2] = {}, *p64v[2] = {}, *q64v = nullptr, *z64v = nullptr, *rhs64v = nullptr;
  double *pq = nullptr, *rr = nullptr, *rz = nullptr, *scal = nullptr, *xsave = nullptr;
  int *done = nullptr, *stat = nullptr;
  int g0 = 0;  // grid of the row kernels of level 0
  double c = 0.;
  double valuesFrom = -1.;  // unused marker
};

namespace {

template <typename T>
int upload(fe_hip_ctx *ctx, T **d, const std::vector<T> &h)
{
  if(int rc = feAlloc(ctx, d, std::max<size_t>(1, h.size()))) return rc;
  if(!h.empty()) FE_HIP_TRY(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return FE_STATUS_OK;
}

template <typename T, typename TV>
PLev<T, TV> pview(const FeAmgPcgLevel &h);
template <>
PLev<float, float> pview<float, float>(const FeAmgPcgLevel &h)
{
  PLev<float, float> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.val32, h.l1_32, h.agg, h.aggStart, h.tileStart,
                       h.tileAgg, h.b32, h.u32, h.e32};
  return L;
}
template <>
PLev<double, float> pview<double, float>(const FeAmgPcgLevel &h)
{
  PLev<double, float> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.val32, h.l1_32, h.agg, h.aggStart, h.tileStart,
                        h.tileAgg, h.b64, h.u64, h.e64};
  return L;
}
template <>
PLev<double, double> pview<double, double>(const FeAmgPcgLevel &h)
{
  PLev<double, double> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.valD, h.l1_64, h.agg, h.aggStart, h.tileStart,
                         h.tileAgg, h.b64, h.u64, h.e64};
  return L;
}

template <typename T, typename TV>
PBlockPlan pcgPlan(const FeAmgPcg &A)
{
  const int Lc = (int)A.lv.size() - 1;
  PBlockPlan pl;
  for(int from = 1; from <= Lc; ++from) {
    if(Lc - from > kPBlockLev) continue;
    bool ok = true;
    for(int l = from; l < Lc; ++l) ok = ok && A.lv[l].n <= kPBlockRows;
    if(!ok) continue;
    int bytes = 0;
    auto take = [&](size_t b) {
      const int o = bytes;
      bytes += (int)((b + 15) & ~(size_t)15);
      return o;
    };
    pl = PBlockPlan();
    pl.from = from;
    pl.coarsest = Lc;
    for(int l = from; l < Lc; ++l) {
      const int s = l - from;
      const FeAmgPcgLevel &h = A.lv[l];
      pl.n[s] = h.n;
      pl.nc[s] = h.nc;
      pl.nnz[s] = (int)h.nnz;
      pl.oIa[s] = take(sizeof(int) * (h.n + 1));
      pl.oJa[s] = take(2 * (size_t)h.nnz);
      pl.oVal[s] = take(sizeof(TV) * (size_t)h.nnz);
      pl.oL1[s] = take(sizeof(TV) * h.n);
      pl.oAgg[s] = take(2 * (size_t)h.n);
      pl.oAs[s] = take(2 * (size_t)(h.nc + 1));
      pl.oB[s] = take(sizeof(T) * h.n);
      pl.oU[s] = take(sizeof(T) * h.n);
      pl.oY[s] = take(sizeof(T) * h.n);
      pl.oE[s] = take(sizeof(T) * h.n);
    }
    const int nC = A.lv[Lc].n;
    pl.n[Lc - from] = nC;
    pl.oB[Lc - from] = take(sizeof(T) * nC);
    pl.oCx = take(sizeof(T) * nC);
    pl.bytes = bytes;
    if(bytes <= kPBlockLds) return pl;
  }
  pl.bytes = kPBlockLds + 1;  // nothing fits
  return pl;
}

// Nested-ordered copy of the host hierarchy (once per CSR pattern).
int pcgSetupPattern(fe_hip_ctx *ctx)
{
  const std::vector<FeAmgHostLevel> &H = ctx->amgHost;
  const int Lc = (int)H.size() - 1;
  FE_REQUIRE(ctx, Lc >= 1 && ctx->amgHostSweeps == 0 && H[Lc].n <= kPDense,
             "nested multigrid PCG: the hierarchy needs a dense coarsest level");
  // pos[l][i]: permuted index of row i of level l; order[l]: its inverse
  std::vector<std::vector<int>> pos(Lc + 1), order(Lc + 1);
  order[Lc].resize(H[Lc].n);
  std::iota(order[Lc].begin(), order[Lc].end(), 0);
  pos[Lc] = order[Lc];
  std::vector<std::vector<int>> aggStart(Lc);
  for(int l = Lc - 1; l >= 0; --l) {
    const FeAmgHostLevel &h = H[l];
    std::vector<int> cnt(h.nc + 1, 0);
    for(int i = 0; i < h.n; ++i) {
      FE_REQUIRE(ctx, h.agg[i] >= 0 && h.agg[i] < h.nc, "nested multigrid PCG: a row without aggregate");
      cnt[pos[l + 1][h.agg[i]] + 1]++;
    }
    for(int I = 0; I < h.nc; ++I) cnt[I + 1] += cnt[I];
    aggStart[l] = cnt;
    order[l].resize(h.n);
    pos[l].resize(h.n);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for(int i = 0; i < h.n; ++i) {
      const int p = fill[pos[l + 1][h.agg[i]]]++;
      order[l][p] = i;
      pos[l][i] = p;
    }
  }
  FeAmgPcg *A = new FeAmgPcg();
  ctx->amgPcg = A;
  A->lv.resize(Lc + 1);
  // permuted CSR of every level and, per level, original entry -> permuted entry
  std::vector<std::vector<int>> entPos(Lc + 1);
  for(int l = 0; l <= Lc; ++l) {
    const FeAmgHostLevel &h = H[l];
    FeAmgPcgLevel &d = A->lv[l];
    d.n = h.n;
    d.nc = l < Lc ? h.nc : 0;
    d.nnz = h.ia[h.n];
    std::vector<int> ia(h.n + 1, 0), ja(d.nnz), ent(d.nnz);
    entPos[l].assign(d.nnz, -1);
    std::vector<std::pair<int, int>> row;
    for(int ip = 0; ip < h.n; ++ip) {
      const int i = order[l][ip];
      row.clear();
      for(int k = h.ia[i]; k < h.ia[i + 1]; ++k) row.emplace_back(pos[l][h.ja[k]], k);
      std::sort(row.begin(), row.end());
      for(size_t m = 0; m < row.size(); ++m) {
        const int kp = ia[ip] + (int)m;
        ja[kp] = row[m].first;
        ent[kp] = row[m].second;
        entPos[l][row[m].second] = kp;
      }
      ia[ip + 1] = ia[ip] + (int)row.size();
    }
    if(int rc = upload(ctx, &d.ia, ia)) return rc;
    if(int rc = upload(ctx, &d.ja, ja)) return rc;
    if(l == 0) {
      if(int rc = upload(ctx, &A->perm, order[0])) return rc;
      if(int rc = upload(ctx, &A->map, ent)) return rc;
    } else {
      // Galerkin lists in permuted numbering: level 1 sums entries of the
      // system's original CSR, deeper levels entries of the permuted level
      // above (same list order as the original lists)
      std::vector<int> cptr(d.nnz + 1, 0), clist;
      clist.reserve(h.clist.size());
      for(int kp = 0; kp < d.nnz; ++kp) {
        const int k = ent[kp];
        for(int m = h.cptr[k]; m < h.cptr[k + 1]; ++m) clist.push_back(l == 1 ? h.clist[m] : entPos[l - 1][h.clist[m]]);
        cptr[kp + 1] = (int)clist.size();
      }
      if(int rc = upload(ctx, &d.cptr, cptr)) return rc;
      if(int rc = upload(ctx, &d.clist, clist)) return rc;
      if(int rc = feAlloc(ctx, &d.val64, (size_t)d.nnz)) return rc;
      d.valD = d.val64;
    }
    if(l < Lc) {
      std::vector<int> aggP(h.n);
      for(int ip = 0; ip < h.n; ++ip) aggP[ip] = pos[l + 1][h.agg[order[l][ip]]];
      if(int rc = upload(ctx, &d.agg, aggP)) return rc;
      if(int rc = upload(ctx, &d.aggStart, aggStart[l])) return rc;
      // tiles of whole aggregates, at most kPRPB rows
      std::vector<int> ts(1, 0), ta(1, 0);
      int rows = 0;
      for(int I = 0; I < h.nc; ++I) {
        const int sz = aggStart[l][I + 1] - aggStart[l][I];
        FE_REQUIRE(ctx, sz <= kPRPB, "nested multigrid PCG: an aggregate is larger than a tile");
        if(rows + sz > kPRPB) {
          ts.push_back(aggStart[l][I]);
          ta.push_back(I);
          rows = 0;
        }
        rows += sz;
      }
      ts.push_back(h.n);
      ta.push_back(h.nc);
      d.nTiles = (int)ts.size() - 1;
      if(int rc = upload(ctx, &d.tileStart, ts)) return rc;
      if(int rc = upload(ctx, &d.tileAgg, ta)) return rc;
    }
    if(int rc = feAlloc(ctx, &d.val32, (size_t)d.nnz)) return rc;
    if(int rc = feAlloc(ctx, &d.l1_32, (size_t)d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.l1_64, (size_t)d.n)) return rc;
    if(l == 0)
      if(int rc = feAlloc(ctx, &d.valD, (size_t)d.nnz)) return rc;
    if(l > 0) {
      if(int rc = feAlloc(ctx, &d.b32, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.e32, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.b64, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.e64, d.n)) return rc;
    }
    if(int rc = feAlloc(ctx, &d.u32, d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.u64, d.n)) return rc;
  }
  const size_t nC = (size_t)A->lv[Lc].n;
  if(int rc = feAlloc(ctx, &A->cinv32, nC * nC)) return rc;
  if(int rc = feAlloc(ctx, &A->cinv64, nC * nC)) return rc;
  if(int rc = feAlloc(ctx, &A->flag, 1)) return rc;
  const int n = A->lv[0].n;
  for(int s = 0; s < 2; ++s) {
    if(int rc = feAlloc(ctx, &A->r32[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->p32[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->r64v[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->p64v[s], n)) return rc;
  }
  if(feAlloc(ctx, &A->x32, n) || feAlloc(ctx, &A->q32, n) || feAlloc(ctx, &A->z32, n) || feAlloc(ctx, &A->rhs32, n) ||
     feAlloc(ctx, &A->x64v, n) || feAlloc(ctx, &A->q64v, n) || feAlloc(ctx, &A->z64v, n) ||
     feAlloc(ctx, &A->rhs64v, n) || feAlloc(ctx, &A->xsave, n))
    return FE_STATUS_OUT_OF_MEMORY;
  A->g0 = (n + kPRPB - 1) / kPRPB;
  const int gMax = std::max(A->g0, A->lv[0].nTiles);
  if(feAlloc(ctx, &A->pq, gMax) || feAlloc(ctx, &A->rr, gMax) || feAlloc(ctx, &A->rz, gMax) ||
     feAlloc(ctx, &A->scal, 8) || feAlloc(ctx, &A->done, 4) || feAlloc(ctx, &A->stat, 4))
    return FE_STATUS_OUT_OF_MEMORY;
  A->plan[0] = pcgPlan<float, float>(*A);
  A->plan[1] = pcgPlan<double, float>(*A);
  A->plan[2] = pcgPlan<double, double>(*A);
  {
    std::vector<PLev<float, float>> a;
    std::vector<PLev<double, float>> b;
    std::vector<PLev<double, double>> c;
    for(const FeAmgPcgLevel &h : A->lv) {
      a.push_back(pview<float, float>(h));
      b.push_back(pview<double, float>(h));
      c.push_back(pview<double, double>(h));
    }
    PLev<float, float> *da = nullptr;
    PLev<double, float> *db = nullptr;
    PLev<double, double> *dc = nullptr;
    if(int rc = upload(ctx, &da, a)) return rc;
    if(int rc = upload(ctx, &db, b)) return rc;
    if(int rc = upload(ctx, &dc, c)) return rc;
    A->desc[0] = da;
    A->desc[1] = db;
    A->desc[2] = dc;
  }
  return FE_STATUS_OK;
}

// Numeric set-up for the current values (every solve): permuted level-0
// values and l1, Galerkin sums and l1 per level, dense coarsest inverse.
int pcgNumeric(fe_hip_ctx *ctx, bool fp64Values)
{
  FeAmgPcg &A = *ctx->amgPcg;
  A.c = ctx->amgOmega;
  FeAmgPcgLevel &f = A.lv[0];
  const int g0 = (f.n + kPRPB - 1) / kPRPB;
  if(fp64Values)
    hipLaunchKernelGGL(pcgSetup0<double>, dim3(g0), dim3(kPTB), 0, ctx->stream, f.n, f.ia, A.map, ctx->d_val, A.c,
                       f.valD, f.l1_64);
  else
    hipLaunchKernelGGL(pcgSetup0<float>, dim3(g0), dim3(kPTB), 0, ctx->stream, f.n, f.ia, A.map, ctx->d_val, A.c,
                       f.val32, f.l1_32);
  for(size_t l = 1; l < A.lv.size(); ++l) {
    FeAmgPcgLevel &c = A.lv[l];
    const double *src = l == 1 ? ctx->d_val : A.lv[l - 1].val64;
    const int g = (c.n + kPRPB - 1) / kPRPB;
    if(fp64Values)
      hipLaunchKernelGGL(pcgSetupL<double>, dim3(g), dim3(kPTB), 0, ctx->stream, c.n, c.ia, c.cptr, c.clist, src,
                         A.c, c.val64, c.valD, c.l1_64);
    else
      hipLaunchKernelGGL(pcgSetupL<float>, dim3(g), dim3(kPTB), 0, ctx->stream, c.n, c.ia, c.cptr, c.clist, src, A.c,
                         c.val64, c.val32, c.l1_32);
  }
  const FeAmgPcgLevel &C = A.lv.back();
  hipLaunchKernelGGL(pcgZero<int>, dim3(1), dim3(64), 0, ctx->stream, 1, A.flag, A.done, A.stat);
  hipLaunchKernelGGL(pcgDenseInverse, dim3(1), dim3(1024), 0, ctx->stream, C.n, C.ia, C.ja, C.val64, A.cinv64,
                     A.cinv32, A.flag);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

template <typename T>
struct PcgVecs {
  T *x, *r[2], *p[2], *q, *z, *rhs;
  const T *cinv;
};
PcgVecs<float> vecs32(FeAmgPcg &A)
{
  return PcgVecs<float>{A.x32, {A.r32[0], A.r32[1]}, {A.p32[0], A.p32[1]}, A.q32, A.z32, A.rhs32, A.cinv32};
}
PcgVecs<double> vecs64(FeAmgPcg &A)
{
  return PcgVecs<double>{A.x64v, {A.r64v[0], A.r64v[1]}, {A.p64v[0], A.p64v[1]}, A.q64v, A.z64v, A.rhs64v, A.cinv64};
}

// Launches the V-cycle after down0: down_l (levels 1 .. B-1), the block
// kernel, up_l (levels B-1 .. 1), then up0 on rNew into z.
template <typename T, typename TV>
void vcycleRest(fe_hip_ctx *ctx, FeAmgPcg &A, int slot, const PcgVecs<T> &v, const T *rNew)
{
  const PBlockPlan &pl = A.plan[slot];
  const int B = pl.from;
  const T oc = (T)ctx->amgOverCorrect;
  const PLev<T, TV> *d = (const PLev<T, TV> *)A.desc[slot];
  for(int l = 1; l < B; ++l) {
    const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
    const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
    hipLaunchKernelGGL((pcgDownL<T, TV>), dim3(L.nTiles), dim3(kPTB), 0, ctx->stream, L, N.b, A.done);
  }
  hipLaunchKernelGGL((pcgBlock<T, TV>), dim3(1), dim3(kPBlockTB), 0, ctx->stream, d, pl, v.cinv, oc, A.done);
  for(int l = B - 1; l >= 1; --l) {
    const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
    const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
    hipLaunchKernelGGL((pcgUpL<T, TV>), dim3((L.n + kPRPB - 1) / kPRPB), dim3(kPTB), 0, ctx->stream, L, N.e, oc,
                       A.done);
  }
  const PLev<T, TV> L0 = pview<T, TV>(A.lv[0]);
  const PLev<T, TV> L1 = pview<T, TV>(A.lv[1]);
  hipLaunchKernelGGL((pcgUp0<T, TV>), dim3(A.g0), dim3(kPTB), 0, ctx->stream, L0, (const T *)L0.u, (const T *)L1.e,
                     rNew, v.z, oc, A.done, A.rz);
}

// One inner solve of A x = rhs (permuted, vector type T) from x = 0 until
// (r, r) <= tol2, maxit iterations or a breakdown.  Returns the iterations
// in *iters and the stop code (1 converged, 2 iteration cap, 3 breakdown).
template <typename T, typename TV>
int innerSolve(fe_hip_ctx *ctx, FeAmgPcg &A, int slot, const PcgVecs<T> &v, double tol2, int maxit, int *iters,
               int *code)
{
  const int n = A.lv[0].n;
  const PLev<T, TV> L0 = pview<T, TV>(A.lv[0]);
  const PLev<T, TV> L1 = pview<T, TV>(A.lv[1]);
  hipLaunchKernelGGL(pcgZero<T>, dim3(256), dim3(256), 0, ctx->stream, n, v.x, A.done, A.stat);
  // z_0 = M r_0 (down0 with alpha = 0 copies r_0 into r[0])
  hipLaunchKernelGGL((pcgDown0<T, TV>), dim3(L0.nTiles), dim3(kPTB), 0, ctx->stream, L0, v.x, (const T *)v.p[0],
                     (const T *)v.q, (const T *)v.rhs, v.r[0], L0.u, L1.b, (const double *)A.pq, A.g0,
                     (const double *)A.scal, 0, 1, A.done, A.stat, 0, A.rr);
  vcycleRest<T, TV>(ctx, A, slot, v, v.r[0]);
  int *hs = (int *)ctx->h_state;  // pinned: stat[0..1], done
  int k = 0, batch = 12, kPrev = -1;
  double rrPrev = -1.;
  for(;;) {
    const int kEnd = std::min(maxit, k + batch);
    for(; k < kEnd; ++k) {
      const int s = k & 1;
      // p_k from z_k and p_k-1; q_k = A p_k
      hipLaunchKernelGGL((pcgProduct<T, TV>), dim3(A.g0), dim3(kPTB), 0, ctx->stream, n, L0.ia, L0.ja, L0.val,
                         (const T *)v.z, (const T *)v.p[s ^ 1], v.p[s], v.q, (const double *)A.rz, A.g0,
                         (const double *)A.rr, L0.nTiles, A.scal, s, k == 0 ? 1 : 0, tol2, A.done, A.stat, k, A.pq);
      // r_k+1 = r_k - alpha q_k into r[(k + 1) & 1], x += alpha p_k, V-cycle down
      hipLaunchKernelGGL((pcgDown0<T, TV>), dim3(L0.nTiles), dim3(kPTB), 0, ctx->stream, L0, v.x, (const T *)v.p[s],
                         (const T *)v.q, (const T *)v.r[s], v.r[s ^ 1], L0.u, L1.b, (const double *)A.pq, A.g0,
                         (const double *)A.scal, s, 0, A.done, A.stat, k, A.rr);
      vcycleRest<T, TV>(ctx, A, slot, v, v.r[s ^ 1]);
    }
    FE_HIP_TRY(ctx, hipGetLastError());
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs, A.stat, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync(hs + 2, A.done, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipMemcpyAsync((char *)ctx->h_state + 64, A.scal + 2, sizeof(double), hipMemcpyDeviceToHost,
                                   ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const int dn = hs[2];
    if(dn) {
      *code = dn;
      *iters = hs[1] >= 0 ? hs[1] : k;
      return FE_STATUS_OK;
    }
    if(k >= maxit) {
      *code = 2;
      *iters = k;
      return FE_STATUS_OK;
    }
    // next batch from the observed reduction of (r, r) (the value at the
    // start of the last iteration launched)
    const double rr = *(double *)((char *)ctx->h_state + 64);
    int next = 8;
    if(kPrev >= 0 && rr > 0. && rr < rrPrev && k > kPrev) {
      const double rate = std::log(rr / rrPrev) / (double)(k - kPrev);  // per iteration, of (r, r)
      next = (int)std::ceil(std::log(tol2 / rr) / rate) + 1;
    }
    kPrev = k;
    rrPrev = rr;
    batch = std::max(2, std::min(64, next));
  }
}

}  // namespace

int feAmgPcgSolve(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double bnorm, double target, double *x64,
                  double *r64, FeAmgPcgResult *res)
{
  *res = FeAmgPcgResult();
  if(!ctx->amg)
    if(int rc = feAmgSetup(ctx)) return rc;
  if(!ctx->amgPcg)
    if(int rc = pcgSetupPattern(ctx)) return rc;
  FeAmgPcg &A = *ctx->amgPcg;
  const bool fp32 = o->precision == FE_PREC_FP32_IR;
  const bool val64 = !fp32 && !ctx->pcValues32;
  const int slot = fp32 ? 0 : (val64 ? 2 : 1);
  FE_REQUIRE(ctx, A.plan[slot].bytes <= kPBlockLds, "nested multigrid PCG: the small levels do not fit the LDS");
  if(int rc = pcgNumeric(ctx, val64)) return rc;
  if(!ctx->h_state) FE_HIP_TRY(ctx, hipHostMalloc(&ctx->h_state, 4096));
  const int n = ctx->n;
  double rn = bnorm;
  bool inner32 = fp32;
  int iters = 0, outer = 0;
  while(outer < 60) {
    res->relres = rn / bnorm;
    if(rn <= target) {
      res->converged = 1;
      res->reason = 2;
      break;
    }
    if(iters >= o->max_iter) {
      res->reason = -3;
      break;
    }
    // inner target: FP32 inner solves reduce by inner_rtol; FP64 ones aim at
    // the outer target directly
    double tol2 = inner32 ? (o->inner_rtol * rn) * (o->inner_rtol * rn) : 0.25 * target * target;
    tol2 = std::max(tol2, 0.01 * target * target);
    int it = 0, code = 0;
    const int gp = std::min(1024, (n + 255) / 256);
    FE_HIP_TRY(ctx, hipMemcpyAsync(A.xsave, x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    if(inner32) {
      PcgVecs<float> v = vecs32(A);
      hipLaunchKernelGGL(pcgPermuteRhs<float>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r64, v.rhs);
      if(int rc = innerSolve<float, float>(ctx, A, 0, v, tol2, o->max_iter - iters, &it, &code)) return rc;
      hipLaunchKernelGGL(pcgAddBack<float>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, (const float *)v.x, x64);
    } else {
      PcgVecs<double> v = vecs64(A);
      hipLaunchKernelGGL(pcgPermuteRhs<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r64, v.rhs);
      if(val64) {
        if(int rc = innerSolve<double, double>(ctx, A, 2, v, tol2, o->max_iter - iters, &it, &code)) return rc;
      } else if(int rc = innerSolve<double, float>(ctx, A, 1, v, tol2, o->max_iter - iters, &it, &code)) {
        return rc;
      }
      hipLaunchKernelGGL(pcgAddBack<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, (const double *)v.x, x64);
    }
    iters += it;
    outer++;
    // FP64 residual of the original system: r64 = b - A x64
    double rr = 0.;
    if(int rc = feResidualNorm2(ctx, x64, r64, &rr)) return rc;
    const double rnew = std::sqrt(rr);
    if(code == 3) res->breakdowns++;
    if(inner32 && !(rnew < 0.9 * rn)) {
      // the FP32 correction did not reduce the FP64 residual enough: continue
      // with FP64 inner solves, from the previous iterate when it grew
      inner32 = false;
      if(!(rnew <= rn)) {
        FE_HIP_TRY(ctx, hipMemcpyAsync(x64, A.xsave, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        double rr0 = 0.;
        if(int rc = feResidualNorm2(ctx, x64, r64, &rr0)) return rc;
        rn = std::sqrt(rr0);
        continue;
      }
    }
    if(!(rnew == rnew)) {
      res->reason = -9;
      rn = rnew;
      break;
    }
    rn = rnew;
    if(res->breakdowns > 20) {
      res->reason = -3;
      break;
    }
    if(rn / bnorm > o->dtol) {
      res->reason = -4;
      break;
    }
  }
  res->relres = rn / bnorm;
  if(!res->converged && rn <= target) {
    res->converged = 1;
    res->reason = 2;
  }
  res->iters = iters;
  res->outer = outer;
  int flag = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&flag, A.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, flag == 0, "AMG: the coarsest operator is not positive definite");
  return FE_STATUS_OK;
}

bool feAmgNestedUsable(const fe_hip_ctx *ctx)
{
  return ctx->amgKernel == 2 && ctx->amgHostSweeps == 0 && ctx->amgHost.size() >= 2 &&
         ctx->amgHost.back().n <= kPDense && ctx->amgHostMaxAgg <= kPRPB;
}

int feAmgPcgApply(fe_hip_ctx *ctx, const double *r, double *z)
{
  if(!ctx->amgPcg)
    if(int rc = pcgSetupPattern(ctx)) return rc;
  FeAmgPcg &A = *ctx->amgPcg;
  const bool val64 = !ctx->pcValues32;
  const int slot = val64 ? 2 : 1;
  FE_REQUIRE(ctx, A.plan[slot].bytes <= kPBlockLds, "nested multigrid PCG: the small levels do not fit the LDS");
  if(int rc = pcgNumeric(ctx, val64)) return rc;
  const int n = A.lv[0].n;
  const int gp = std::min(1024, (n + 255) / 256);
  PcgVecs<double> v = vecs64(A);
  hipLaunchKernelGGL(pcgPermuteRhs<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r, v.rhs);
  hipLaunchKernelGGL(pcgZero<double>, dim3(gp), dim3(256), 0, ctx->stream, n, v.x, A.done, A.stat);
  if(val64) {
    const PLev<double, double> L0 = pview<double, double>(A.lv[0]), L1 = pview<double, double>(A.lv[1]);
    hipLaunchKernelGGL((pcgDown0<double, double>), dim3(L0.nTiles), dim3(kPTB), 0, ctx->stream, L0, v.x,
                       (const double *)v.p[0], (const double *)v.q, (const double *)v.rhs, v.r[0], L0.u, L1.b,
                       (const double *)A.pq, A.g0, (const double *)A.scal, 0, 1, A.done, A.stat, 0, A.rr);
    vcycleRest<double, double>(ctx, A, 2, v, v.r[0]);
  } else {
    const PLev<double, float> L0 = pview<double, float>(A.lv[0]), L1 = pview<double, float>(A.lv[1]);
    hipLaunchKernelGGL((pcgDown0<double, float>), dim3(L0.nTiles), dim3(kPTB), 0, ctx->stream, L0, v.x,
                       (const double *)v.p[0], (const double *)v.q, (const double *)v.rhs, v.r[0], L0.u, L1.b,
                       (const double *)A.pq, A.g0, (const double *)A.scal, 0, 1, A.done, A.stat, 0, A.rr);
    vcycleRest<double, float>(ctx, A, 1, v, v.r[0]);
  }
  hipLaunchKernelGGL(pcgScatterBack<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, (const double *)v.z, z);
  FE_HIP_TRY(ctx, hipGetLastError());
  int flag = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&flag, A.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, flag == 0, "AMG: the coarsest operator is not positive definite");
  return FE_STATUS_OK;
}

int feAmgPcgLevels(const fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap)
{
  *levels = 0;
  if(!ctx->amgPcg) return FE_STATUS_OK;
  const FeAmgPcg &A = *ctx->amgPcg;
  *levels = (int)A.lv.size();
  for(int l = 0; l < (int)A.lv.size() && l < cap; ++l) {
    rows[l] = A.lv[l].n;
    nnz[l] = A.lv[l].nnz;
  }
  return FE_STATUS_OK;
}

void feAmgPcgRelease(fe_hip_ctx *ctx)
{
  delete ctx->amgPcg;  // device buffers are owned by the context
  ctx->amgPcg = nullptr;
}
