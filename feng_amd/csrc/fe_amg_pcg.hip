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
// converged.  FP32 iterations run under FP64 iterative refinement on the
// original system with residual replacement: when the recursive residual has
// dropped by the inner tolerance, the FP32 iterate is added to the FP64 one,
// the FP64 residual replaces the recursive one and the stopped iteration is
// redone from the kept search direction (no restart, the Krylov history
// carries over).
#include "fe_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>
#include <vector>

namespace {

constexpr int kPTB = 256;             // threads of the level kernels (levels >= 1)
constexpr int kFTB = 256;             // threads of the level-0 kernels: one row per thread (about n / 256 partials)
constexpr int kPLPR = 4;              // lanes per row of the level kernels
constexpr int kPRPB = kPTB / kPLPR;   // rows per block of a level kernel: the tile limit of levels >= 1
constexpr int kFRPB = kFTB;           // rows per block of a level-0 kernel: the tile limit of level 0
constexpr int kSlice = 64;            // level-0 SELL slice: 64 rows, slot k of the slice's rows contiguous
constexpr int kPBlockRows = 1024;     // levels of at most this many rows run in the single-workgroup kernel
constexpr int kPBlockTB = 1024;
constexpr int kPBlockLds = 150 * 1024;
constexpr int kPDense = 64;           // the coarsest level is inverted densely up to this size
constexpr int kPBlockLev = 8;

template <typename T, typename TV>
struct PLev {
  int n, nc, nTiles;
  const int *ia, *ja;   // level 0: ja holds the SELL columns
  const TV *val, *l1;   // level 0: SELL values
  const int2 *sell;     // level 0: per row (first SELL position, length); entry k at first + 64 k
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
  int oCx = 0, oCi = 0;
  int imgBytes = 0;  // the static part (matrices, aggregates, coarsest inverse) comes first: its image is
                     // packed once per numeric set-up and copied into LDS in one pass
};

// Host-mapped (pinned, coherent) words the product kernel posts to, so the
// host can keep a bounded number of iterations queued ahead of the device and
// stop launching as soon as the device flags a stop, without synchronizing
// the stream: progress = the iteration whose product last started, done = the
// stop code (1 converged, 3 breakdown), stopK = its iteration, rr = the
// recursive (r, r) at the stop; barErr = a grid barrier of the fused
// coarse-level launch timed out.
struct PcgMailbox {
  int progress, done, stopK, barErr;
  double rr;
};

// ---------------------------------------------------------------------------
// Device helpers

// System-scope stores into the mailbox (vector stores, no ordering: the host
// reads rr only after synchronizing the stream, and progress and done are
// independent hints).
__device__ __forceinline__ void mbPost(int *p, int v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void mbStop(PcgMailbox *mb, int code, int k, double rr)
{
  mb->rr = rr;
  mbPost(&mb->stopK, k);
  mbPost(&mb->done, code);
}

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

// The two halves of foldPartials, so the partial loads can be issued with the
// kernel's other independent loads and reduced later: this thread's strided
// share, then the block-wide sum (the same value in every block).
template <int TB>
__device__ __forceinline__ double partialsShare(const double *__restrict__ part, int G)
{
  double s = 0.;
#pragma unroll 4
  for(int g = threadIdx.x; g < G; g += TB) s += part[g];
  return s;
}
template <int TB>
__device__ __forceinline__ double blockAllSum(double s, double *red)
{
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

// Level 0, one row per thread: sum_k val f(col) over the row's SELL entries,
// in batches of kBatch entries whose column, value and gathered loads are
// all issued before any of them is used (a row is one or two batches), so a
// row costs a few memory round trips instead of one per entry.
constexpr int kBatch = 8;
// Two row sums over one pass of the SELL columns and values.
template <typename T, typename TV>
__device__ __forceinline__ void sellSum2(const int *__restrict__ col, const TV *__restrict__ val, int2 bl,
                                         const T *__restrict__ f1, const T *__restrict__ f2, T &s1, T &s2)
{
  s1 = (T)0;
  s2 = (T)0;
  for(int k0 = 0; k0 < bl.y; k0 += kBatch) {
    int c[kBatch];
    TV v[kBatch];
#pragma unroll
    for(int u = 0; u < kBatch; ++u) {
      const bool in = k0 + u < bl.y;
      const int q = bl.x + (in ? k0 + u : 0) * kSlice;
      c[u] = col[q];
      v[u] = in ? val[q] : (TV)0;
    }
    T g1[kBatch], g2[kBatch];
#pragma unroll
    for(int u = 0; u < kBatch; ++u) {
      g1[u] = f1[c[u]];
      g2[u] = f2[c[u]];
    }
#pragma unroll
    for(int u = 0; u < kBatch; ++u) {
      s1 += (T)v[u] * g1[u];
      s2 += (T)v[u] * g2[u];
    }
  }
}
template <typename T, typename TV, typename F>
__device__ __forceinline__ T sellSum(const int *__restrict__ col, const TV *__restrict__ val, int2 bl, F f)
{
  T acc = (T)0;
  for(int k0 = 0; k0 < bl.y; k0 += kBatch) {
    int c[kBatch];
    TV v[kBatch];
#pragma unroll
    for(int u = 0; u < kBatch; ++u) {
      const bool in = k0 + u < bl.y;
      const int q = bl.x + (in ? k0 + u : 0) * kSlice;
      c[u] = col[q];
      v[u] = in ? val[q] : (TV)0;
    }
    T g[kBatch];
#pragma unroll
    for(int u = 0; u < kBatch; ++u) g[u] = f(c[u]);
#pragma unroll
    for(int u = 0; u < kBatch; ++u) acc += (T)v[u] * g[u];
  }
  return acc;
}

// sum_k val[k] f(ja[k]) over [kb, ke), kPLPR lanes per row (lanes sub of a
// group of kPLPR consecutive lanes); every lane of the group gets the sum.
template <typename T, typename TV, typename F>
__device__ __forceinline__ T rowSum(const int *__restrict__ ja, const TV *__restrict__ val, int kb, int ke, int sub,
                                    F f)
{
  constexpr int U = 4;  // entries per lane per batch, loads issued together
  T acc = (T)0;
  for(int k0 = kb + sub; k0 < ke; k0 += U * kPLPR) {
    int c[U];
    TV v[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int k = k0 + u * kPLPR;
      const bool in = k < ke;
      c[u] = ja[in ? k : kb];
      v[u] = in ? val[k] : (TV)0;
    }
    T g[U];
#pragma unroll
    for(int u = 0; u < U; ++u) g[u] = f(c[u]);
#pragma unroll
    for(int u = 0; u < U; ++u) acc += (T)v[u] * g[u];
  }
#pragma unroll
  for(int o = 1; o < kPLPR; o <<= 1) acc += __shfl_xor(acc, o, 64);
  return acc;
}

// Row sum over LDS-staged entries [k0, k1) in batches of 8 whose loads are
// issued together.
template <typename T, typename TV, typename F>
__device__ __forceinline__ T ldsRow(int k0, int k1, const unsigned short *ja, const TV *val, F f)
{
  T acc = (T)0;
  for(int kb = k0; kb < k1; kb += 8) {
    int c[8];
    TV v[8];
#pragma unroll
    for(int u = 0; u < 8; ++u) {
      const bool in = kb + u < k1;
      c[u] = ja[in ? kb + u : k0];
      v[u] = in ? val[kb + u] : (TV)0;
    }
    T g[8];
#pragma unroll
    for(int u = 0; u < 8; ++u) g[u] = f(c[u]);
#pragma unroll
    for(int u = 0; u < 8; ++u) acc += (T)v[u] * g[u];
  }
  return acc;
}

// Copy of count items into LDS by the whole block, B loads per thread
// issued before their stores.
template <int B, typename TD, typename TS>
__device__ __forceinline__ void stageLds(TD *dst, const TS *__restrict__ src, int count)
{
  for(int base = threadIdx.x; base < count; base += B * blockDim.x) {
    TS v[B];
#pragma unroll
    for(int u = 0; u < B; ++u) {
      const int i = base + u * (int)blockDim.x;
      v[u] = i < count ? src[i] : (TS)0;
    }
#pragma unroll
    for(int u = 0; u < B; ++u) {
      const int i = base + u * (int)blockDim.x;
      if(i < count) dst[i] = (TD)v[u];
    }
  }
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

// rho = (r, z) from the previous up0 partials; p = z + beta p_old and
// q = A z + beta q_old (= A p; the product gathers z only, beta is applied
// after the row sum, so the fold of the partials overlaps the gathers; p and
// q are double-buffered, so an iteration can be redone from a replaced
// residual with the previous direction);
// w = l1 q for the next restriction; partials of (p, q).  Level-0 SELL
// values of type TV, one row per thread.
template <typename T, typename TV, typename TL>
__global__ __launch_bounds__(kFTB) void pcgProduct(int n, const int2 *__restrict__ sell, const int *__restrict__ col,
                                                   const TV *__restrict__ val, const TL *__restrict__ l1,
                                                   const T *__restrict__ z, const T *__restrict__ pOld,
                                                   T *__restrict__ pNew, const T *__restrict__ qOld,
                                                   T *__restrict__ q, T *__restrict__ w,
                                                   const double *__restrict__ rzPart, int Grz,
                                                   const double *__restrict__ rrPart, int Grr, double *scal, int slot,
                                                   int first, double tol2, int *done, int *stat,
                                                   PcgMailbox *mb, int k, double *__restrict__ pqPart)
{
  __shared__ double red[kFTB / 64];
  // independent loads first: flag, row bounds, own-row values, partial shares
  const int dn = *done;
  const int row = blockIdx.x * kFTB + (int)threadIdx.x;
  int2 bl = make_int2(0, 0);
  T zi = (T)0, pi = (T)0, qi = (T)0, li = (T)0;
  if(row < n) {
    bl = sell[row];
    zi = z[row];
    li = (T)l1[row];
    if(!first) {
      pi = pOld[row];
      qi = qOld[row];
    }
  }
  const double rzOwn = partialsShare<kFTB>(rzPart, Grz);
  const double rrOwn = blockIdx.x == 0 ? partialsShare<kFTB>(rrPart, Grr) : 0.;
  const double prevRho = first ? 1. : scal[slot ^ 1];
  if(dn) return;
  const T az = sellSum<T, TV>(col, val, bl, [&](int j) { return z[j]; });
  const double rho = blockAllSum<kFTB>(rzOwn, red);
  double beta = 0.;
  if(!first) beta = rho / prevRho;
  if(blockIdx.x == 0) {
    const double rr = blockAllSum<kFTB>(rrOwn, red);
    if(threadIdx.x == 0) {
      scal[slot] = rho;
      scal[2] = rr;
      const int code = !(rho > 0.) || !(rr == rr) ? 3 : (!first && rr <= tol2 ? 1 : 0);
      if(code) {
        *done = code;
        stat[1] = k;
        mbStop(mb, code, k, rr);
      }
      mbPost(&mb->progress, k);
    }
  }
  const T bt = (T)beta;
  double pq = 0.;
  if(row < n) {
    // first: p = z, q = A z (the previous direction is not read)
    const T p = first ? zi : fmaT<T>(bt, pi, zi);
    const T qn = first ? az : fmaT<T>(bt, qi, az);
    pNew[row] = p;
    q[row] = qn;
    w[row] = li * qn;
    pq = (double)p * (double)qn;
  }
  blockPartial<kFTB>(pq, red, pqPart);
}

// alpha = rho / (p, q); x += alpha p; r_new = r_old - alpha q; u = l1 r_new
// (the first smoothing step); the restriction to level 1 of
// r_new - A u = r_new - (A uOld - alpha A w) with uOld = l1 r_old (stored by
// the previous call) and w = l1 q (stored by the product), so alpha is
// applied after the gathers; tiles of whole aggregates, one row per thread;
// partials of (r_new, r_new).  first: alpha = 0 and A (l1 r_old) is gathered
// directly (x, p, q, w and uOld are not read).
template <typename T, typename TV>
__global__ __launch_bounds__(kFTB) void pcgDown0(PLev<T, TV> L, T *__restrict__ x, const T *__restrict__ p,
                                                 const T *__restrict__ q, const T *__restrict__ w,
                                                 const T *__restrict__ rOld, T *__restrict__ rNew,
                                                 const T *__restrict__ uOld, T *__restrict__ u, T *__restrict__ bNext,
                                                 const double *__restrict__ pqPart, int Gpq, const double *scal,
                                                 int slot, int first, int *done, int *stat, PcgMailbox *mb,
                                                 int k, double *__restrict__ rrPart, unsigned *bar)
{
  if(blockIdx.x == 0 && threadIdx.x == 0) *bar = 0u;  // grid-barrier counter of the fused coarse-level launch
  __shared__ double red[kFTB / 64];
  __shared__ T res[kFRPB];
  // independent loads first: flag, tile, row bounds, own-row values, this
  // thread's aggregate bounds, partial shares
  const int dn = *done;
  const int t = blockIdx.x;
  const int r0 = L.tileStart[t], r1 = L.tileStart[t + 1];
  const int I0 = L.tileAgg[t], I1 = L.tileAgg[t + 1];
  const int row = r0 + (int)threadIdx.x;
  int2 bl = make_int2(0, 0);
  T ro = (T)0, qi = (T)0, pi = (T)0, xi = (T)0, li = (T)0;
  if(row < r1) {
    bl = L.sell[row];
    ro = rOld[row];
    li = (T)L.l1[row];
    if(!first) {
      qi = q[row];
      pi = p[row];
      xi = x[row];
    }
  }
  const int I = I0 + (int)threadIdx.x;  // tiles hold at most kFRPB = kFTB aggregates
  int a0 = 0, a1 = 0;
  if(I < I1) {
    a0 = L.aggStart[I] - r0;
    a1 = L.aggStart[I + 1] - r0;
  }
  const double pqOwn = first ? 0. : partialsShare<kFTB>(pqPart, Gpq);
  const double rho = first ? 0. : scal[slot];
  if(dn) return;
  T s1, s2 = (T)0;
  if(first) {
    s1 = sellSum<T, TV>(L.ja, L.val, bl, [&](int j) { return (T)L.l1[j] * rOld[j]; });
  } else {
    sellSum2<T, TV>(L.ja, L.val, bl, uOld, w, s1, s2);
  }
  double alpha = 0.;
  if(!first) {
    const double pq = blockAllSum<kFTB>(pqOwn, red);
    if(pq > 0.) alpha = rho / pq;
    else if(t == 0 && threadIdx.x == 0) {
      *done = 3;  // (p, A p) <= 0: no update, the host restarts from the FP64 residual
      stat[1] = k;
      mbStop(mb, 3, k, 0.);
    }
  }
  const T a = (T)alpha;
  double rr = 0.;
  if(row < r1) {
    const T rn = first ? ro : fmaT<T>(-a, qi, ro);
    rNew[row] = rn;
    if(!first) x[row] = fmaT<T>(a, pi, xi);
    u[row] = li * rn;
    res[threadIdx.x] = rn - (first ? s1 : fmaT<T>(-a, s2, s1));
    rr = (double)rn * (double)rn;
  }
  __syncthreads();
  if(I < I1) {
    T sum = (T)0;
    for(int m = a0; m < a1; ++m) sum += res[m];
    bNext[I] = sum;
  }
  blockPartial<kFTB>(rr, red, rrPart);
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

// Level l >= 1: y = u + oc e_next[agg]; e = y + S (b - A y).  On level 1
// (y0 given), also the level-0 y0 = u0 + oc e over the rows of each
// aggregate (contiguous in the nested order: aggStart0).
template <typename T, typename TV>
__global__ __launch_bounds__(kPTB) void pcgUpL(PLev<T, TV> L, const T *__restrict__ eNext, T oc, const int *done,
                                               T *__restrict__ y0, const T *__restrict__ u0,
                                               const int *__restrict__ aggStart0)
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
  if(row < L.n) {
    const T y = fmaT<T>(oc, eNext[L.agg[row]], L.u[row]);
    const T e = fmaT<T>((T)L.l1[row], L.b[row] - acc, y);
    if(sub == 0) L.e[row] = e;
    if(y0)
      for(int m = aggStart0[row] + sub; m < aggStart0[row + 1]; m += kPLPR) y0[m] = fmaT<T>(oc, e, u0[m]);
  }
}

// Level 0: z = y + S (r - A y) with y = u + oc P e_1 written by the level-1
// prolongation; partials of (r, z).  One row per thread.
template <typename T, typename TV>
__global__ __launch_bounds__(kFTB) void pcgUp0(PLev<T, TV> L, const T *__restrict__ y, const T *__restrict__ r,
                                               T *__restrict__ z, const int *done, double *__restrict__ rzPart)
{
  __shared__ double red[kFTB / 64];
  const int dn = *done;
  const int row = blockIdx.x * kFTB + (int)threadIdx.x;
  int2 bl = make_int2(0, 0);
  T ri = (T)0, yi = (T)0, li = (T)0;
  if(row < L.n) {
    bl = L.sell[row];
    ri = r[row];
    yi = y[row];
    li = (T)L.l1[row];
  }
  if(dn) return;
  const T acc = sellSum<T, TV>(L.ja, L.val, bl, [&](int j) { return y[j]; });
  double rz = 0.;
  if(row < L.n) {
    const T zi = fmaT<T>(li, ri - acc, yi);
    z[row] = zi;
    rz = (double)ri * (double)zi;
  }
  blockPartial<kFTB>(rz, red, rzPart);
}

// Levels from .. coarsest in one workgroup, staged in LDS: b of level from is
// read from global memory, its correction e written back.
// The static part of the small-level kernel's LDS (matrices with 16-bit
// columns, values, l1, aggregates, the coarsest inverse) packed into one
// global image in the LDS layout, once per numeric set-up.
template <typename T, typename TV>
__global__ __launch_bounds__(kPBlockTB) void pcgPackBlock(const PLev<T, TV> *__restrict__ lv, PBlockPlan pl,
                                                          const T *__restrict__ cinvG, unsigned char *img)
{
  const int nLev = pl.coarsest - pl.from;
  for(int s = 0; s < nLev; ++s) {
    const PLev<T, TV> L = lv[pl.from + s];
    stageLds<2>((int *)(img + pl.oIa[s]), L.ia, pl.n[s] + 1);
    stageLds<8>((unsigned short *)(img + pl.oJa[s]), L.ja, pl.nnz[s]);
    stageLds<8>((TV *)(img + pl.oVal[s]), L.val, pl.nnz[s]);
    stageLds<2>((TV *)(img + pl.oL1[s]), L.l1, pl.n[s]);
    stageLds<2>((unsigned short *)(img + pl.oAgg[s]), L.agg, pl.n[s]);
    stageLds<2>((unsigned short *)(img + pl.oAs[s]), L.aggStart, pl.nc[s] + 1);
  }
  const int nC = pl.n[nLev];
  stageLds<2>((T *)(img + pl.oCi), cinvG, nC * nC);
}

// Static image of the single-workgroup levels into LDS: 16-byte loads, all
// of a batch issued before their stores.
__device__ __forceinline__ void blockImage(const PBlockPlan &pl, const unsigned char *__restrict__ img,
                                           unsigned char *lds)
{
  constexpr int U = 10;  // 16-byte chunks per thread per batch (160 KB per batch)
  const int t = threadIdx.x, nt = blockDim.x;
  const int chunks = pl.imgBytes / 16;
  for(int c0 = t; c0 < chunks; c0 += U * nt) {
    uint4 v[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int c = c0 + u * nt;
      v[u] = c < chunks ? ((const uint4 *)img)[c] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for(int u = 0; u < U; ++u) {
      const int c = c0 + u * nt;
      if(c < chunks) ((uint4 *)lds)[c] = v[u];
    }
  }
}

// Levels from .. coarsest with the static image in LDS: b of level from is
// read from global memory, its correction e written back (and, from = 1,
// the level-0 y0 = u0 + oc P e_1).
template <typename T, typename TV>
__device__ __forceinline__ void blockSolve(const PLev<T, TV> *__restrict__ lv, const PBlockPlan &pl,
                                           unsigned char *lds, T oc, T *__restrict__ y0, const T *__restrict__ u0)
{
  const int t = threadIdx.x, nt = blockDim.x;
  const int nLev = pl.coarsest - pl.from;
  stageLds<2>((T *)(lds + pl.oB[0]), lv[pl.from].b, pl.n[0]);
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
      const T acc = ldsRow<T>(ia[i], ia[i + 1], ja, val, [&](int j) { return (T)l1[j] * b[j]; });
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
    const T *cinv = (const T *)(lds + pl.oCi);  // from the image
    T *x = (T *)(lds + pl.oCx);
    for(int i = t; i < nC; i += nt) {
      T acc = (T)0;
#pragma unroll 8
      for(int j = 0; j < nC; ++j) acc += cinv[i * nC + j] * b[j];
      x[i] = acc;
      if(nLev == 0) lv[pl.from].e[i] = acc;
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
    T *e = (T *)(lds + pl.oE[s]);
    T *eg = s == 0 ? lv[pl.from].e : nullptr;
    for(int i = t; i < n; i += nt) y[i] = fmaT<T>(oc, en[agg[i]], u[i]);
    __syncthreads();
    for(int i = t; i < n; i += nt) {
      const T acc = ldsRow<T>(ia[i], ia[i + 1], ja, val, [&](int j) { return y[j]; });
      const T ei = fmaT<T>((T)l1[i], b[i] - acc, y[i]);
      e[i] = ei;
      if(eg) eg[i] = ei;
    }
    __syncthreads();
  }
  if(pl.from == 1 && y0) {
    // level 1 is in this kernel: the level-0 y0 = u0 + oc P e_1
    const T *e1 = (const T *)(lds + (nLev == 0 ? pl.oCx : pl.oE[0]));
    const int n0 = lv[0].n;
    const int *agg0 = lv[0].agg;
    for(int m = t; m < n0; m += nt) y0[m] = fmaT<T>(oc, e1[agg0[m]], u0[m]);
  }
}

template <typename T, typename TV>
__global__ __launch_bounds__(kPBlockTB) void pcgBlock(const PLev<T, TV> *__restrict__ lv, PBlockPlan pl,
                                                      const unsigned char *__restrict__ img, T oc, const int *done,
                                                      T *__restrict__ y0, const T *__restrict__ u0)
{
  __shared__ __align__(16) unsigned char lds[kPBlockLds];
  const int dn = *done;
  blockImage(pl, img, lds);
  if(dn) return;
  blockSolve<T, TV>(lv, pl, lds, oc, y0, u0);
}

// ---------------------------------------------------------------------------
// Levels F .. coarsest in one launch
//
// The restriction and prolongation of the levels F .. B-1 above the
// single-workgroup ones and the single-workgroup levels (B .. coarsest) run in
// one launch of G workgroups separated by grid barriers, instead of
// 2 (B - F) + 1 dependent launches; F is the first level whose shares fit one
// round of the workgroup (levels 1 .. F-1 keep their own launches, where the
// rows of a level are spread over many more workgroups).  Workgroup w owns, on every level F <= l < B,
// the rows nested in its contiguous share of level-B rows (tiles[l][w] ..
// tiles[l][w + 1]): its restrictions are sums inside its own rows, and the
// neighbour values a row needs (l1 b, u + oc e_next) are recomputed from
// vectors completed before the last barrier.  Workgroup 0 loads the
// single-workgroup LDS image first and solves levels B .. coarsest between
// the down and up sweeps.  The grid is small (G <= kFuseMaxG) and every
// workgroup holds one CU, so all are resident; each barrier wait is bounded
// and a timeout raises a flag (no wait is endless).
constexpr int kFuseMaxG = 64;
constexpr int kFuseRows = kPBlockTB / kPLPR;  // rows of one level per workgroup: one round of kPLPR lanes per row
constexpr int kFuseLds = kPBlockLds + kFuseRows * 8;
constexpr int kFuseSpin = 1 << 21;

struct FusePlan {
  int G = 0, B = 0;  // G = 0: not used
};

__device__ __forceinline__ void gridBarrier(unsigned *cnt, unsigned target, int *err)
{
  __syncthreads();
  if(threadIdx.x == 0) {
    __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0;
    while(__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if(++spins > kFuseSpin) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
}

// tiles: (B + 1) x (G + 1) row starts, level l at l * (G + 1).  bar: counter
// zeroed by the preceding down0 launch.  err: flag raised on a barrier timeout
// (host-mapped).  y0 = u0 + oc P e_1 for the level-0 rows of the own level-1
// rows.
template <typename T, typename TV>
__global__ __launch_bounds__(kPBlockTB) void pcgCoarseV(const PLev<T, TV> *__restrict__ lv, PBlockPlan pl,
                                                        const unsigned char *__restrict__ img,
                                                        const int *__restrict__ tiles, int G, int F, T oc,
                                                        const int *done,
                                                        unsigned *bar, int *err, T *__restrict__ y0,
                                                        const T *__restrict__ u0, const int *__restrict__ aggStart0)
{
  __shared__ __align__(16) unsigned char lds[kFuseLds];
  const int t = threadIdx.x, w = blockIdx.x, B = pl.from;
  const int dn = *done;
  if(w == 0) blockImage(pl, img, lds);
  if(dn) return;
  T *res = (T *)(lds + kPBlockLds);
  const int rl = t / kPLPR, sub = t % kPLPR;
  constexpr int RPR = kPBlockTB / kPLPR;  // rows per round
  unsigned target = 0;
  for(int l = F; l < B; ++l) {  // restrictions: b_l+1 = P^T (b_l - A (l1 b_l))
    const PLev<T, TV> L = lv[l];
    T *bn = lv[l + 1].b;
    const int r0 = tiles[l * (G + 1) + w], r1 = tiles[l * (G + 1) + w + 1];
    for(int base = r0; base < r1; base += RPR) {
      const int row = base + rl;
      int kb = 0, ke = 0;
      if(row < r1) {
        kb = L.ia[row];
        ke = L.ia[row + 1];
      }
      const T acc = rowSum<T, TV>(L.ja, L.val, kb, ke, sub, [&](int j) { return (T)L.l1[j] * L.b[j]; });
      if(row < r1 && sub == 0) {
        const T bi = L.b[row];
        L.u[row] = (T)L.l1[row] * bi;
        res[row - r0] = bi - acc;
      }
    }
    __syncthreads();
    const int I0 = tiles[(l + 1) * (G + 1) + w], I1 = tiles[(l + 1) * (G + 1) + w + 1];
    for(int I = I0 + t; I < I1; I += kPBlockTB) {
      T s = (T)0;
      for(int m = L.aggStart[I] - r0; m < L.aggStart[I + 1] - r0; ++m) s += res[m];
      bn[I] = s;
    }
    gridBarrier(bar, target += G, err);
  }
  if(w == 0) blockSolve<T, TV>(lv, pl, lds, oc, (T *)nullptr, (const T *)nullptr);
  gridBarrier(bar, target += G, err);
  for(int l = B - 1; l >= F; --l) {  // prolongations: e_l = y + S (b_l - A y), y = u + oc P e_l+1
    const PLev<T, TV> L = lv[l];
    const T *eNext = lv[l + 1].e;
    const int r0 = tiles[l * (G + 1) + w], r1 = tiles[l * (G + 1) + w + 1];
    for(int base = r0; base < r1; base += RPR) {
      const int row = base + rl;
      int kb = 0, ke = 0;
      if(row < r1) {
        kb = L.ia[row];
        ke = L.ia[row + 1];
      }
      const T acc =
          rowSum<T, TV>(L.ja, L.val, kb, ke, sub, [&](int j) { return fmaT<T>(oc, eNext[L.agg[j]], L.u[j]); });
      if(row < r1) {
        const T y = fmaT<T>(oc, eNext[L.agg[row]], L.u[row]);
        const T e = fmaT<T>((T)L.l1[row], L.b[row] - acc, y);
        if(l > 1) {
          if(sub == 0) L.e[row] = e;
        } else {
          for(int m = aggStart0[row] + sub; m < aggStart0[row + 1]; m += kPLPR) y0[m] = fmaT<T>(oc, e, u0[m]);
        }
      }
    }
    if(l > F) gridBarrier(bar, target += G, err);
  }
}

// ---------------------------------------------------------------------------
// W-cycle of the coarse levels in one workgroup
//
// Levels R .. D-1 (one or two "register levels") are each solved by two
// cycles, e = e1 + C(b - A e1), C the cycle of the level over the solve of the
// next one; level D (at most kWDense rows) by its dense inverse.  One launch
// replaces the restriction, single-workgroup and prolongation launches of
// those levels, and the second cycles cut the PCG iterations (the coarse
// problem of level R - 1 is solved almost exactly).
//
// A register level's matrix is held in the registers of the workgroup: its
// rows, each padded to whole groups of kWGroup entries, are cut into runs of
// K entries, one run per thread (slots 4g .. 4g + 3 of thread t are the
// 16-byte word g * kWTB + t).  A word holds the LDS byte address of the
// gathered entry of the level's X vector (low 16 bits) and the binary16 value
// h_ij = a_ij w_i (high 16 bits), w the level's l1 weights rounded to
// bfloat16; the operator is diag(w)^-1 H: r = b - (H x) / w (the reciprocal
// of w to 1 ulp), e = y + w b - H y.  Per thread: its first row, whether that row began in the
// previous thread, and one bit per group where a row ends.  A row spans at
// most two threads (K >= its padded length); the two partial sums are added
// after a barrier, in a fixed order.  Vectors, weights, aggregate starts and
// the dense inverse are in LDS.
constexpr int kWTB = 512;          // threads of the W-cycle kernel
constexpr int kWGroup = 4;         // row entries per group (rows end on a group boundary)
constexpr int kWLds = 160 * 1024;  // LDS of the W-cycle kernel
constexpr int kWDense = 128;       // its dense level: at most this many rows
constexpr int kWKA[2] = {64, 160};  // slots per thread of the first register level (kernel variants)
constexpr int kWKB = 32;            // slots per thread of the second register level

struct WLev {
  int n = 0, nc = 0;                    // rows; rows of the next level (= aggregates)
  int oX = 0, oB = 0, oE = 0, oR = 0;   // LDS byte offsets of the float vectors
  int oW = 0, oAs = 0;                  // bfloat16 weights, nc + 1 aggregate starts (static image)
};
// Steps of the W-cycle, run in order by the kernel (ops[k] = step | level << 4);
// one instance of each level's pass code serves all its passes, which keeps
// the kernel inside the instruction cache.
enum WStep : unsigned char {
  kWScaleB,      // X = w B                                (u of the first cycle)
  kWPassE,       // E = S = H X (raw row sums)
  kWPassR,       // R = S
  kWPassB,       // B = S
  kWRestrictE,   // next rhs = P^T (B - E / w)             (r = b - A u)
  kWRestrictB,   // next rhs = P^T (R - B / w)             (r' = b' - A u')
  kWProlong,     // X += oc P e_next                       (y)
  kWPost1,       // E = X + w B - E; X = E                 (e1)
  kWResRScale,   // R = B - R / w; X = w R                 (b' = b - A e1, u')
  kWPost2,       // E += X + w R - B                       (e1 + e2)
  kWDenseSolve   // e_D = Dinv b_D
};
constexpr int kWMaxOps = 48;

struct WPlan {
  int R = 0, D = 0, nReg = 0, KA = 0, KB = 0;  // R = 0: the kernel is not used
  WLev lv[2];
  int nOps = 0;
  unsigned char ops[kWMaxOps];
  int nD = 0, oDinv = 0, oBD = 0, oED = 0, oCarry = 0;
  int oImg = 0, imgBytes = 0, bytes = 0;  // static image: LDS bytes [oImg, oImg + imgBytes)
};

__device__ __forceinline__ float wBf16(unsigned short b)
{
  return __uint_as_float((uint32_t)b << 16);
}
__device__ __forceinline__ unsigned short bf16Bits(float x)  // nearest, ties to even
{
  const uint32_t u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Decoding of a register word, kept inside each pass: the words stay in
// registers for the whole kernel, and decoded copies hoisted out of the
// passes would need registers of their own (volatile, so not merged across
// passes; ALU only).
__device__ __forceinline__ uint32_t wAddr(uint32_t w)  // LDS byte address (low 16 bits)
{
  uint32_t a;
  asm volatile("v_and_b32 %0, 0xffff, %1" : "=v"(a) : "v"(w));
  return a;
}
__device__ __forceinline__ float wFma(uint32_t w, float x, float acc)  // acc + h x, h the high binary16
{
  asm volatile("v_fma_mix_f32 %0, %1, %2, %0 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(acc) : "v"(w), "v"(x));
  return acc;
}
__device__ __forceinline__ uint32_t wBit(uint32_t m, uint32_t bit)
{
  uint32_t r;
  asm volatile("v_and_b32 %0, %1, %2" : "=v"(r) : "v"(m), "v"(bit));
  return r;
}

// S_i = sum_j h_ij X_j over the thread's run, stored to dst[i] (the row the
// thread continues after the barrier).  Gathers run one group ahead; the
// row ends only store.
template <int K>
__device__ __forceinline__ void wPass(const uint4 (&w)[K / kWGroup], int4 info, const unsigned char *lds, float *carry,
                                      float *dst)
{
  constexpr int G = K / kWGroup;
  const bool cont = info.y != 0;
  int row = info.x;
  float acc = 0.f, head = 0.f;
  float xs[2][kWGroup];
  auto gather = [&](const uint4 &q, float *x) {
    x[0] = *(const float *)(lds + wAddr(q.x));
    x[1] = *(const float *)(lds + wAddr(q.y));
    x[2] = *(const float *)(lds + wAddr(q.z));
    x[3] = *(const float *)(lds + wAddr(q.w));
  };
  gather(w[0], xs[0]);
#pragma unroll
  for(int g = 0; g < G; ++g) {
    if(g + 1 < G) gather(w[g + 1], xs[(g + 1) & 1]);
    acc = wFma(w[g].x, xs[g & 1][0], acc);
    acc = wFma(w[g].y, xs[g & 1][1], acc);
    acc = wFma(w[g].z, xs[g & 1][2], acc);
    acc = wFma(w[g].w, xs[g & 1][3], acc);
    if(wBit(g < 32 ? (uint32_t)info.z : (uint32_t)info.w, 1u << (g & 31))) {
      if(cont && row == info.x) head = acc;
      else dst[row] = acc;
      ++row;
      acc = 0.f;
    }
  }
  carry[threadIdx.x] = acc;
  __syncthreads();
  if(cont) dst[info.x] = carry[threadIdx.x - 1] + head;
  __syncthreads();
}

// Rows i = t, t + kWTB, ... of a level in batches of kWU: every load of a
// batch is issued before its stores (LDS latency once per batch).
constexpr int kWU = 4;
template <int NV, typename Load, typename Store>
__device__ __forceinline__ void wRows(int n, Load load, Store store)
{
  for(int base = threadIdx.x; base < n; base += kWU * kWTB) {
    float v[kWU][NV];
#pragma unroll
    for(int u = 0; u < kWU; ++u) {
      const int i = base + u * kWTB;
      if(i < n) load(i, v[u]);
    }
#pragma unroll
    for(int u = 0; u < kWU; ++u) {
      const int i = base + u * kWTB;
      if(i < n) store(i, v[u]);
    }
  }
  __syncthreads();
}

// Aggregates I = t, t + kWTB, ...: f(I, i0, i1) over member rows [i0, i1).
template <typename F>
__device__ __forceinline__ void wAggs(int nc, const int *as, F f)
{
  for(int I = threadIdx.x; I < nc; I += kWTB) f(I, as[I], as[I + 1]);
  __syncthreads();
}

// e = Dinv b on the dense level, 4 lanes per row, 8 entries per batch.
__device__ __forceinline__ void wDense(const float *Dinv, int nD, const float *b, float *e)
{
  const int t = threadIdx.x, i = t >> 2, sub = t & 3;
  float s = 0.f;
  if(i < nD)
    for(int j0 = sub; j0 < nD; j0 += 32) {
      float d[8], x[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) {
        const int j = j0 + 4 * u;
        d[u] = j < nD ? Dinv[i * nD + j] : 0.f;
        x[u] = j < nD ? b[j] : 0.f;
      }
#pragma unroll
      for(int u = 0; u < 8; ++u) s = __builtin_fmaf(d[u], x[u], s);
    }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if(i < nD && sub == 0) e[i] = s;
  __syncthreads();
}

// The W-cycle kernel: b_R (T) in, e_R (T) out; with R = 1 also the level-0
// y0 = u0 + oc P e_1.  KB = 0: level R + 1 is the dense level.  Each
// register level is solved by two cycles on its rhs B (solution E, work
// vectors R and X), the next level's rhs and solution being its B and E or
// the dense level's b_D and e_D; the steps come from the plan's op list.
template <typename T, int KA, int KB>
__global__ __launch_bounds__(kWTB) void pcgCoarseW(WPlan pl, const uint4 *__restrict__ wA,
                                                   const int4 *__restrict__ infoA, const uint4 *__restrict__ wB,
                                                   const int4 *__restrict__ infoB, const unsigned char *__restrict__ img,
                                                   float oc, const int *done, const T *__restrict__ bR,
                                                   T *__restrict__ eR, T *__restrict__ y0, const T *__restrict__ u0,
                                                   const int *__restrict__ aggStart0)
{
  __shared__ __align__(16) unsigned char lds[kWLds];
  const int t = threadIdx.x;
  const int dn = *done;
  uint4 a[KA / kWGroup];
#pragma unroll
  for(int g = 0; g < KA / kWGroup; ++g) a[g] = wA[g * kWTB + t];
  constexpr int GB = KB > 0 ? KB / kWGroup : 1;
  uint4 b[GB];
  int4 ib = make_int4(0, 0, 0, 0);
  if constexpr(KB > 0) {
#pragma unroll
    for(int g = 0; g < GB; ++g) b[g] = wB[g * kWTB + t];
    ib = infoB[t];
  }
  const int4 ia = infoA[t];
  {
    constexpr int U = 4;
    const int chunks = pl.imgBytes / 16;
    uint4 *dst = (uint4 *)(lds + pl.oImg);
    for(int c0 = t; c0 < chunks; c0 += U * kWTB) {
      uint4 v[U];
#pragma unroll
      for(int u = 0; u < U; ++u) {
        const int c = c0 + u * kWTB;
        v[u] = c < chunks ? ((const uint4 *)img)[c] : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for(int u = 0; u < U; ++u)
        if(c0 + u * kWTB < chunks) dst[c0 + u * kWTB] = v[u];
    }
  }
  if(dn) return;
  for(int i = t; i < pl.lv[0].n; i += kWTB) ((float *)(lds + pl.lv[0].oB))[i] = (float)bR[i];
  __syncthreads();
  float *carry = (float *)(lds + pl.oCarry);
  float *bD = (float *)(lds + pl.oBD), *eD = (float *)(lds + pl.oED);
  for(int k = 0; k < pl.nOps; ++k) {
    const int step = pl.ops[k] & 15, L = pl.ops[k] >> 4;
    if(step == kWDenseSolve) {
      wDense((const float *)(lds + pl.oDinv), pl.nD, bD, eD);
      continue;
    }
    const WLev &V = pl.lv[L];
    const int n = V.n, nc = V.nc;
    float *X = (float *)(lds + V.oX), *B = (float *)(lds + V.oB), *E = (float *)(lds + V.oE), *R = (float *)(lds + V.oR);
    const unsigned short *wt = (const unsigned short *)(lds + V.oW);
    const int *as = (const int *)(lds + V.oAs);
    const bool regChild = L + 1 < pl.nReg;
    float *cB = regChild ? (float *)(lds + pl.lv[L + 1].oB) : bD;
    const float *cE = regChild ? (const float *)(lds + pl.lv[L + 1].oE) : eD;
    switch(step) {
    case kWScaleB:
      wRows<2>(
          n, [&](int i, float *v) { v[0] = wBf16(wt[i]); v[1] = B[i]; }, [&](int i, const float *v) { X[i] = v[0] * v[1]; });
      break;
    case kWPost1:
      wRows<4>(
          n, [&](int i, float *v) { v[0] = wBf16(wt[i]); v[1] = B[i]; v[2] = X[i]; v[3] = E[i]; },
          [&](int i, const float *v) {
            const float e = v[2] + v[0] * v[1] - v[3];
            E[i] = e;
            X[i] = e;
          });
      break;
    case kWResRScale:
      wRows<3>(
          n, [&](int i, float *v) { v[0] = wBf16(wt[i]); v[1] = B[i]; v[2] = R[i]; },
          [&](int i, const float *v) {
            const float r = v[1] - v[2] * __builtin_amdgcn_rcpf(v[0]);
            R[i] = r;
            X[i] = v[0] * r;
          });
      break;
    case kWPost2:
      wRows<5>(
          n, [&](int i, float *v) { v[0] = wBf16(wt[i]); v[1] = R[i]; v[2] = X[i]; v[3] = B[i]; v[4] = E[i]; },
          [&](int i, const float *v) { E[i] = v[4] + (v[2] + v[0] * v[1] - v[3]); });
      break;
    case kWRestrictE:
    case kWRestrictB: {
      const float *Y = step == kWRestrictE ? B : R, *Sm = step == kWRestrictE ? E : B;
      wAggs(nc, as, [&](int I, int i0, int i1) {
        float s = 0.f;
        for(int b0 = i0; b0 < i1; b0 += 8) {
          float y[8], sv[8], w[8];
#pragma unroll
          for(int u = 0; u < 8; ++u) {
            const int i = b0 + u < i1 ? b0 + u : i0;
            y[u] = Y[i];
            sv[u] = Sm[i];
            w[u] = wBf16(wt[i]);
          }
#pragma unroll
          for(int u = 0; u < 8; ++u)
            if(b0 + u < i1) s += y[u] - sv[u] * __builtin_amdgcn_rcpf(w[u]);
        }
        cB[I] = s;
      });
      break;
    }
    case kWProlong:
      wAggs(nc, as, [&](int I, int i0, int i1) {
        const float e = oc * cE[I];
        for(int b0 = i0; b0 < i1; b0 += 8) {
          float x[8];
#pragma unroll
          for(int u = 0; u < 8; ++u) x[u] = X[b0 + u < i1 ? b0 + u : i0];
#pragma unroll
          for(int u = 0; u < 8; ++u)
            if(b0 + u < i1) X[b0 + u] = x[u] + e;
        }
      });
      break;
    default: {  // a pass over the level's matrix: raw sums into E, R or B
      float *dst = step == kWPassE ? E : step == kWPassR ? R : B;
      if constexpr(KB > 0) {
        if(L == 1) {
          wPass<KB>(b, ib, lds, carry, dst);
          break;
        }
      }
      wPass<KA>(a, ia, lds, carry, dst);
      break;
    }
    }
  }
  const float *E = (const float *)(lds + pl.lv[0].oE);
  for(int i = t; i < pl.lv[0].n; i += kWTB) eR[i] = (T)E[i];
  if(y0)
    for(int I = t; I < pl.lv[0].n; I += kWTB) {
      const T e = (T)oc * (T)E[I];
      for(int m = aggStart0[I]; m < aggStart0[I + 1]; ++m) y0[m] = u0[m] + e;
    }
}

// Register words and static image of one register level from the FP32
// values and l1 of the level (every numeric set-up): map[q] = (CSR entry or
// -1, row | byte address << 16) per word slot q.
__global__ void pcgWPack(int nq, const int2 *__restrict__ map, const float *__restrict__ val32,
                         const float *__restrict__ l132, uint32_t *__restrict__ words, int n,
                         unsigned short *__restrict__ wImg, int nc1, const int *__restrict__ as, int *__restrict__ asImg)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if(q < nq) {
    const int2 m = map[q];
    uint32_t word = (uint32_t)m.y >> 16;  // padding: value 0 at the level's X[0]
    if(m.x >= 0) {
      const float w = wBf16(bf16Bits(l132[m.y & 0xFFFF]));
      const _Float16 h = (_Float16)(val32[m.x] * w);
      word |= (uint32_t)__builtin_bit_cast(unsigned short, h) << 16;
    }
    words[q] = word;
  }
  if(q < n) wImg[q] = bf16Bits(l132[q]);
  if(q < nc1) asImg[q] = as[q];
}

// FP32 inverse of the dense level (n <= kWDense) by Gauss-Jordan in FP64
// (SPD: no pivoting) with the matrix in registers: thread t holds row t >> 3,
// columns (t & 7) + 8k; each step's pivot column and normalised pivot row go
// through LDS.  flag = 1 on a non-positive pivot.
__global__ __launch_bounds__(1024) void pcgWDenseInverse(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                         const double *__restrict__ val, float *__restrict__ inv,
                                                         int *flag)
{
  __shared__ double Md[kWDense * kWDense];
  __shared__ double rowP[kWDense], colP[kWDense];
  const int t = threadIdx.x, i = t >> 3, c0 = t & 7;
  constexpr int KC = kWDense / 8;
  for(int q = t; q < n * n; q += 1024) Md[q] = 0.;
  __syncthreads();
  for(int r = t; r < n; r += 1024)
    for(int k = ia[r]; k < ia[r + 1]; ++k) Md[r * n + ja[k]] = val[k];
  __syncthreads();
  double m[KC];
#pragma unroll
  for(int k = 0; k < KC; ++k) {
    const int c = c0 + 8 * k;
    m[k] = i < n && c < n ? Md[i * n + c] : 0.;
  }
  for(int p = 0; p < n; ++p) {
    double mp = 0.;
#pragma unroll
    for(int k = 0; k < KC; ++k)
      if(k == (p >> 3)) mp = m[k];
    if(c0 == (p & 7) && i < n) colP[i] = mp;  // column p before the step
    __syncthreads();
    const double piv = colP[p];
    if(!(piv > 0.)) {
      if(t == 0) *flag = 1;
      return;
    }
    if(i == p) {
      const double rp = 1. / piv;
#pragma unroll
      for(int k = 0; k < KC; ++k) {
        const int c = c0 + 8 * k;
        m[k] = (c == p ? 1. : m[k]) * rp;
        if(c < n) rowP[c] = m[k];
      }
    }
    __syncthreads();
    if(i != p && i < n) {
      const double f = colP[i];
#pragma unroll
      for(int k = 0; k < KC; ++k) {
        const int c = c0 + 8 * k;
        if(c < n) m[k] = (c == p ? 0. : m[k]) - f * rowP[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for(int k = 0; k < KC; ++k) {
    const int c = c0 + 8 * k;
    if(i < n && c < n) inv[i * n + c] = (float)m[k];
  }
}

// ---------------------------------------------------------------------------
// Numeric set-up

// Level 0: the SELL values (the original CSR entry of every SELL entry),
// one row per thread, and l1 = c / sum |a_ij|, in FP32 (W32) and / or FP64
// (W64).
template <bool W32, bool W64>
__global__ __launch_bounds__(kPTB) void pcgSetup0(int n, const int2 *__restrict__ sell, const int *__restrict__ map,
                                                  const double *__restrict__ val, double c,
                                                  float *__restrict__ val32, float *__restrict__ l1_32,
                                                  double *__restrict__ val64, double *__restrict__ l1_64)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if(row >= n) return;
  const int2 bl = sell[row];
  double s = 0.;
  for(int k = 0; k < bl.y; ++k) {
    const int q = bl.x + k * kSlice;
    const double v = val[map[q]];
    if(W32) val32[q] = (float)v;
    if(W64) val64[q] = v;
    s += fabs(v);
  }
  if(W32) l1_32[row] = (float)(c / s);
  if(W64) l1_64[row] = c / s;
}

// Level l >= 1: one thread per entry, the Galerkin entry as the fixed-order
// sum of the previous level's FP64 entries and its TV copy; the same launch
// sets l1 = c / sum |a_ij| of the previous level's rows when that level is a
// coarse one (its entries were finished by the previous launch; pn = 0 for
// level 1, whose previous level is the system).
template <typename TV>
__global__ __launch_bounds__(kPTB) void pcgSetupL(int nnz, const int *__restrict__ cptr, const int *__restrict__ clist,
                                                  const double *__restrict__ src, double *__restrict__ val64,
                                                  TV *__restrict__ valT, int pn, const int *__restrict__ pia,
                                                  double c, TV *__restrict__ pl1)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t < nnz) {
    // batches of 8 contributions: list entries, then their values, in flight together
    double v = 0.;
    const int m0 = cptr[t], m1 = cptr[t + 1];
    for(int mb = m0; mb < m1; mb += 8) {
      int c8[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) c8[u] = clist[mb + u < m1 ? mb + u : m0];
      double s8[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) s8[u] = src[c8[u]];
#pragma unroll
      for(int u = 0; u < 8; ++u)
        if(mb + u < m1) v += s8[u];
    }
    val64[t] = v;
    valT[t] = (TV)v;
  }
  if(t < pn) {
    double s = 0.;
    const int k0 = pia[t], k1 = pia[t + 1];
    for(int kb = k0; kb < k1; kb += 8) {
      double s8[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) s8[u] = src[kb + u < k1 ? kb + u : k0];
#pragma unroll
      for(int u = 0; u < 8; ++u)
        if(kb + u < k1) s += fabs(s8[u]);
    }
    pl1[t] = (TV)(c / s);
  }
}

// Dense inverse of the coarsest level (n <= kPDense) by Gauss-Jordan in LDS
// (FP64, SPD: no pivoting); flag = 1 on a non-positive pivot.
template <typename TV>
__global__ __launch_bounds__(1024) void pcgDenseInverse(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                        const double *__restrict__ val, double *__restrict__ inv64,
                                                        float *__restrict__ inv32, int *flag, double c,
                                                        TV *__restrict__ l1)
{
  __shared__ double M[kPDense][2 * kPDense + 1];
  __shared__ double f[kPDense];
  const int t = threadIdx.x, nt = blockDim.x, w = 2 * n;
  for(int i = t; i < n; i += nt) {  // l1 of the coarsest level
    double s = 0.;
    for(int k = ia[i]; k < ia[i + 1]; ++k) s += fabs(val[k]);
    l1[i] = (TV)(c / s);
  }
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
  int64_t nnz = 0;   // level 0: SELL entries (padding included)
  int2 *sell = nullptr;  // level 0: (first SELL position, length) per row
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
  int *map = nullptr;   // level-0 SELL entry -> original CSR entry
  int2 *sell = nullptr;  // level-0 rows: (first SELL position, length)
  float *cinv32 = nullptr;
  double *cinv64 = nullptr;
  int *flag = nullptr;  // non-positive coarsest pivot
  PBlockPlan plan[3];   // (vectors, values): 0 float/float, 1 double/float, 2 double/double
  unsigned char *img[3] = {nullptr, nullptr, nullptr};  // packed static LDS image per plan
  void *desc[3] = {nullptr, nullptr, nullptr};
  // level-0 solver vectors (permuted), one set per vector type
  float *x32 = nullptr, *r32[2] = {}, *p32[2] = {}, *u32[2] = {}, *q32[2] = {}, *z32 = nullptr, *rhs32 = nullptr,
        *y32 = nullptr, *w32 = nullptr;
  double *x64v = nullptr, *r64v[2] = {}, *p64v[2] = {}, *u64v[2] = {}, *q64v[2] = {}, *z64v = nullptr,
         *rhs64v = nullptr, *y64v = nullptr, *w64v = nullptr;
  double *pq = nullptr, *rr = nullptr, *rz = nullptr, *scal = nullptr, *xsave = nullptr;
  int *done = nullptr, *stat = nullptr;
  PcgMailbox *mbHost = nullptr, *mbDev = nullptr;  // pinned, mapped: host and device addresses
  // W-cycle kernel of the coarse levels (wplan.R = 0: not used): per register
  // level the words, per-thread (first row, continued, group-end mask) and
  // the word -> (entry, row | byte address << 16) map; the static LDS image
  WPlan wplan;
  uint32_t *wWords[2] = {nullptr, nullptr};
  int4 *wInfo[2] = {nullptr, nullptr};
  int2 *wMap[2] = {nullptr, nullptr};
  unsigned char *wImg = nullptr;
  // fused coarse-level launch per slot (fuseG = 0: not used): tiles of
  // (B + 1) x (G + 1) row starts; the grid-barrier counter
  int fuseG[3] = {0, 0, 0}, fuseF[3] = {0, 0, 0};
  int *fuseTiles[3] = {nullptr, nullptr, nullptr};
  unsigned *bar = nullptr;
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
  PLev<float, float> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.val32, h.l1_32, h.sell, h.agg, h.aggStart, h.tileStart,
                       h.tileAgg, h.b32, h.u32, h.e32};
  return L;
}
template <>
PLev<double, float> pview<double, float>(const FeAmgPcgLevel &h)
{
  PLev<double, float> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.val32, h.l1_32, h.sell, h.agg, h.aggStart, h.tileStart,
                        h.tileAgg, h.b64, h.u64, h.e64};
  return L;
}
template <>
PLev<double, double> pview<double, double>(const FeAmgPcgLevel &h)
{
  PLev<double, double> L{h.n, h.nc, h.nTiles, h.ia, h.ja, h.valD, h.l1_64, h.sell, h.agg, h.aggStart, h.tileStart,
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
    const int nC = A.lv[Lc].n;
    for(int l = from; l < Lc; ++l) {  // static arrays
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
    }
    pl.oCi = take(sizeof(T) * nC * nC);
    pl.imgBytes = bytes;
    for(int l = from; l < Lc; ++l) {  // vectors
      const int s = l - from;
      const FeAmgPcgLevel &h = A.lv[l];
      pl.oB[s] = take(sizeof(T) * h.n);
      pl.oU[s] = take(sizeof(T) * h.n);
      pl.oY[s] = take(sizeof(T) * h.n);
      pl.oE[s] = take(sizeof(T) * h.n);
    }
    pl.n[Lc - from] = nC;
    pl.oB[Lc - from] = take(sizeof(T) * nC);
    pl.oCx = take(sizeof(T) * nC);
    pl.bytes = bytes;
    if(bytes <= kPBlockLds) return pl;
  }
  pl.bytes = kPBlockLds + 1;  // nothing fits
  return pl;
}

// Steps of the two cycles on register level L (the next level solved in
// between, recursively, or the dense level).
void wProgram(WPlan &pl, int L)
{
  auto add = [&](int step, int lev) { pl.ops[pl.nOps++] = (unsigned char)(step | lev << 4); };
  auto child = [&]() {
    if(L + 1 < pl.nReg) wProgram(pl, L + 1);
    else add(kWDenseSolve, 0);
  };
  add(kWScaleB, L);
  add(kWPassE, L);
  add(kWRestrictE, L);
  child();
  add(kWProlong, L);
  add(kWPassE, L);
  add(kWPost1, L);
  add(kWPassR, L);
  add(kWResRScale, L);
  add(kWPassB, L);
  add(kWRestrictB, L);
  child();
  add(kWProlong, L);
  add(kWPassB, L);
  add(kWPost2, L);
}

// Plan, register maps and thread data of the W-cycle kernel (once per CSR
// pattern): R is the first level whose matrix (and, when the level below it
// has more than kWDense rows, that level's too) fits the registers and whose
// vectors, weights, aggregate starts and dense inverse fit the LDS; none: R = 0.
int wSetup(fe_hip_ctx *ctx, FeAmgPcg &A, const std::vector<std::vector<int>> &hIa,
           const std::vector<std::vector<int>> &hJa)
{
  const int Lc = (int)A.lv.size() - 1;
  for(int R = 1; R < Lc; ++R) {
    const int nReg = A.lv[R + 1].n <= kWDense ? 1 : 2;
    if(nReg == 2 && (R + 2 > Lc || A.lv[R + 2].n > kWDense)) continue;
    int K[2] = {0, 0};
    bool ok = true;
    for(int r = 0; r < nReg && ok; ++r) {
      const std::vector<int> &ia = hIa[R + r];
      int64_t P = 0;
      int maxPad = 0;
      for(size_t i = 0; i + 1 < ia.size(); ++i) {
        const int pd = (ia[i + 1] - ia[i] + kWGroup - 1) / kWGroup * kWGroup;
        P += pd;
        maxPad = std::max(maxPad, pd);
      }
      const int need = std::max(maxPad, (int)((P + kWTB - 1) / kWTB + kWGroup - 1) / kWGroup * kWGroup);
      K[r] = r == 0 ? (need <= kWKA[0] ? kWKA[0] : need <= kWKA[1] ? kWKA[1] : 0) : (need <= kWKB ? kWKB : 0);
      ok = K[r] > 0;
    }
    if(!ok) continue;
    WPlan pl;
    pl.R = R;
    pl.D = R + nReg;
    pl.nReg = nReg;
    pl.KA = K[0];
    pl.KB = nReg == 2 ? K[1] : 0;
    int bytes = 0;
    auto take = [&](size_t b) {
      const int o = bytes;
      bytes += (int)((b + 15) & ~(size_t)15);
      return o;
    };
    for(int r = 0; r < nReg; ++r) {
      pl.lv[r].n = A.lv[R + r].n;
      pl.lv[r].nc = A.lv[R + r].nc;
      pl.lv[r].oX = take(sizeof(float) * pl.lv[r].n);
    }
    if(bytes > 65536) continue;  // gathered entries are addressed with 16 bits
    pl.oImg = bytes;
    for(int r = 0; r < nReg; ++r) {
      pl.lv[r].oW = take(2 * (size_t)pl.lv[r].n);
      pl.lv[r].oAs = take(sizeof(int) * (size_t)(pl.lv[r].nc + 1));
    }
    pl.nD = A.lv[pl.D].n;
    pl.oDinv = take(sizeof(float) * (size_t)pl.nD * pl.nD);
    pl.imgBytes = bytes - pl.oImg;
    for(int r = 0; r < nReg; ++r) {
      pl.lv[r].oB = take(sizeof(float) * pl.lv[r].n);
      pl.lv[r].oE = take(sizeof(float) * pl.lv[r].n);
      pl.lv[r].oR = take(sizeof(float) * pl.lv[r].n);
    }
    pl.oBD = take(sizeof(float) * pl.nD);
    pl.oED = take(sizeof(float) * pl.nD);
    pl.oCarry = take(sizeof(float) * kWTB);
    pl.bytes = bytes;
    if(bytes > kWLds) continue;
    wProgram(pl, 0);
    for(int r = 0; r < nReg; ++r) {
      const int Kr = K[r], n = pl.lv[r].n, oX = pl.lv[r].oX;
      const std::vector<int> &ia = hIa[R + r], &ja = hJa[R + r];
      std::vector<int2> map((size_t)Kr * kWTB, make_int2(-1, oX << 16));
      std::vector<int> first(kWTB, n), cont(kWTB, 0);
      std::vector<unsigned long long> mask(kWTB, 0ull);
      auto word = [&](int64_t q) {  // padded position -> word slot
        const int t = (int)(q / Kr), s = (int)(q % Kr);
        return ((size_t)(s / kWGroup) * kWTB + t) * kWGroup + s % kWGroup;
      };
      int64_t p = 0;
      for(int i = 0; i < n; ++i) {
        const int len = ia[i + 1] - ia[i];
        const int pd = (len + kWGroup - 1) / kWGroup * kWGroup;
        for(int m = 0; m < len; ++m)
          map[word(p + m)] = make_int2(ia[i] + m, i | ((oX + (int)sizeof(float) * ja[ia[i] + m]) << 16));
        const int64_t e = p + pd - 1;
        mask[e / Kr] |= 1ull << ((e % Kr) / kWGroup);
        for(int64_t t = (p + Kr - 1) / Kr; t * Kr < p + pd; ++t) {
          first[t] = i;
          cont[t] = t * Kr > p;
        }
        p += pd;
      }
      FE_REQUIRE(ctx, p <= (int64_t)Kr * kWTB && cont[0] == 0, "nested multigrid PCG: W-cycle layout overflow");
      std::vector<int4> info(kWTB);
      for(int t = 0; t < kWTB; ++t)
        info[t] = make_int4(first[t], cont[t], (int)(uint32_t)mask[t], (int)(uint32_t)(mask[t] >> 32));
      if(int rc = upload(ctx, &A.wMap[r], map)) return rc;
      if(int rc = upload(ctx, &A.wInfo[r], info)) return rc;
      if(int rc = feAlloc(ctx, &A.wWords[r], (size_t)Kr * kWTB)) return rc;
    }
    if(int rc = feAlloc(ctx, &A.wImg, (size_t)pl.imgBytes)) return rc;
    A.wplan = pl;
    return FE_STATUS_OK;
  }
  return FE_STATUS_OK;
}

// Tiles of the fused coarse-level launch for a single-workgroup start level
// B: the level-B rows cut into contiguous shares so that no level F <= l < B
// gets more than kFuseRows rows in a share; level-l tiles are the rows nested
// in each share, for l = F .. B.  F is the lowest level (>= 1) with at most
// kFuseMaxG shares.  Returns G (0: B <= 1 or no such level).
int fuseTilesFor(const std::vector<std::vector<int>> &aggStart, const std::vector<FeAmgPcgLevel> &lv, int B,
                 std::vector<int> &tiles, int *Fout)
{
  if(B <= 1 || B >= (int)lv.size()) return 0;
  auto start = [&](int l, int r) {  // first level-l row nested in level-B row r
    for(int m = B - 1; m >= l; --m) r = aggStart[m][r];
    return r;
  };
  const int nB = lv[B].n;
  for(int F = 1; F < B; ++F) {
    std::vector<int> cuts(1, 0);
    bool bad = false;
    for(int r = 1; r <= nB && !bad; ++r) {
      bool over = false;
      for(int l = F; l < B; ++l) over = over || start(l, r) - start(l, cuts.back()) > kFuseRows;
      if(over) {
        if(r - 1 == cuts.back()) bad = true;  // one level-B row alone is too large
        cuts.push_back(r - 1);
        --r;
      }
    }
    if(bad) continue;
    cuts.push_back(nB);
    const int G = (int)cuts.size() - 1;
    if(G > kFuseMaxG) continue;
    tiles.assign((size_t)(B + 1) * (G + 1), 0);
    for(int l = F; l <= B; ++l)
      for(int w = 0; w <= G; ++w) tiles[(size_t)l * (G + 1) + w] = l == B ? cuts[w] : start(l, cuts[w]);
    *Fout = F;
    return G;
  }
  return 0;
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
  std::vector<std::vector<int>> entPos(Lc + 1), hIa(Lc + 1), hJa(Lc + 1);
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
    if(l > 0) {
      hIa[l] = ia;
      hJa[l] = ja;
    }
    if(l == 0) {
      // SELL layout: slices of kSlice rows, slot k of a slice's rows
      // contiguous; each original entry's SELL position for level 1's sums
      const int ns = (h.n + kSlice - 1) / kSlice;
      std::vector<int2> sell(h.n);
      int64_t tot = 0;
      for(int sl = 0; sl < ns; ++sl) {
        int w = 0;
        for(int r = sl * kSlice; r < std::min(h.n, (sl + 1) * kSlice); ++r) w = std::max(w, ia[r + 1] - ia[r]);
        for(int r = sl * kSlice; r < std::min(h.n, (sl + 1) * kSlice); ++r)
          sell[r] = make_int2((int)(tot + r - sl * kSlice), ia[r + 1] - ia[r]);
        tot += (int64_t)w * kSlice;
        FE_REQUIRE(ctx, tot < ((int64_t)1 << 31), "nested multigrid PCG: SELL layout overflows int32");
      }
      std::vector<int> scol(tot, 0), smap(tot, 0);
      for(int r = 0; r < h.n; ++r)
        for(int m = 0; m < sell[r].y; ++m) {
          const int q = sell[r].x + m * kSlice;
          scol[q] = ja[ia[r] + m];
          smap[q] = ent[ia[r] + m];
          entPos[0][ent[ia[r] + m]] = q;
        }
      d.nnz = tot;
      if(int rc = upload(ctx, &d.ja, scol)) return rc;
      if(int rc = upload(ctx, &A->perm, order[0])) return rc;
      if(int rc = upload(ctx, &A->map, smap)) return rc;
      if(int rc = upload(ctx, &A->sell, sell)) return rc;
      d.sell = A->sell;
    } else {
      if(int rc = upload(ctx, &d.ja, ja)) return rc;
      // Galerkin lists in permuted numbering: level 1 sums entries of the
      // system's original CSR, deeper levels entries of the permuted level
      // above (same list order as the original lists)
      std::vector<int> cptr(d.nnz + 1, 0), clist;
      clist.reserve(h.clist.size());
      for(int kp = 0; kp < d.nnz; ++kp) {
        const int k = ent[kp];
        for(int m = h.cptr[k]; m < h.cptr[k + 1]; ++m) clist.push_back(entPos[l - 1][h.clist[m]]);
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
      // tiles of whole aggregates, at most kFRPB rows on level 0 (its kernels
      // run kFTB threads) and kPRPB rows on the others
      const int lim = l == 0 ? kFRPB : kPRPB;
      std::vector<int> ts(1, 0), ta(1, 0);
      int rows = 0;
      for(int I = 0; I < h.nc; ++I) {
        const int sz = aggStart[l][I + 1] - aggStart[l][I];
        FE_REQUIRE(ctx, sz <= lim, "nested multigrid PCG: an aggregate is larger than a tile");
        if(rows + sz > lim) {
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
    if(int rc = feAlloc(ctx, &A->u32[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->r64v[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->p64v[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->u64v[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->q32[s], n)) return rc;
    if(int rc = feAlloc(ctx, &A->q64v[s], n)) return rc;
  }
  if(feAlloc(ctx, &A->x32, n) || feAlloc(ctx, &A->z32, n) || feAlloc(ctx, &A->rhs32, n) ||
     feAlloc(ctx, &A->y32, n) || feAlloc(ctx, &A->y64v, n) || feAlloc(ctx, &A->w32, n) ||
     feAlloc(ctx, &A->w64v, n) ||
     feAlloc(ctx, &A->x64v, n) || feAlloc(ctx, &A->z64v, n) ||
     feAlloc(ctx, &A->rhs64v, n) || feAlloc(ctx, &A->xsave, n))
    return FE_STATUS_OUT_OF_MEMORY;
  A->g0 = (n + kFRPB - 1) / kFRPB;
  const int gMax = std::max(A->g0, A->lv[0].nTiles);
  if(feAlloc(ctx, &A->pq, gMax) || feAlloc(ctx, &A->rr, gMax) || feAlloc(ctx, &A->rz, gMax) ||
     feAlloc(ctx, &A->scal, 8) || feAlloc(ctx, &A->done, 4) || feAlloc(ctx, &A->stat, 4))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipHostMalloc((void **)&A->mbHost, sizeof(PcgMailbox), hipHostMallocMapped | hipHostMallocCoherent));
  FE_HIP_TRY(ctx, hipHostGetDevicePointer((void **)&A->mbDev, A->mbHost, 0));
  A->plan[0] = pcgPlan<float, float>(*A);
  A->plan[1] = pcgPlan<double, float>(*A);
  A->plan[2] = pcgPlan<double, double>(*A);
  for(int s = 0; s < 3; ++s)
    if(A->plan[s].bytes <= kPBlockLds)
      if(int rc = feAlloc(ctx, &A->img[s], (size_t)std::max(16, A->plan[s].imgBytes))) return rc;
  for(int s = 0; s < 3; ++s) {
    std::vector<int> tiles;
    if(A->plan[s].bytes > kPBlockLds) continue;
    A->fuseG[s] = fuseTilesFor(aggStart, A->lv, A->plan[s].from, tiles, &A->fuseF[s]);
    if(A->fuseG[s] > 0)
      if(int rc = upload(ctx, &A->fuseTiles[s], tiles)) return rc;
  }
  if(int rc = feAlloc(ctx, &A->bar, 1)) return rc;
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
  return wSetup(ctx, *A, hIa, hJa);
}

// Numeric set-up for the current values (every solve): permuted level-0
// values and l1, Galerkin sums and l1 per level, dense coarsest inverse.
// fp64Values: FP64 level values; fp64Product: FP64 level-0 values for the
// A-product of FP64 inner solves (with FP32 level values otherwise).
int pcgNumeric(fe_hip_ctx *ctx, bool fp64Values, bool fp64Product, int slot)
{
  FeAmgPcg &A = *ctx->amgPcg;
  A.c = ctx->amgOmega;
  FeAmgPcgLevel &f = A.lv[0];
  // the FP64 SELL copy is always written: level 1 sums its entries
  const int g0 = (f.n + kPTB - 1) / kPTB;
  (void)fp64Product;
  if(fp64Values)
    hipLaunchKernelGGL((pcgSetup0<false, true>), dim3(g0), dim3(kPTB), 0, ctx->stream, f.n, A.sell, A.map, ctx->d_val,
                       A.c, f.val32, f.l1_32, f.valD, f.l1_64);
  else
    hipLaunchKernelGGL((pcgSetup0<true, true>), dim3(g0), dim3(kPTB), 0, ctx->stream, f.n, A.sell, A.map, ctx->d_val,
                       A.c, f.val32, f.l1_32, f.valD, f.l1_64);
  for(size_t l = 1; l < A.lv.size(); ++l) {
    FeAmgPcgLevel &c = A.lv[l];
    const FeAmgPcgLevel &p = A.lv[l - 1];
    const double *src = l == 1 ? p.valD : p.val64;  // level 1: the FP64 SELL values
    const int pn = l == 1 ? 0 : p.n;
    const int g = (int)((std::max<int64_t>(c.nnz, pn) + kPTB - 1) / kPTB);
    if(fp64Values)
      hipLaunchKernelGGL(pcgSetupL<double>, dim3(g), dim3(kPTB), 0, ctx->stream, (int)c.nnz, c.cptr, c.clist, src,
                         c.val64, c.valD, pn, p.ia, A.c, p.l1_64);
    else
      hipLaunchKernelGGL(pcgSetupL<float>, dim3(g), dim3(kPTB), 0, ctx->stream, (int)c.nnz, c.cptr, c.clist, src,
                         c.val64, c.val32, pn, p.ia, A.c, p.l1_32);
  }
  const FeAmgPcgLevel &C = A.lv.back();
  hipLaunchKernelGGL(pcgZero<int>, dim3(1), dim3(64), 0, ctx->stream, 1, A.flag, A.done, A.stat);
  if(fp64Values)
    hipLaunchKernelGGL(pcgDenseInverse<double>, dim3(1), dim3(1024), 0, ctx->stream, C.n, C.ia, C.ja, C.val64,
                       A.cinv64, A.cinv32, A.flag, A.c, C.l1_64);
  else
    hipLaunchKernelGGL(pcgDenseInverse<float>, dim3(1), dim3(1024), 0, ctx->stream, C.n, C.ia, C.ja, C.val64,
                       A.cinv64, A.cinv32, A.flag, A.c, C.l1_32);
  // the W-cycle kernel's register words and static image (FP32 level values)
  const WPlan &W = A.wplan;
  if(W.R > 0 && !fp64Values && ctx->amgWcycle) {
    for(int r = 0; r < W.nReg; ++r) {
      const FeAmgPcgLevel &h = A.lv[W.R + r];
      const int nq = (r == 0 ? W.KA : W.KB) * kWTB;
      const int cnt = std::max(nq, std::max(h.n, h.nc + 1));
      hipLaunchKernelGGL(pcgWPack, dim3((cnt + kPTB - 1) / kPTB), dim3(kPTB), 0, ctx->stream, nq,
                         (const int2 *)A.wMap[r], (const float *)h.val32, (const float *)h.l1_32, A.wWords[r], h.n,
                         (unsigned short *)(A.wImg + W.lv[r].oW - W.oImg), h.nc + 1, (const int *)h.aggStart,
                         (int *)(A.wImg + W.lv[r].oAs - W.oImg));
    }
    const FeAmgPcgLevel &Dl = A.lv[W.D];
    hipLaunchKernelGGL(pcgWDenseInverse, dim3(1), dim3(1024), 0, ctx->stream, Dl.n, (const int *)Dl.ia,
                       (const int *)Dl.ja, (const double *)Dl.val64, (float *)(A.wImg + W.oDinv - W.oImg), A.flag);
  }
  // the static LDS image of the small-level kernel for this slot
  const PBlockPlan &pl = A.plan[slot];
  if(slot == 0)
    hipLaunchKernelGGL((pcgPackBlock<float, float>), dim3(1), dim3(kPBlockTB), 0, ctx->stream,
                       (const PLev<float, float> *)A.desc[0], pl, (const float *)A.cinv32, A.img[0]);
  else if(slot == 1)
    hipLaunchKernelGGL((pcgPackBlock<double, float>), dim3(1), dim3(kPBlockTB), 0, ctx->stream,
                       (const PLev<double, float> *)A.desc[1], pl, (const double *)A.cinv64, A.img[1]);
  else
    hipLaunchKernelGGL((pcgPackBlock<double, double>), dim3(1), dim3(kPBlockTB), 0, ctx->stream,
                       (const PLev<double, double> *)A.desc[2], pl, (const double *)A.cinv64, A.img[2]);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

template <typename T>
const T *productValues(const FeAmgPcgLevel &h);
template <>
const float *productValues<float>(const FeAmgPcgLevel &h)
{
  return h.val32;
}
template <>
const double *productValues<double>(const FeAmgPcgLevel &h)
{
  return h.valD;
}

template <typename T>
struct PcgVecs {
  T *x, *r[2], *p[2], *u[2], *q[2], *z, *rhs, *y, *w;
  const T *cinv;
};
PcgVecs<float> vecs32(FeAmgPcg &A)
{
  return PcgVecs<float>{A.x32,  {A.r32[0], A.r32[1]}, {A.p32[0], A.p32[1]}, {A.u32[0], A.u32[1]}, {A.q32[0], A.q32[1]},
                        A.z32,  A.rhs32,              A.y32,                A.w32,                A.cinv32};
}
PcgVecs<double> vecs64(FeAmgPcg &A)
{
  return PcgVecs<double>{A.x64v,  {A.r64v[0], A.r64v[1]}, {A.p64v[0], A.p64v[1]}, {A.u64v[0], A.u64v[1]}, {A.q64v[0], A.q64v[1]},
                         A.z64v,  A.rhs64v,               A.y64v,                 A.w64v,                 A.cinv64};
}

// The W-cycle kernel on levels R .. D (FP32 level values): b_R in, e_R out,
// with R = 1 also the level-0 y.
template <typename T>
void launchW(fe_hip_ctx *ctx, FeAmgPcg &A, const PcgVecs<T> &v, const T *u0)
{
  const WPlan &W = A.wplan;
  const PLev<T, float> LR = pview<T, float>(A.lv[W.R]);
  T *y0 = W.R == 1 ? v.y : nullptr;
  const float oc = (float)ctx->amgOverCorrect;
#define FE_W_LAUNCH(KA_, KB_)                                                                                         \
  hipLaunchKernelGGL((pcgCoarseW<T, KA_, KB_>), dim3(1), dim3(kWTB), 0, ctx->stream, W, (const uint4 *)A.wWords[0],  \
                     (const int4 *)A.wInfo[0], (const uint4 *)A.wWords[1], (const int4 *)A.wInfo[1],                  \
                     (const unsigned char *)A.wImg, oc, (const int *)A.done, (const T *)LR.b, LR.e, y0, u0,          \
                     (const int *)A.lv[0].aggStart)
  if(W.KA == kWKA[0] && W.KB == 0) FE_W_LAUNCH(kWKA[0], 0);
  else if(W.KA == kWKA[0]) FE_W_LAUNCH(kWKA[0], kWKB);
  else if(W.KB == 0) FE_W_LAUNCH(kWKA[1], 0);
  else FE_W_LAUNCH(kWKA[1], kWKB);
#undef FE_W_LAUNCH
}

// Launches the V-cycle after down0: down_l (levels 1 .. B-1), the block
// kernel (or the W-cycle kernel from level B = R), up_l (levels B-1 .. 1;
// level 1 also writes the level-0 y), then up0 on rNew into z.
template <typename T, typename TV>
void vcycleRest(fe_hip_ctx *ctx, FeAmgPcg &A, int slot, const PcgVecs<T> &v, const T *rNew, const T *u0)
{
  const PBlockPlan &pl = A.plan[slot];
  const bool useW = std::is_same<TV, float>::value && A.wplan.R > 0 && ctx->amgWcycle;
  const int B = useW ? A.wplan.R : pl.from;
  const T oc = (T)ctx->amgOverCorrect;
  if(!useW && ctx->amgFuse && A.fuseG[slot] > 0) {
    const int F = A.fuseF[slot];
    const PLev<T, TV> L0 = pview<T, TV>(A.lv[0]);
    for(int l = 1; l < F; ++l) {
      const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
      const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
      hipLaunchKernelGGL((pcgDownL<T, TV>), dim3(L.nTiles), dim3(kPTB), 0, ctx->stream, L, N.b, A.done);
    }
    hipLaunchKernelGGL((pcgCoarseV<T, TV>), dim3(A.fuseG[slot]), dim3(kPBlockTB), 0, ctx->stream,
                       (const PLev<T, TV> *)A.desc[slot], pl, (const unsigned char *)A.img[slot],
                       (const int *)A.fuseTiles[slot], A.fuseG[slot], F, oc, (const int *)A.done, A.bar,
                       &A.mbDev->barErr, v.y, u0, (const int *)L0.aggStart);
    for(int l = F - 1; l >= 1; --l) {
      const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
      const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
      hipLaunchKernelGGL((pcgUpL<T, TV>), dim3((L.n + kPRPB - 1) / kPRPB), dim3(kPTB), 0, ctx->stream, L, N.e, oc,
                         A.done, l == 1 ? v.y : (T *)nullptr, u0, L0.aggStart);
    }
    hipLaunchKernelGGL((pcgUp0<T, TV>), dim3(A.g0), dim3(kFTB), 0, ctx->stream, L0, (const T *)v.y, rNew, v.z,
                       A.done, A.rz);
    return;
  }
  const PLev<T, TV> *d = (const PLev<T, TV> *)A.desc[slot];
  const PLev<T, TV> L0 = pview<T, TV>(A.lv[0]);
  for(int l = 1; l < B; ++l) {
    const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
    const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
    hipLaunchKernelGGL((pcgDownL<T, TV>), dim3(L.nTiles), dim3(kPTB), 0, ctx->stream, L, N.b, A.done);
  }
  if(useW)
    launchW<T>(ctx, A, v, u0);
  else
    hipLaunchKernelGGL((pcgBlock<T, TV>), dim3(1), dim3(kPBlockTB), 0, ctx->stream, d, pl,
                       (const unsigned char *)A.img[slot], oc, A.done, v.y, u0);
  for(int l = B - 1; l >= 1; --l) {
    const PLev<T, TV> L = pview<T, TV>(A.lv[l]);
    const PLev<T, TV> N = pview<T, TV>(A.lv[l + 1]);
    hipLaunchKernelGGL((pcgUpL<T, TV>), dim3((L.n + kPRPB - 1) / kPRPB), dim3(kPTB), 0, ctx->stream, L, N.e, oc,
                       A.done, l == 1 ? v.y : (T *)nullptr, u0, L0.aggStart);
  }
  hipLaunchKernelGGL((pcgUp0<T, TV>), dim3(A.g0), dim3(kFTB), 0, ctx->stream, L0, (const T *)v.y, rNew, v.z, A.done,
                     A.rz);
}

// Iterations queued ahead of the one whose product the device last started:
// enough to keep it busy while the host launches, and the most that run as
// early exits after a stop.
constexpr int kPcgAhead = 2;

// State of one solve across its segments: the next iteration (its parity
// picks the double buffers).
struct PcgRun {
  int k = 0;
};

// One segment of PCG iterations on A x = rhs (permuted, vector type T), from
// x = 0, until (r, r) <= tol2, maxit iterations or a breakdown.
//   resume = false: a fresh solve, iteration 0 with r = rhs.
//   resume = true:  a residual replacement.  r = rhs (the caller's FP64
//     residual of the iterate it has accumulated), z = M r, and iteration
//     run.k, at which the previous segment stopped, is redone from the kept
//     p_k-1, q_k-1 = A p_k-1 and rho_k-1, so the Krylov history carries over
//     instead of restarting.
// The host launches iteration k once the device has started the product of
// iteration k - kPcgAhead, and stops launching when the mailbox shows a stop.
// Returns the segment's iterations in *iters, the stop code (1 converged,
// 2 iteration cap, 3 breakdown) and the recursive (r, r) at a stop in *rrLast;
// run.k is left at the iteration that stopped.
template <typename T, typename TV>
int pcgSegment(fe_hip_ctx *ctx, FeAmgPcg &A, int slot, const PcgVecs<T> &v, double tol2, int maxit, bool resume,
               PcgRun &run, int *iters, int *code, double *rrLast)
{
  const int n = A.lv[0].n;
  const PLev<T, TV> L0 = pview<T, TV>(A.lv[0]);
  const PLev<T, TV> L1 = pview<T, TV>(A.lv[1]);
  if(!resume) run.k = 0;
  const int k0 = run.k;
  volatile PcgMailbox *mb = A.mbHost;  // no kernel of this context is running: the stream is idle here
  mb->done = 0;
  mb->stopK = -1;
  mb->rr = 0.;
  mb->progress = k0 - 1;
  mb->barErr = 0;
  {
    const int s = k0 & 1;
    hipLaunchKernelGGL(pcgZero<T>, dim3(256), dim3(256), 0, ctx->stream, n, v.x, A.done, A.stat);
    // z = M r (down0 in first mode copies rhs into r[s] and forms u[s] = l1 r)
    hipLaunchKernelGGL((pcgDown0<T, TV>), dim3(L0.nTiles), dim3(kFTB), 0, ctx->stream, L0, v.x, (const T *)v.p[s],
                       (const T *)v.q[s], (const T *)v.w, (const T *)v.rhs, v.r[s], (const T *)v.u[s ^ 1], v.u[s],
                       L1.b, (const double *)A.pq, A.g0, (const double *)A.scal, 0, 1, A.done, A.stat, A.mbDev, k0, A.rr, A.bar);
    vcycleRest<T, TV>(ctx, A, slot, v, v.r[s], v.u[s]);
  }
  const T *prodVal = productValues<T>(A.lv[0]);
  const int kMax = k0 + std::max(0, maxit);
  // profiling: events around every down0 and around every whole iteration
  FeKernelTimer &tk = ctx->spmvTimer, &ti = ctx->iterTimer;
  tk.used = ti.used = 0;
  auto mark = [&](FeKernelTimer &t) -> int {
    if(t.used + 1 > t.ev.size())
      for(int q = 0; q < 64; ++q) {
        hipEvent_t e;
        FE_HIP_TRY(ctx, hipEventCreate(&e));
        t.ev.push_back(e);
      }
    FE_HIP_TRY(ctx, hipEventRecord(t.ev[t.used++], ctx->stream));
    return FE_STATUS_OK;
  };
  int k = k0;
  for(; k < kMax; ++k) {
    // wait until the device is at most kPcgAhead iterations behind; the
    // stream draining ends the wait too (nothing left to wait for)
    for(int spin = 1; mb->progress < k - kPcgAhead && !mb->done; ++spin)
      if((spin & 255) == 0 && hipStreamQuery(ctx->stream) != hipErrorNotReady) break;
    if(mb->done) break;
    const int s = k & 1;
    if(ctx->profiling)
      if(int rc = mark(ti)) return rc;
    // p_k = z_k + beta p_k-1, q_k = A z_k + beta q_k-1 (level-0 values in the
    // vector precision), w = l1 q_k
    hipLaunchKernelGGL((pcgProduct<T, T, TV>), dim3(A.g0), dim3(kFTB), 0, ctx->stream, n, (const int2 *)L0.sell,
                       L0.ja, prodVal, L0.l1, (const T *)v.z, (const T *)v.p[s ^ 1], v.p[s], (const T *)v.q[s ^ 1],
                       v.q[s], v.w, (const double *)A.rz, A.g0, (const double *)A.rr, L0.nTiles, A.scal, s,
                       k == 0 ? 1 : 0, tol2, A.done, A.stat, A.mbDev, k, A.pq);
    if(ctx->profiling)
      if(int rc = mark(tk)) return rc;
    // r_k+1 = r_k - alpha q_k into r[(k + 1) & 1], x += alpha p_k, u = l1 r_k+1
    // into u[(k + 1) & 1] (u[k & 1] = l1 r_k), V-cycle down
    hipLaunchKernelGGL((pcgDown0<T, TV>), dim3(L0.nTiles), dim3(kFTB), 0, ctx->stream, L0, v.x, (const T *)v.p[s],
                       (const T *)v.q[s], (const T *)v.w, (const T *)v.r[s], v.r[s ^ 1], (const T *)v.u[s], v.u[s ^ 1],
                       L1.b, (const double *)A.pq, A.g0, (const double *)A.scal, s, 0, A.done, A.stat, A.mbDev, k,
                       A.rr, A.bar);
    if(ctx->profiling)
      if(int rc = mark(tk)) return rc;
    vcycleRest<T, TV>(ctx, A, slot, v, v.r[s ^ 1], v.u[s ^ 1]);
    if(ctx->profiling)
      if(int rc = mark(ti)) return rc;
  }
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, !mb->barErr, "nested multigrid PCG: a grid barrier of the coarse-level launch timed out");
  const int dn = mb->done;
  if(dn) {
    run.k = mb->stopK >= k0 ? mb->stopK : k;
    *code = dn;
    *rrLast = mb->rr;
  } else {
    run.k = k;
    *code = 2;
    *rrLast = 0.;
  }
  *iters = run.k - k0;
  if(ctx->profiling) {
    // the iterations that ran in full (the launches after a stop exit at once)
    const size_t full = (size_t)std::max(0, run.k - k0);
    for(FeKernelTimer *t : {&tk, &ti}) {
      for(size_t q = 0; q < std::min(full, t->used / 2); ++q) {
        float ms = 0.f;
        FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t->ev[2 * q], t->ev[2 * q + 1]));
        t->total_ms += ms;
        t->launches++;
      }
      t->used = 0;
    }
  }
  return FE_STATUS_OK;
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
  FE_REQUIRE(ctx, (slot != 2 && A.wplan.R > 0 && ctx->amgWcycle) || A.plan[slot].bytes <= kPBlockLds,
             "nested multigrid PCG: the small levels do not fit the LDS");
  if(int rc = pcgNumeric(ctx, val64, !fp32, slot)) return rc;
  if(!ctx->h_state) FE_HIP_TRY(ctx, hipHostMalloc(&ctx->h_state, 4096));
  const int n = ctx->n;
  double rn = bnorm;
  bool inner32 = fp32, resume = false;
  PcgRun run;
  int iters = 0, outer = 0, lastIt = 0;
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
    // segment target: FP32 segments reduce the residual by inner_rtol, FP64
    // ones aim at the outer target directly; neither goes below half of it
    const double t = inner32 ? std::max(o->inner_rtol * rn, 0.5 * target) : 0.5 * target;
    // a resumed segment that needs more than twice its predecessor's
    // iterations has lost the conjugacy it carried over: it is stopped and
    // the next segment restarts
    int segMax = o->max_iter - iters;
    if(resume) segMax = std::min(segMax, 2 * lastIt + 10);
    int it = 0, code = 0;
    double rrRec = 0.;
    const int gp = std::min(1024, (n + 255) / 256);
    FE_HIP_TRY(ctx, hipMemcpyAsync(A.xsave, x64, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    if(inner32) {
      PcgVecs<float> v = vecs32(A);
      hipLaunchKernelGGL(pcgPermuteRhs<float>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r64, v.rhs);
      if(int rc = pcgSegment<float, float>(ctx, A, 0, v, t * t, segMax, resume, run, &it, &code, &rrRec))
        return rc;
      hipLaunchKernelGGL(pcgAddBack<float>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, (const float *)v.x, x64);
    } else {
      PcgVecs<double> v = vecs64(A);
      hipLaunchKernelGGL(pcgPermuteRhs<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r64, v.rhs);
      if(val64) {
        if(int rc = pcgSegment<double, double>(ctx, A, 2, v, t * t, segMax, resume, run, &it, &code,
                                               &rrRec))
          return rc;
      } else if(int rc = pcgSegment<double, float>(ctx, A, 1, v, t * t, segMax, resume, run, &it, &code,
                                                   &rrRec)) {
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
    // a segment that reached its target continues with the replaced
    // residual when the FP64 residual is close to the recursive one (the
    // directions carried over are then still conjugate enough); after a
    // breakdown, the cap or a larger gap the next segment starts afresh
    resume = code == 1 && rnew <= 2. * std::sqrt(std::max(rrRec, 0.));
    lastIt = it;
    if(inner32 && !(rnew < 0.9 * rn)) {
      // the FP32 segment did not reduce the FP64 residual enough: continue
      // with FP64 inner solves (FP64 product values set up now), from the
      // previous iterate when it grew
      inner32 = false;
      resume = false;
      FE_REQUIRE(ctx, A.plan[1].bytes <= kPBlockLds, "nested multigrid PCG: the small levels do not fit the LDS");
      if(int rc = pcgNumeric(ctx, false, true, 1)) return rc;
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

int feAmgPcgWLevel(const fe_hip_ctx *ctx)
{
  return ctx->amgPcg && ctx->amgWcycle ? ctx->amgPcg->wplan.R : 0;
}

int feAmgPcgFuseGrid(const fe_hip_ctx *ctx)
{
  if(!ctx->amgPcg || !ctx->amgFuse) return 0;
  const int slot = ctx->pcValues32 ? 1 : 2;
  return ctx->amgPcg->fuseG[slot];
}

int feAmgPcgWDense(const fe_hip_ctx *ctx)
{
  return ctx->amgPcg && ctx->amgWcycle ? ctx->amgPcg->wplan.D : 0;
}

bool feAmgNestedUsable(const fe_hip_ctx *ctx)
{
  return ctx->amgKernel == 2 && ctx->amgHostSweeps == 0 && ctx->amgHost.size() >= 2 &&
         ctx->amgHost.back().n <= kPDense && ctx->amgHostMaxAgg <= kPRPB;
}

int feAmgPcgPrepare(fe_hip_ctx *ctx)
{
  if(!ctx->amgPcg)
    if(int rc = pcgSetupPattern(ctx)) return rc;
  FeAmgPcg &A = *ctx->amgPcg;
  const bool val64 = !ctx->pcValues32;
  const int slot = val64 ? 2 : 1;
  FE_REQUIRE(ctx, (slot != 2 && A.wplan.R > 0 && ctx->amgWcycle) || A.plan[slot].bytes <= kPBlockLds,
             "nested multigrid PCG: the small levels do not fit the LDS");
  if(int rc = pcgNumeric(ctx, val64, false, slot)) return rc;
  int flag = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&flag, A.flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, flag == 0, "AMG: the coarsest operator is not positive definite");
  return FE_STATUS_OK;
}

int feAmgPcgVcycle(fe_hip_ctx *ctx, const double *r, double *z)
{
  FE_REQUIRE(ctx, ctx->amgPcg, "nested multigrid: set up before the V-cycle");
  FeAmgPcg &A = *ctx->amgPcg;
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the mailbox is written by the device
  ((volatile PcgMailbox *)A.mbHost)->barErr = 0;
  const bool val64 = !ctx->pcValues32;
  const int n = A.lv[0].n;
  const int gp = std::min(1024, (n + 255) / 256);
  PcgVecs<double> v = vecs64(A);
  hipLaunchKernelGGL(pcgPermuteRhs<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, r, v.rhs);
  hipLaunchKernelGGL(pcgZero<double>, dim3(gp), dim3(256), 0, ctx->stream, n, v.x, A.done, A.stat);
  if(val64) {
    const PLev<double, double> L0 = pview<double, double>(A.lv[0]), L1 = pview<double, double>(A.lv[1]);
    hipLaunchKernelGGL((pcgDown0<double, double>), dim3(L0.nTiles), dim3(kFTB), 0, ctx->stream, L0, v.x,
                       (const double *)v.p[0], (const double *)v.q[0], (const double *)v.w, (const double *)v.rhs, v.r[0],
                       (const double *)v.u[1], v.u[0], L1.b, (const double *)A.pq, A.g0, (const double *)A.scal, 0, 1,
                       A.done, A.stat, A.mbDev, 0, A.rr, A.bar);
    vcycleRest<double, double>(ctx, A, 2, v, v.r[0], v.u[0]);
  } else {
    const PLev<double, float> L0 = pview<double, float>(A.lv[0]), L1 = pview<double, float>(A.lv[1]);
    hipLaunchKernelGGL((pcgDown0<double, float>), dim3(L0.nTiles), dim3(kFTB), 0, ctx->stream, L0, v.x,
                       (const double *)v.p[0], (const double *)v.q[0], (const double *)v.w, (const double *)v.rhs, v.r[0],
                       (const double *)v.u[1], v.u[0], L1.b, (const double *)A.pq, A.g0, (const double *)A.scal, 0, 1,
                       A.done, A.stat, A.mbDev, 0, A.rr, A.bar);
    vcycleRest<double, float>(ctx, A, 1, v, v.r[0], v.u[0]);
  }
  hipLaunchKernelGGL(pcgScatterBack<double>, dim3(gp), dim3(256), 0, ctx->stream, n, A.perm, (const double *)v.z, z);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, !A.mbHost->barErr, "nested multigrid: a grid barrier of the coarse-level launch timed out");
  return FE_STATUS_OK;
}

int feAmgPcgApply(fe_hip_ctx *ctx, const double *r, double *z)
{
  if(int rc = feAmgPcgPrepare(ctx)) return rc;
  return feAmgPcgVcycle(ctx, r, z);
}

void feAmgPcgRelease(fe_hip_ctx *ctx)
{
  if(ctx->amgPcg && ctx->amgPcg->mbHost) (void)hipHostFree(ctx->amgPcg->mbHost);
  delete ctx->amgPcg;  // device buffers are owned by the context
  ctx->amgPcg = nullptr;
}
