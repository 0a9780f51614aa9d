// Aggregation algebraic multigrid preconditioner for CG (FE_PC_AMG), the
// counterpart of the algebraic multigrid PETSc selects with `-pc_type gamg`
// through the runtime options the reference's PETSc back end reads
// (feLinearSystemPETSc.cpp:900-1031; the solver is configured by options, not
// by code, so only the role is alike).
//
// Hierarchy (built once per CSR pattern, on the host, like the pattern
// itself): aggregates are a maximal independent set of the matrix graph taken
// greedily in a hashed priority order, each root with its neighbours, every
// other node joining the largest-numbered neighbouring root.  Prolongation is
// the piecewise-constant tentative one (unsmoothed aggregation), so the
// Galerkin operator A_c = P^T A P is a plain sum of fine entries: each coarse
// entry keeps the list of fine entries that add into it, and the numeric
// set-up per matrix is one gather-sum per level.
//
// V-cycle (symmetric, so CG can use it): l1-Jacobi smoothing S = c D_l1^-1,
// D_l1 = diag(sum_j |a_ij|) (A <= D_l1, so S converges for c < 2), one step
// before and after the coarse correction, which is over-corrected by a
// factor oc < 2:
//   u = S b;  b_c = P^T (b - A u);  e = V(b_c);  y = u + oc P e;
//   x = y + S (b - A y).
// The first step is fused into the restriction (one kernel per level reads
// the aggregate's rows and writes b_c), the prolongation into the second
// smoothing step (one kernel per level gathers u_j + oc e[agg_j]).  Levels of
// at most `blockRows` rows and the coarsest solve run in one
// single-workgroup kernel, so a V-cycle is 2 L + 1 launches for L large
// levels.  The coarsest level (at most kAmgDense rows) is inverted densely per
// matrix; when coarsening stalls above that size (a block-diagonal batch of
// small problems ends at a node or two per problem) it is solved by
// kAmgCoarseSweeps l1-Jacobi sweeps instead, a fixed SPD polynomial.

#include <algorithm>
#include <cmath>
#include <numeric>

#include "fe_hip_internal.h"
#include "fe_tile.h"

using namespace fe_tile;

namespace {

constexpr int kAmgDense = 64;       // the coarsest level is inverted densely up to this size
constexpr int kAmgMaxLevels = 20;
constexpr double kAmgStall = 0.9;   // coarsening stops when a level keeps more than this fraction of rows
constexpr int kAmgCoarseSweeps = 4; // l1-Jacobi sweeps of a stalled coarsest level larger than kAmgDense
constexpr int kAmgTileMax = 2048;   // rows of one aggregate tile (LDS)
constexpr int kAmgTileRows = 256;   // target rows per aggregate tile
constexpr int kAmgTB = 256;         // block of the level kernels
constexpr int kAmgLPR = 4;          // lanes per row in the level kernels
constexpr int kAmgBlockTB = 1024;   // the small-level kernel
constexpr int kAmgBlockLds = 152 * 1024;  // LDS of the small-level kernel (bytes)
constexpr int kAmgBlockLev = 8;           // small levels (coarsest excluded) it can hold

// LDS layout of the small-level kernel: levels from .. coarsest - 1 (slot
// s = l - from) with their CSR (16-bit columns), values, l1, aggregates and
// the vectors b (then x), u, r (then y); the coarsest right-hand side and
// solution (its dense inverse is read from global memory).
struct AmgBlockPlan {
  int from = 0, coarsest = 0;
  int n[kAmgBlockLev + 1], nc[kAmgBlockLev], nnz[kAmgBlockLev];
  int oIa[kAmgBlockLev], oJa[kAmgBlockLev], oVal[kAmgBlockLev], oL1[kAmgBlockLev], oAgg[kAmgBlockLev];
  int oMptr[kAmgBlockLev], oMlist[kAmgBlockLev], oU[kAmgBlockLev], oR[kAmgBlockLev];
  int oB[kAmgBlockLev + 1];
  int oCx = 0, bytes = 0;
  int mglob[kAmgBlockLev];  // 1: the level's matrix and aggregates are read from global memory, not staged
  int coarseSweeps = 0;  // 0: dense inverse of the coarsest level; > 0: l1-Jacobi sweeps (global memory)
  float coarseScale = 1.f;  // 1 / c: turns the stored l1 = c / sum |a_ij| into 1 / sum |a_ij|
};

// Tile of this block: with xcd set and a grid that is a multiple of 8, the
// 8 XCDs (block b runs on XCD b % 8) take 8 contiguous ranges of tiles, so
// the gathers of neighbouring rows stay in one XCD's L2.
__device__ __forceinline__ int amgTile(int xcd)
{
  const int G = gridDim.x;
  if(!xcd || (G & 7)) return blockIdx.x;
  return (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
}

// A tile whose problems [r.x, r.y] have all stopped (stop[p] != 0, the
// batched solver's per-problem flags) is skipped; a tile over more than
// kAmgSkipSpan problems (a coarse level) always runs.
constexpr int kAmgSkipSpan = 16;
__device__ __forceinline__ bool amgSkip(const int2 *__restrict__ tp, const int *__restrict__ stop, int tile)
{
  if(!stop) return false;
  const int2 r = tp[tile];
  if(r.y - r.x >= kAmgSkipSpan) return false;
  for(int p = r.x; p <= r.y; ++p)
    if(!stop[p]) return false;
  return true;
}

template <typename T, typename TV>
struct AmgLev {
  int n, nc;
  const int *ia, *ja;
  const TV *val, *l1;  // l1[i] = c / sum_j |a_ij|
  const int *agg, *mptr, *mlist;
  const int2 *mrange;  // row range [ia[i], ia[i + 1]) of member i = mlist[q]
  T *b, *u, *x, *r;
};

}  // namespace

struct FeAmgLevel {
  int n = 0, nc = 0;
  int64_t nnz = 0;
  int *ia = nullptr, *ja = nullptr;  // level 0: the context's CSR
  double *val64 = nullptr;           // level 0: the context's values
  float *val32 = nullptr;
  double *l1_64 = nullptr;
  float *l1_32 = nullptr;
  int *agg = nullptr, *mptr = nullptr, *mlist = nullptr;  // to level + 1
  int2 *mrange = nullptr;
  // aggregate-aligned tiles of kAmgTB member positions (no aggregate crosses
  // a tile; padding positions have row -1): row, row range, aggregate
  int padTiles = 0;
  int *prow = nullptr, *pagg = nullptr;
  int2 *prg = nullptr;
  int *cptr = nullptr, *clist = nullptr;                  // Galerkin sums from level - 1
  int nTiles = 0;
  int *tileRow = nullptr, *tileAgg = nullptr;
  // problems touched by each padded tile, aggregate tile and row block of
  // the level kernels (block-diagonal systems of independent problems)
  int2 *probPad = nullptr, *probTile = nullptr, *probUp = nullptr;
  std::vector<int2> hPadAgg;   // first and last aggregate of each padded tile
  std::vector<int> hTileAgg;   // aggregate range boundaries of the aggregate tiles
  double *b64 = nullptr, *u64 = nullptr, *x64 = nullptr, *r64 = nullptr;
  float *b32 = nullptr, *u32 = nullptr, *x32 = nullptr, *r32 = nullptr;
  int maxAgg = 0;
};

struct FeAmg {
  std::vector<FeAmgLevel> lv;
  AmgBlockPlan plan[3];  // per (vector, value) type combination of feAmgApply
  int coarseSweeps = 0;  // stalled coarsening: the coarsest level is swept, not inverted
  double *cinv64 = nullptr;
  float *cinv32 = nullptr;
  void *desc[3] = {nullptr, nullptr, nullptr};  // AmgLev<double,double>, <double,float>, <float,float>
  int *flag = nullptr;
  double c = 0., oc = 0.;
  const int *act = nullptr;    // per-problem stop flags (tile skipping on)
  std::vector<int> probRows;   // the problem rows the tile ranges were built for
};

namespace {

// ---------------------------------------------------------------------------
// Host: aggregation and the Galerkin pattern

uint64_t amgMix(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Greedy maximal independent set in descending hashed priority (the fixed
// point of Luby's rounds with these priorities), roots numbered in row order,
// every other row joining the largest-numbered neighbouring root.
int amgAggregate(int n, const std::vector<int> &ia, const std::vector<int> &ja, std::vector<int> &agg)
{
  std::vector<uint64_t> pri(n);
  std::vector<int> order(n);
  for(int i = 0; i < n; ++i) {
    pri[i] = amgMix((uint64_t)i);
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return pri[a] != pri[b] ? pri[a] > pri[b] : a < b; });
  std::vector<signed char> state(n, 0);
  for(int i : order) {
    if(state[i] != 0) continue;
    state[i] = 1;
    for(int k = ia[i]; k < ia[i + 1]; ++k)
      if(state[ja[k]] == 0) state[ja[k]] = -1;
  }
  agg.assign(n, -1);
  int nc = 0;
  for(int i = 0; i < n; ++i)
    if(state[i] == 1) agg[i] = nc++;
  for(int i = 0; i < n; ++i) {
    if(state[i] == 1) continue;
    int best = -1;
    for(int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if(state[j] == 1) best = std::max(best, agg[j]);
    }
    agg[i] = best;  // a non-root was excluded by a neighbouring root
  }
  return nc;
}

struct HostLevel {
  int n = 0, nc = 0;
  std::vector<int> ia, ja, agg, mptr, mlist, cptr, clist, tileRow, tileAgg;
  int maxAgg = 0;
};

// Coarse pattern of P^T A P and, for each coarse entry, the fine entries that
// sum into it (fine -> coarse position, then grouped by position).
void amgGalerkinPattern(const HostLevel &f, std::vector<int> &iac, std::vector<int> &jac, std::vector<int> &cptr,
                        std::vector<int> &clist)
{
  const int nc = f.nc;
  iac.assign(nc + 1, 0);
  jac.clear();
  std::vector<int> mark(nc, -1), cols;
  for(int I = 0; I < nc; ++I) {
    cols.clear();
    for(int q = f.mptr[I]; q < f.mptr[I + 1]; ++q) {
      const int i = f.mlist[q];
      for(int k = f.ia[i]; k < f.ia[i + 1]; ++k) {
        const int J = f.agg[f.ja[k]];
        if(mark[J] != I) {
          mark[J] = I;
          cols.push_back(J);
        }
      }
    }
    std::sort(cols.begin(), cols.end());
    jac.insert(jac.end(), cols.begin(), cols.end());
    iac[I + 1] = (int)jac.size();
  }
  const int64_t nnzF = f.ia[f.n];
  std::vector<int> pos(nnzF);
  for(int i = 0; i < f.n; ++i) {
    const int I = f.agg[i];
    const int *b = jac.data() + iac[I], *e = jac.data() + iac[I + 1];
    for(int k = f.ia[i]; k < f.ia[i + 1]; ++k) pos[k] = (int)(std::lower_bound(b, e, f.agg[f.ja[k]]) - jac.data());
  }
  const int nnzC = iac[nc];
  cptr.assign(nnzC + 1, 0);
  for(int64_t k = 0; k < nnzF; ++k) cptr[pos[k] + 1]++;
  for(int p = 0; p < nnzC; ++p) cptr[p + 1] += cptr[p];
  clist.resize(nnzF);
  std::vector<int> fill(cptr.begin(), cptr.end() - 1);
  for(int64_t k = 0; k < nnzF; ++k) clist[fill[pos[k]]++] = (int)k;
}

// ---------------------------------------------------------------------------
// Device: numeric set-up

// l1[i] = c / sum_j |a_ij| of the fine level (8 lanes per row), and the FP32
// rounding of the fine values.
__global__ __launch_bounds__(256) void amgFineL1(int n, const int *__restrict__ ia, const double *__restrict__ val,
                                                 double c, double *__restrict__ l1_64, float *__restrict__ l1_32,
                                                 float *__restrict__ val32)
{
  const int sub = threadIdx.x & 7;
  for(int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; i < n; i += (gridDim.x * blockDim.x) >> 3) {
    double a = 0.;
    for(int k = ia[i] + sub; k < ia[i + 1]; k += 8) {
      const double v = val[k];
      val32[k] = (float)v;
      a += fabs(v);
    }
#pragma unroll
    for(int o = 4; o > 0; o >>= 1) a += __shfl_xor(a, o, 8);
    if(sub == 0) {
      const double s = a > 0. ? c / a : 0.;
      l1_64[i] = s;
      l1_32[i] = (float)s;
    }
  }
}

// Coarse values as sums of fine values (fixed order, eight contributions in
// flight per thread), one thread per coarse entry.
__global__ __launch_bounds__(256) void amgGalerkin(int nnzC, const int *__restrict__ cptr, const int *__restrict__ clist,
                                                   const double *__restrict__ valF, double *__restrict__ val64,
                                                   float *__restrict__ val32)
{
  for(int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnzC; p += gridDim.x * blockDim.x) {
    double s = 0.;
    const int q1 = cptr[p + 1];
    for(int q = cptr[p]; q < q1; q += 8) {
      int kk[8];
#pragma unroll
      for(int u = 0; u < 8; ++u) kk[u] = q + u < q1 ? clist[q + u] : -1;
#pragma unroll
      for(int u = 0; u < 8; ++u)
        if(kk[u] >= 0) s += valF[kk[u]];
    }
    val64[p] = s;
    val32[p] = (float)s;
  }
}

// l1[i] = c / sum_j |a_ij| of a coarse level (8 lanes per row).
__global__ __launch_bounds__(256) void amgRowL1(int n, const int *__restrict__ ia, const double *__restrict__ val,
                                                double c, double *__restrict__ l1_64, float *__restrict__ l1_32)
{
  const int sub = threadIdx.x & 7;
  for(int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; i < n; i += (gridDim.x * blockDim.x) >> 3) {
    double a = 0.;
    for(int k = ia[i] + sub; k < ia[i + 1]; k += 8) a += fabs(val[k]);
#pragma unroll
    for(int o = 4; o > 0; o >>= 1) a += __shfl_xor(a, o, 8);
    if(sub == 0) {
      const double s = a > 0. ? c / a : 0.;
      l1_64[i] = s;
      l1_32[i] = (float)s;
    }
  }
}

// Dense inverse of the coarsest operator (n <= kAmgDense = 64, SPD) by
// in-place Gauss-Jordan elimination without pivoting in one workgroup of
// 32 x 32 threads: thread (ty, tx) keeps entries (ty + 32 a, tx + 32 b),
// a, b < 2, in registers; the pivot row and column of every step go through
// LDS.  flag is set when a pivot is not positive.
__global__ __launch_bounds__(1024) void amgDenseInverse(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                        const double *__restrict__ val, double *__restrict__ inv64,
                                                        float *__restrict__ inv32, int *__restrict__ flag)
{
  static_assert(kAmgDense == 64, "2 x 2 entries of 32 x 32 threads");
  __shared__ double D[kAmgDense * kAmgDense];
  __shared__ double prow[kAmgDense], pcol[kAmgDense];
  const int t = threadIdx.x, tx = t & 31, ty = t >> 5;
  for(int e = t; e < n * n; e += 1024) D[e] = 0.;
  __syncthreads();
  for(int i = t; i < n; i += 1024)
    for(int k = ia[i]; k < ia[i + 1]; ++k) D[i * n + ja[k]] = val[k];
  __syncthreads();
  double m[2][2];
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      m[a][b] = (i < n && j < n) ? D[i * n + j] : 0.;
    }
  for(int k = 0; k < n; ++k) {
    const int ka = k >> 5, kr = k & 31;
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
      for(int b = 0; b < 2; ++b) {
        const int i = ty + 32 * a, j = tx + 32 * b;
        if(a == ka && ty == kr && j < n) prow[j] = m[a][b];
        if(b == ka && tx == kr && i < n) pcol[i] = m[a][b];
      }
    __syncthreads();
    const double piv = prow[k];
    const double p = piv > 0. ? 1. / piv : 0.;
    if(t == 0 && !(piv > 0.)) flag[0] = 1;
    double pc[2], pr[2];
#pragma unroll
    for(int a = 0; a < 2; ++a) pc[a] = (ty + 32 * a < n) ? pcol[ty + 32 * a] : 0.;
#pragma unroll
    for(int b = 0; b < 2; ++b) pr[b] = (tx + 32 * b < n) ? prow[tx + 32 * b] * p : 0.;
#pragma unroll
    for(int a = 0; a < 2; ++a)
#pragma unroll
      for(int b = 0; b < 2; ++b) {
        const bool rk = (a == ka && ty == kr), ck = (b == ka && tx == kr);
        if(rk && ck) m[a][b] = p;
        else if(rk) m[a][b] *= p;
        else if(ck) m[a][b] = -m[a][b] * p;
        else m[a][b] -= pc[a] * pr[b];
      }
    __syncthreads();
  }
#pragma unroll
  for(int a = 0; a < 2; ++a)
#pragma unroll
    for(int b = 0; b < 2; ++b) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      if(i < n && j < n) {
        inv64[i * n + j] = m[a][b];
        inv32[i * n + j] = (float)m[a][b];
      }
    }
}

// ---------------------------------------------------------------------------
// Device: the V-cycle

// Restriction fused with the first smoothing step over aggregate-aligned
// tiles of TB member positions: u_i = l1_i b_i, r_i = b_i - sum_j a_ij l1_j b_j
// for the member rows (through LDS), then the first position of every
// aggregate sums its members into bNext.  The position data are the first
// loads of the kernel, so the row products start one load later than
// through a tile header.
template <typename T, typename TV, int LPR, int TB, bool PRE = false>
__global__ __launch_bounds__(TB) void amgDownP(AmgLev<T, TV> L, const T *__restrict__ b, T *__restrict__ bNext,
                                               const int *__restrict__ prow, const int2 *__restrict__ prg,
                                               const int *__restrict__ pagg, int nTiles, int xcd, const int *done,
                                               const int2 *__restrict__ tp, const int *__restrict__ act)
{
  __shared__ T ldsY[TB];
  __shared__ T rb[TB];
  __shared__ int ag[TB];
  if(done && *done) return;
  const int tile = amgTile(xcd);
  if(tile >= nTiles || amgSkip(tp, act, tile)) return;
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR, g = threadIdx.x / LPR;
  const int base = tile * TB;
  int kb[LPR], ke[LPR];
#pragma unroll
  for(int u = 0; u < LPR; ++u) {
    const int2 rg = prg[base + u * RPB + g];
    kb[u] = rg.x + sub;
    ke[u] = rg.y;
  }
  const int t = threadIdx.x;
  const int i = prow[base + t], a = pagg[base + t];
  T bi = (T)0, li = (T)0;
  if(i >= 0) {
    bi = b[i];
    li = (T)L.l1[i];
  }
  // PRE: u = l1 b is already stored for every row (amgScaleRows), one
  // gather per entry instead of two
  if constexpr(PRE)
    tileProductF<T, LPR, TB>(L.ja, L.val, kb, ke, ldsY, [&](int j) { return L.u[j]; });
  else
    tileProductF<T, LPR, TB>(L.ja, L.val, kb, ke, ldsY, [&](int j) { return (T)L.l1[j] * b[j]; });
  __syncthreads();
  if(i >= 0) {
    if constexpr(!PRE) L.u[i] = li * bi;
    rb[t] = bi - ldsY[t];
  }
  ag[t] = a;
  __syncthreads();
  if(a >= 0 && (t == 0 || ag[t - 1] != a)) {
    T s = (T)0;
    for(int q = t; q < TB && ag[q] == a; ++q) s += rb[q];
    bNext[a] = s;
  }
}

// Restriction fused with the first smoothing step, over aggregate tiles:
// u_i = l1_i b_i, r_i = b_i - sum_j a_ij l1_j b_j for the member rows of the
// tile (through LDS), then bNext[I] = sum of r over the members of I.
template <typename T, typename TV, int LPR, int TB>
__global__ __launch_bounds__(TB) void amgDown(AmgLev<T, TV> L, const T *__restrict__ b, T *__restrict__ bNext,
                                              const int *__restrict__ tileRow, const int *__restrict__ tileAgg,
                                              const int *done, const int2 *__restrict__ tp, const int *__restrict__ act)
{
  __shared__ T ldsY[TB];
  __shared__ T rb[kAmgTileMax];
  if(done && *done) return;
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR, g = threadIdx.x / LPR;
  const int tile = blockIdx.x;
  if(amgSkip(tp, act, tile)) return;
  const int r0 = tileRow[tile], r1 = tileRow[tile + 1];
  for(int pass = r0; pass < r1; pass += TB) {
    int kb[LPR], ke[LPR];
#pragma unroll
    for(int u = 0; u < LPR; ++u) {
      const int q = pass + u * RPB + g;
      if(q < r1) {
        const int2 rg = L.mrange[q];
        kb[u] = rg.x + sub;
        ke[u] = rg.y;
      } else {
        kb[u] = ke[u] = 0;
      }
    }
    tileProductF<T, LPR, TB>(L.ja, L.val, kb, ke, ldsY, [&](int j) { return (T)L.l1[j] * b[j]; });
    __syncthreads();
    const int q = pass + threadIdx.x;
    if(q < r1) {
      const int i = L.mlist[q];
      const T bi = b[i];
      L.u[i] = (T)L.l1[i] * bi;
      rb[q - r0] = bi - ldsY[threadIdx.x];
    }
    __syncthreads();
  }
  const int I1 = tileAgg[tile + 1];
  for(int I = tileAgg[tile] + threadIdx.x; I < I1; I += TB) {
    T s = (T)0;
    for(int q = L.mptr[I]; q < L.mptr[I + 1]; ++q) s += rb[q - r0];
    bNext[I] = s;
  }
}

// Prolongation fused with the second smoothing step, over row tiles:
// y_i = u_i + oc e[agg_i], x_i = y_i + l1_i (b_i - sum_j a_ij y_j).
template <typename T, typename TV, int LPR, int TB, bool PRE = false>
__global__ __launch_bounds__(TB) void amgUp(AmgLev<T, TV> L, const T *__restrict__ b, const T *__restrict__ e, T oc,
                                            T *__restrict__ x, int xcd, const int *done, const int2 *__restrict__ tp,
                                            const int *__restrict__ act)
{
  __shared__ T ldsY[TB];
  if(done && *done) return;
  const int n = L.n;
  const int nTiles = (n + TB - 1) / TB;
  for(int tile = amgTile(xcd); tile < nTiles; tile += gridDim.x) {
    if(amgSkip(tp, act, tile)) continue;  // uniform over the block
    const int base = tile * TB;
    int kb[LPR], ke[LPR];
    tileBounds<LPR, TB>(n, base, L.ia, kb, ke);
    // PRE: y = u + oc e[agg] is already stored in L.r (amgUpStart), one
    // gather per entry instead of three
    if constexpr(PRE)
      tileProductF<T, LPR, TB>(L.ja, L.val, kb, ke, ldsY, [&](int j) { return L.r[j]; });
    else
      tileProductF<T, LPR, TB>(L.ja, L.val, kb, ke, ldsY, [&](int j) { return L.u[j] + oc * e[L.agg[j]]; });
    __syncthreads();
    const int row = base + threadIdx.x;
    if(row < n) {
      const T yi = PRE ? L.r[row] : L.u[row] + oc * e[L.agg[row]];
      x[row] = yi + (T)L.l1[row] * (b[row] - ldsY[threadIdx.x]);
    }
    __syncthreads();
  }
}

// Row passes of the large levels (the gathered vector stored once, so that
// the level kernels' products gather one vector per entry): over the same row
// blocks, skip flags and block order as amgUp.
template <typename T, typename TV, int TB>
__global__ __launch_bounds__(TB) void amgScaleRows(AmgLev<T, TV> L, const T *__restrict__ b, int xcd, const int *done,
                                                   const int2 *__restrict__ tp, const int *__restrict__ act)
{
  if(done && *done) return;
  const int nTiles = (L.n + TB - 1) / TB;
  for(int tile = amgTile(xcd); tile < nTiles; tile += gridDim.x) {
    if(amgSkip(tp, act, tile)) continue;
    const int row = tile * TB + threadIdx.x;
    if(row < L.n) L.u[row] = (T)L.l1[row] * b[row];
  }
}

template <typename T, typename TV, int TB>
__global__ __launch_bounds__(TB) void amgUpStart(AmgLev<T, TV> L, const T *__restrict__ e, T oc, int xcd,
                                                 const int *done, const int2 *__restrict__ tp,
                                                 const int *__restrict__ act)
{
  if(done && *done) return;
  const int nTiles = (L.n + TB - 1) / TB;
  for(int tile = amgTile(xcd); tile < nTiles; tile += gridDim.x) {
    if(amgSkip(tp, act, tile)) continue;
    const int row = tile * TB + threadIdx.x;
    if(row < L.n) L.r[row] = L.u[row] + oc * e[L.agg[row]];
  }
}

// One restriction step of the small-level kernel (all threads): u = l1 b,
// r = b - A u on the rows, then bn[I] = sum of r over the members of I.
template <typename T, typename TV, typename IA, typename IJ, typename IX>
__device__ __forceinline__ void amgBlockDown(int n, int nc, const IA *ia, const IJ *ja, const TV *val, const TV *l1,
                                             const IX *mptr, const IX *mlist, const T *b, T *u, T *r, T *bn)
{
  const int t = threadIdx.x, nt = blockDim.x;
  for(int i = t; i < n; i += nt) {
    T acc = (T)0;
    const int k1 = ia[i + 1];
    for(int k = ia[i]; k < k1; k += 4) {
      int j[4];
      T v[4];
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        j[q] = k + q < k1 ? (int)ja[k + q] : 0;
        v[q] = k + q < k1 ? (T)val[k + q] : (T)0;
      }
#pragma unroll
      for(int q = 0; q < 4; ++q) acc += v[q] * ((T)l1[j[q]] * b[j[q]]);
    }
    const T bi = b[i];
    u[i] = (T)l1[i] * bi;
    r[i] = bi - acc;
  }
  __syncthreads();
  for(int I = t; I < nc; I += nt) {
    T acc = (T)0;
    for(int q = mptr[I]; q < mptr[I + 1]; ++q) acc += r[mlist[q]];
    bn[I] = acc;
  }
  __syncthreads();
}

// One prolongation step (all threads): y = u + oc e[agg], x = y + l1 (b - A y);
// x goes to xg (global) when given, else over b.
template <typename T, typename TV, typename IA, typename IJ, typename IX>
__device__ __forceinline__ void amgBlockUp(int n, const IA *ia, const IJ *ja, const TV *val, const TV *l1,
                                           const IX *agg, T *b, const T *u, T *y, const T *e, T oc, T *xg)
{
  const int t = threadIdx.x, nt = blockDim.x;
  for(int i = t; i < n; i += nt) y[i] = u[i] + oc * e[agg[i]];
  __syncthreads();
  for(int i = t; i < n; i += nt) {
    T acc = (T)0;
    const int k1 = ia[i + 1];
    for(int k = ia[i]; k < k1; k += 4) {
      int j[4];
      T v[4];
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        j[q] = k + q < k1 ? (int)ja[k + q] : 0;
        v[q] = k + q < k1 ? (T)val[k + q] : (T)0;
      }
#pragma unroll
      for(int q = 0; q < 4; ++q) acc += v[q] * y[j[q]];
    }
    const T xi = y[i] + (T)l1[i] * (b[i] - acc);
    if(xg) xg[i] = xi;
    else b[i] = xi;
  }
  __syncthreads();
}

// Levels from .. coarsest - 1 and the coarsest solve in one workgroup.  The
// vectors of every level live in LDS; a level's matrix, l1 and aggregates are
// staged in LDS too (16-bit column ids) unless plan.mglob marks them to be
// read from global memory (L2-resident: the larger levels).  Phases are
// separated by workgroup barriers.  Level `from` reads b (b0 when from == 0)
// and writes its result to x (x0 when from == 0).
template <typename T, typename TV>
__global__ __launch_bounds__(kAmgBlockTB) void amgBlock(const AmgLev<T, TV> *__restrict__ lv, AmgBlockPlan pl,
                                                        const T *__restrict__ cinv, T oc, const T *b0, T *x0,
                                                        const int *done)
{
  __shared__ __align__(16) unsigned char lds[kAmgBlockLds];
  if(done && *done) return;
  const int t = threadIdx.x, nt = blockDim.x;
  const int nLev = pl.coarsest - pl.from;
  for(int s = 0; s < nLev; ++s) {
    if(pl.mglob[s]) continue;
    const AmgLev<T, TV> L = lv[pl.from + s];
    const int n = pl.n[s], nnz = pl.nnz[s], nc = pl.nc[s];
    int *ia = (int *)(lds + pl.oIa[s]);
    unsigned short *ja = (unsigned short *)(lds + pl.oJa[s]);
    TV *val = (TV *)(lds + pl.oVal[s]), *l1 = (TV *)(lds + pl.oL1[s]);
    unsigned short *agg = (unsigned short *)(lds + pl.oAgg[s]), *mptr = (unsigned short *)(lds + pl.oMptr[s]),
                   *mlist = (unsigned short *)(lds + pl.oMlist[s]);
    for(int i = t; i <= n; i += nt) ia[i] = L.ia[i];
    for(int k = t; k < nnz; k += nt) {
      ja[k] = (unsigned short)L.ja[k];
      val[k] = L.val[k];
    }
    for(int i = t; i < n; i += nt) {
      l1[i] = L.l1[i];
      agg[i] = (unsigned short)L.agg[i];
      mlist[i] = (unsigned short)L.mlist[i];
    }
    for(int I = t; I <= nc; I += nt) mptr[I] = (unsigned short)L.mptr[I];
  }
  {
    const T *bsrc = pl.from == 0 ? b0 : lv[pl.from].b;
    T *b = (T *)(lds + pl.oB[0]);
    for(int i = t; i < pl.n[0]; i += nt) b[i] = bsrc[i];
  }
  __syncthreads();
  for(int s = 0; s < nLev; ++s) {
    const int n = pl.n[s], nc = pl.nc[s];
    const T *b = (const T *)(lds + pl.oB[s]);
    T *u = (T *)(lds + pl.oU[s]), *r = (T *)(lds + pl.oR[s]);
    T *bn = (T *)(lds + pl.oB[s + 1]);
    if(pl.mglob[s]) {
      const AmgLev<T, TV> L = lv[pl.from + s];
      amgBlockDown<T, TV>(n, nc, L.ia, L.ja, L.val, L.l1, L.mptr, L.mlist, b, u, r, bn);
    } else {
      amgBlockDown<T, TV>(n, nc, (const int *)(lds + pl.oIa[s]), (const unsigned short *)(lds + pl.oJa[s]),
                          (const TV *)(lds + pl.oVal[s]), (const TV *)(lds + pl.oL1[s]),
                          (const unsigned short *)(lds + pl.oMptr[s]), (const unsigned short *)(lds + pl.oMlist[s]),
                          b, u, r, bn);
    }
  }
  {
    const int nC = pl.n[nLev];
    const T *b = (const T *)(lds + pl.oB[nLev]);
    T *x = nLev == 0 ? (pl.from == 0 ? x0 : lv[pl.from].x) : (T *)(lds + pl.oCx);
    if(pl.coarseSweeps == 0) {
      const T *ci = cinv;  // global: row i of the inverse per thread
      for(int i = t; i < nC; i += nt) {
        T acc = (T)0;
        for(int j = 0; j < nC; ++j) acc += ci[i * nC + j] * b[j];
        x[i] = acc;
      }
    } else {
      // stalled coarsening: l1-Jacobi sweeps from zero on the coarsest CSR
      // (global memory; its u and r vectors as the two iterates)
      const AmgLev<T, TV> C = lv[pl.coarsest];
      const T sc = (T)pl.coarseScale;
      T *xa = C.u, *xb = C.r;
      for(int i = t; i < nC; i += nt) xa[i] = sc * (T)C.l1[i] * b[i];
      __syncthreads();
      for(int sw = 1; sw < pl.coarseSweeps; ++sw) {
        for(int i = t; i < nC; i += nt) {
          T acc = (T)0;
          for(int k = C.ia[i]; k < C.ia[i + 1]; ++k) acc += (T)C.val[k] * xa[C.ja[k]];
          xb[i] = xa[i] + sc * (T)C.l1[i] * (b[i] - acc);
        }
        __syncthreads();
        T *tmp = xa;
        xa = xb;
        xb = tmp;
      }
      for(int i = t; i < nC; i += nt) x[i] = xa[i];
    }
    __syncthreads();
  }
  for(int s = nLev - 1; s >= 0; --s) {
    const int n = pl.n[s];
    T *b = (T *)(lds + pl.oB[s]);
    const T *u = (const T *)(lds + pl.oU[s]);
    T *y = (T *)(lds + pl.oR[s]);
    // e = x of level s + 1: the coarsest solution, or level s + 1's b slot
    const T *e = (s + 1 == nLev) ? (const T *)(lds + pl.oCx) : (const T *)(lds + pl.oB[s + 1]);
    T *xg = s == 0 ? (pl.from == 0 ? x0 : lv[pl.from].x) : nullptr;
    if(pl.mglob[s]) {
      const AmgLev<T, TV> L = lv[pl.from + s];
      amgBlockUp<T, TV>(n, L.ia, L.ja, L.val, L.l1, L.agg, b, u, y, e, oc, xg);
    } else {
      amgBlockUp<T, TV>(n, (const int *)(lds + pl.oIa[s]), (const unsigned short *)(lds + pl.oJa[s]),
                        (const TV *)(lds + pl.oVal[s]), (const TV *)(lds + pl.oL1[s]),
                        (const unsigned short *)(lds + pl.oAgg[s]), b, u, y, e, oc, xg);
    }
  }
}

template <typename T>
T *pick(double *a, float *b);
template <>
double *pick<double>(double *a, float *) { return a; }
template <>
float *pick<float>(double *, float *b) { return b; }

template <typename TV>
const TV *pickV(const double *a, const float *b);
template <>
const double *pickV<double>(const double *a, const float *) { return a; }
template <>
const float *pickV<float>(const double *, const float *b) { return b; }

template <typename T, typename TV>
AmgLev<T, TV> amgView(const FeAmgLevel &h)
{
  AmgLev<T, TV> L;
  L.n = h.n;
  L.nc = h.nc;
  L.ia = h.ia;
  L.ja = h.ja;
  L.val = pickV<TV>(h.val64, h.val32);
  L.l1 = pickV<TV>(h.l1_64, h.l1_32);
  L.agg = h.agg;
  L.mptr = h.mptr;
  L.mlist = h.mlist;
  L.mrange = h.mrange;
  L.b = pick<T>(h.b64, h.b32);
  L.u = pick<T>(h.u64, h.u32);
  L.x = pick<T>(h.x64, h.x32);
  L.r = pick<T>(h.r64, h.r32);
  return L;
}

template <typename T, typename TV>
int amgUploadDesc(fe_hip_ctx *ctx, FeAmg &A, int slot)
{
  std::vector<AmgLev<T, TV>> h;
  for(const FeAmgLevel &l : A.lv) h.push_back(amgView<T, TV>(l));
  AmgLev<T, TV> *d = nullptr;
  if(int rc = feAlloc(ctx, &d, h.size())) return rc;
  FE_HIP_TRY(ctx, hipMemcpy(d, h.data(), sizeof(AmgLev<T, TV>) * h.size(), hipMemcpyHostToDevice));
  A.desc[slot] = d;
  return FE_STATUS_OK;
}

template <typename T>
int upload(fe_hip_ctx *ctx, T **d, const std::vector<T> &h)
{
  if(int rc = feAlloc(ctx, d, h.size())) return rc;
  if(!h.empty()) FE_HIP_TRY(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return FE_STATUS_OK;
}

}  // namespace

// Hierarchy of the current CSR pattern (once per context).
int feAmgSetup(fe_hip_ctx *ctx)
{
  if(ctx->amg) return FE_STATUS_OK;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR before an AMG solve");
  const int n = ctx->n;
  HostLevel f;
  f.n = n;
  f.ia.resize(n + 1);
  FE_HIP_TRY(ctx, hipMemcpy(f.ia.data(), ctx->d_ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  f.ja.resize(f.ia[n]);
  FE_HIP_TRY(ctx, hipMemcpy(f.ja.data(), ctx->d_ja, sizeof(int) * f.ja.size(), hipMemcpyDeviceToHost));
  std::vector<HostLevel> H;
  H.push_back(std::move(f));
  int coarseSweeps = 0;
  while(H.back().n > kAmgDense) {
    HostLevel &c = H.back();
    c.nc = amgAggregate(c.n, c.ia, c.ja, c.agg);
    if((int)H.size() >= kAmgMaxLevels || c.nc > kAmgStall * c.n) {
      // coarsening has stalled (e.g. every problem of a block-diagonal system
      // is down to a node or two): this level is the coarsest, swept
      c.nc = 0;
      c.agg.clear();
      coarseSweeps = kAmgCoarseSweeps;
      break;
    }
    c.mptr.assign(c.nc + 1, 0);
    for(int i = 0; i < c.n; ++i) c.mptr[c.agg[i] + 1]++;
    for(int I = 0; I < c.nc; ++I) c.mptr[I + 1] += c.mptr[I];
    c.mlist.resize(c.n);
    {
      std::vector<int> fill(c.mptr.begin(), c.mptr.end() - 1);
      for(int i = 0; i < c.n; ++i) c.mlist[fill[c.agg[i]]++] = i;
    }
    // aggregate tiles of about kAmgTileRows member rows
    c.tileRow.assign(1, 0);
    c.tileAgg.assign(1, 0);
    int rows = 0;
    for(int I = 0; I < c.nc; ++I) {
      const int sz = c.mptr[I + 1] - c.mptr[I];
      c.maxAgg = std::max(c.maxAgg, sz);
      FE_REQUIRE(ctx, sz <= kAmgTileMax, "AMG: an aggregate is larger than the tile limit");
      if(rows > 0 && rows + sz > kAmgTileRows) {
        c.tileRow.push_back(c.mptr[I]);
        c.tileAgg.push_back(I);
        rows = 0;
      }
      rows += sz;
    }
    c.tileRow.push_back(c.n);
    c.tileAgg.push_back(c.nc);
    HostLevel nx;
    nx.n = c.nc;
    amgGalerkinPattern(c, nx.ia, nx.ja, nx.cptr, nx.clist);
    H.push_back(std::move(nx));
  }
  FeAmg *A = new FeAmg();
  ctx->amg = A;
  A->coarseSweeps = coarseSweeps;
  ctx->amgHost.assign(H.size(), FeAmgHostLevel());
  ctx->amgHostSweeps = coarseSweeps;
  ctx->amgHostMaxAgg = 0;
  for(const HostLevel &h : H) ctx->amgHostMaxAgg = std::max(ctx->amgHostMaxAgg, h.maxAgg);
  for(size_t l = 0; l < H.size(); ++l) {
    FeAmgHostLevel &k = ctx->amgHost[l];
    k.n = H[l].n;
    k.nc = H[l].nc;
    k.ia = H[l].ia;
    k.ja = H[l].ja;
    k.agg = H[l].agg;
    k.cptr = H[l].cptr;
    k.clist = H[l].clist;
  }
  A->lv.resize(H.size());
  for(size_t l = 0; l < H.size(); ++l) {
    HostLevel &h = H[l];
    FeAmgLevel &d = A->lv[l];
    d.n = h.n;
    d.nc = h.nc;
    d.nnz = h.ia[h.n];
    d.maxAgg = h.maxAgg;
    if(l == 0) {
      d.ia = ctx->d_ia;
      d.ja = ctx->d_ja;
      d.val64 = ctx->d_val;
    } else {
      if(int rc = upload(ctx, &d.ia, h.ia)) return rc;
      if(int rc = upload(ctx, &d.ja, h.ja)) return rc;
      if(int rc = upload(ctx, &d.cptr, h.cptr)) return rc;
      if(int rc = upload(ctx, &d.clist, h.clist)) return rc;
      if(int rc = feAlloc(ctx, &d.val64, d.nnz)) return rc;
    }
    if(int rc = feAlloc(ctx, &d.val32, d.nnz)) return rc;
    if(int rc = feAlloc(ctx, &d.l1_64, d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.l1_32, d.n)) return rc;
    if(l + 1 < H.size()) {
      if(int rc = upload(ctx, &d.agg, h.agg)) return rc;
      if(int rc = upload(ctx, &d.mptr, h.mptr)) return rc;
      if(int rc = upload(ctx, &d.mlist, h.mlist)) return rc;
      if(h.maxAgg <= kAmgTB) {
        std::vector<int> prow, pagg;
        std::vector<int2> prg;
        for(int I = 0; I < h.nc; ++I) {
          const int sz = h.mptr[I + 1] - h.mptr[I];
          const int used = (int)prow.size() % kAmgTB;
          if(used + sz > kAmgTB)
            for(int q = used; q < kAmgTB; ++q) {
              prow.push_back(-1);
              pagg.push_back(-1);
              prg.push_back(make_int2(0, 0));
            }
          for(int q = h.mptr[I]; q < h.mptr[I + 1]; ++q) {
            const int i = h.mlist[q];
            prow.push_back(i);
            pagg.push_back(I);
            prg.push_back(make_int2(h.ia[i], h.ia[i + 1]));
          }
        }
        while(prow.size() % kAmgTB) {
          prow.push_back(-1);
          pagg.push_back(-1);
          prg.push_back(make_int2(0, 0));
        }
        d.padTiles = (int)(prow.size() / kAmgTB);
        d.hPadAgg.assign(d.padTiles, make_int2(-1, -1));
        for(size_t q = 0; q < pagg.size(); ++q)
          if(pagg[q] >= 0) {
            int2 &r = d.hPadAgg[q / kAmgTB];
            if(r.x < 0) r.x = pagg[q];
            r.y = pagg[q];
          }
        if(int rc = upload(ctx, &d.prow, prow)) return rc;
        if(int rc = upload(ctx, &d.pagg, pagg)) return rc;
        if(int rc = upload(ctx, &d.prg, prg)) return rc;
      }
      {
        std::vector<int2> rg(h.n);
        for(int q = 0; q < h.n; ++q) rg[q] = make_int2(h.ia[h.mlist[q]], h.ia[h.mlist[q] + 1]);
        if(int rc = upload(ctx, &d.mrange, rg)) return rc;
      }
      if(int rc = upload(ctx, &d.tileRow, h.tileRow)) return rc;
      if(int rc = upload(ctx, &d.tileAgg, h.tileAgg)) return rc;
      d.nTiles = (int)h.tileRow.size() - 1;
      d.hTileAgg = h.tileAgg;
    }
    // level 0 reads the Krylov vectors for b and x
    if(l > 0) {
      if(int rc = feAlloc(ctx, &d.b64, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.x64, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.b32, d.n)) return rc;
      if(int rc = feAlloc(ctx, &d.x32, d.n)) return rc;
    }
    if(int rc = feAlloc(ctx, &d.u64, d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.r64, d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.u32, d.n)) return rc;
    if(int rc = feAlloc(ctx, &d.r32, d.n)) return rc;
  }
  const size_t nC = coarseSweeps ? 1 : (size_t)A->lv.back().n;
  if(int rc = feAlloc(ctx, &A->cinv64, nC * nC)) return rc;
  if(int rc = feAlloc(ctx, &A->cinv32, nC * nC)) return rc;
  if(int rc = feAlloc(ctx, &A->flag, 1)) return rc;
  if(int rc = amgUploadDesc<double, double>(ctx, *A, 0)) return rc;
  if(int rc = amgUploadDesc<double, float>(ctx, *A, 1)) return rc;
  if(int rc = amgUploadDesc<float, float>(ctx, *A, 2)) return rc;
  return FE_STATUS_OK;
}

int feAmgProblemTiles(fe_hip_ctx *ctx, int nprob, const int *rows, const int *act)
{
  if(!act) {
    if(ctx->amg) ctx->amg->act = nullptr;
    return FE_STATUS_OK;
  }
  FE_REQUIRE(ctx, ctx->amg, "AMG problem tiles: set up the hierarchy first");
  FeAmg &A = *ctx->amg;
  A.act = nullptr;
  FE_REQUIRE(ctx, nprob >= 1 && rows, "AMG problem tiles: no problems");
  const std::vector<int> key(rows, rows + nprob + 1);
  if(key != A.probRows) {
    FE_REQUIRE(ctx, key[0] == 0 && key[nprob] == A.lv[0].n && std::is_sorted(key.begin(), key.end()),
               "AMG problem tiles: the problem rows do not partition the system");
    FE_REQUIRE(ctx, ctx->amgHost.size() == A.lv.size(), "AMG problem tiles: no host hierarchy");
    // problem of row r of a level with problem offsets o (empty problems own no row)
    auto probOf = [nprob](const std::vector<int> &o, int r) {
      return std::min(nprob - 1, (int)(std::upper_bound(o.begin(), o.end(), r) - o.begin()) - 1);
    };
    auto put = [ctx](int2 **d, const std::vector<int2> &h) -> int {
      if(h.empty()) return FE_STATUS_OK;
      if(!*d)
        if(int rc = feAlloc(ctx, d, h.size())) return rc;
      FE_HIP_TRY(ctx, hipMemcpy(*d, h.data(), sizeof(int2) * h.size(), hipMemcpyHostToDevice));
      return FE_STATUS_OK;
    };
    std::vector<int> off = key;
    for(size_t l = 0; l < A.lv.size(); ++l) {
      FeAmgLevel &d = A.lv[l];
      std::vector<int> offN;
      if(l + 1 < A.lv.size()) {
        // roots are numbered in row order and aggregates never join rows of
        // two problems (the system is block-diagonal): problem p's aggregates
        // are the range [offN[p], offN[p + 1]) of the next level
        const std::vector<int> &agg = ctx->amgHost[l].agg;
        offN.assign(nprob + 1, d.nc);
        for(int p = 0; p < nprob; ++p)
          for(int i = off[p]; i < off[p + 1]; ++i) offN[p] = std::min(offN[p], agg[i]);
        for(int p = nprob - 1; p >= 0; --p) offN[p] = std::min(offN[p], offN[p + 1]);
        for(int p = 0; p < nprob; ++p)
          for(int i = off[p]; i < off[p + 1]; ++i)
            FE_REQUIRE(ctx, agg[i] >= offN[p] && agg[i] < offN[p + 1],
                       "AMG problem tiles: an aggregate joins rows of two problems");
        std::vector<int2> pad(d.padTiles), til(d.nTiles);
        for(int t = 0; t < d.padTiles; ++t) {
          const int2 r = d.hPadAgg[t];
          pad[t] = r.x < 0 ? make_int2(0, nprob - 1) : make_int2(probOf(offN, r.x), probOf(offN, r.y));
        }
        for(int t = 0; t < d.nTiles; ++t)
          til[t] = make_int2(probOf(offN, d.hTileAgg[t]), probOf(offN, std::max(d.hTileAgg[t], d.hTileAgg[t + 1] - 1)));
        if(int rc = put(&d.probPad, pad)) return rc;
        if(int rc = put(&d.probTile, til)) return rc;
      }
      const int nUp = (d.n + kAmgTB - 1) / kAmgTB;
      std::vector<int2> up(nUp);
      for(int t = 0; t < nUp; ++t)
        up[t] = make_int2(probOf(off, t * kAmgTB), probOf(off, std::min(d.n, (t + 1) * kAmgTB) - 1));
      if(int rc = put(&d.probUp, up)) return rc;
      off.swap(offN);
    }
    // rows of stopped problems keep their last values; start them finite
    // (the coarsest dense inverse multiplies them by exact zeros)
    for(size_t l = 0; l < A.lv.size(); ++l) {
      FeAmgLevel &d = A.lv[l];
      for(double *q : {d.b64, d.x64, d.u64, d.r64})
        if(q) FE_HIP_TRY(ctx, hipMemsetAsync(q, 0, sizeof(double) * d.n, ctx->stream));
      for(float *q : {d.b32, d.x32, d.u32, d.r32})
        if(q) FE_HIP_TRY(ctx, hipMemsetAsync(q, 0, sizeof(float) * d.n, ctx->stream));
    }
    A.probRows = key;
  }
  A.act = act;
  return FE_STATUS_OK;
}

void feAmgRelease(fe_hip_ctx *ctx)
{
  delete ctx->amg;  // device buffers are owned by the context
  ctx->amg = nullptr;
}

namespace {

// LDS plan of the small-level kernel for vectors of T and values of TV: the
// most trailing levels (each of at most blockRows rows) that fit.  The
// coarsest level alone always fits (kAmgDense^2 values of T).
template <typename T, typename TV>
AmgBlockPlan amgPlan(const FeAmg &A, int blockRows)
{
  const int Lc = (int)A.lv.size() - 1;
  AmgBlockPlan pl;
  for(int from = 0; from <= Lc; ++from) {
    if(Lc - from > kAmgBlockLev) continue;
    bool ok = true;
    for(int l = from; l < Lc; ++l) ok = ok && A.lv[l].n <= blockRows && A.lv[l].n <= 65535;
    if(!ok) continue;
    int bytes = 0;
    auto take = [&](size_t b) {
      const int o = bytes;
      bytes += (int)((b + 15) & ~(size_t)15);
      return o;
    };
    pl = AmgBlockPlan();
    pl.from = from;
    pl.coarsest = Lc;
    // vectors of every level
    for(int l = from; l < Lc; ++l) {
      const int s = l - from;
      const FeAmgLevel &h = A.lv[l];
      pl.n[s] = h.n;
      pl.nc[s] = h.nc;
      pl.nnz[s] = (int)h.nnz;
      pl.oB[s] = take(sizeof(T) * h.n);
      pl.oU[s] = take(sizeof(T) * h.n);
      pl.oR[s] = take(sizeof(T) * h.n);
      pl.mglob[s] = 1;
    }
    const int nC = A.lv[Lc].n;
    pl.n[Lc - from] = nC;
    pl.oB[Lc - from] = take(sizeof(T) * nC);
    pl.oCx = take(sizeof(T) * nC);
    if(bytes > kAmgBlockLds) continue;
    // stage matrices and aggregates from the coarsest level up while they fit
    for(int l = Lc - 1; l >= from; --l) {
      const int s = l - from;
      const FeAmgLevel &h = A.lv[l];
      const size_t need = ((sizeof(int) * (h.n + 1) + 15) & ~15) + ((2 * h.nnz + 15) & ~15) +
                          ((sizeof(TV) * h.nnz + 15) & ~15) + ((sizeof(TV) * h.n + 15) & ~15) + ((2 * h.n + 15) & ~15) +
                          ((2 * (h.nc + 1) + 15) & ~15) + ((2 * h.n + 15) & ~15);
      if(bytes + (int)need > kAmgBlockLds) break;
      pl.oIa[s] = take(sizeof(int) * (h.n + 1));
      pl.oJa[s] = take(2 * h.nnz);
      pl.oVal[s] = take(sizeof(TV) * h.nnz);
      pl.oL1[s] = take(sizeof(TV) * h.n);
      pl.oAgg[s] = take(2 * h.n);
      pl.oMptr[s] = take(2 * (h.nc + 1));
      pl.oMlist[s] = take(2 * h.n);
      pl.mglob[s] = 0;
    }
    pl.bytes = bytes;
    pl.coarseSweeps = A.coarseSweeps;
    pl.coarseScale = (float)(1. / A.c);
    return pl;
  }
  return pl;  // not reached: from = Lc fits
}

}  // namespace

// Numeric set-up for the current values: fine l1 and FP32 values, the
// Galerkin sums of every coarser level, the dense coarsest inverse.  A
// non-positive coarsest pivot is reported by feAmgCheck after the solve.
int feAmgNumeric(fe_hip_ctx *ctx)
{
  FeAmg &A = *ctx->amg;
  A.c = ctx->amgOmega;
  A.oc = ctx->amgOverCorrect;
  A.plan[0] = amgPlan<double, double>(A, ctx->amgBlockRows);
  A.plan[1] = amgPlan<double, float>(A, ctx->amgBlockRows);
  A.plan[2] = amgPlan<float, float>(A, ctx->amgBlockRows);
  for(const AmgBlockPlan &pl : A.plan)
    FE_REQUIRE(ctx, pl.bytes <= kAmgBlockLds, "AMG: the coarsest level does not fit the small-level kernel's LDS");
  FeAmgLevel &f = A.lv[0];
  f.val64 = ctx->d_val;
  {
    const int g = std::min(4096, (f.n * 8 + 255) / 256);
    hipLaunchKernelGGL(amgFineL1, dim3(g), dim3(256), 0, ctx->stream, f.n, f.ia, f.val64, A.c, f.l1_64, f.l1_32,
                       f.val32);
  }
  for(size_t l = 1; l < A.lv.size(); ++l) {
    FeAmgLevel &c = A.lv[l];
    const int ge = std::max(1, std::min(4096, (int)((c.nnz + 255) / 256)));
    hipLaunchKernelGGL(amgGalerkin, dim3(ge), dim3(256), 0, ctx->stream, (int)c.nnz, c.cptr, c.clist,
                       A.lv[l - 1].val64, c.val64, c.val32);
    const int gr = std::max(1, std::min(4096, (c.n * 8 + 255) / 256));
    hipLaunchKernelGGL(amgRowL1, dim3(gr), dim3(256), 0, ctx->stream, c.n, c.ia, c.val64, A.c, c.l1_64, c.l1_32);
  }
  FE_HIP_TRY(ctx, hipMemsetAsync(A.flag, 0, sizeof(int), ctx->stream));
  const FeAmgLevel &C = A.lv.back();
  if(!A.coarseSweeps)
    hipLaunchKernelGGL(amgDenseInverse, dim3(1), dim3(1024), 0, ctx->stream, C.n, C.ia, C.ja, C.val64, A.cinv64,
                       A.cinv32, A.flag);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

// Error when the last numeric set-up met a non-positive coarsest pivot
// (synchronises the stream).
int feAmgCheck(fe_hip_ctx *ctx)
{
  if(!ctx->amg) return FE_STATUS_OK;
  int flag = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&flag, ctx->amg->flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, flag == 0, "AMG: the coarsest operator is not positive definite");
  return FE_STATUS_OK;
}

namespace {

template <typename T, typename TV>
int amgApplyT(fe_hip_ctx *ctx, const T *in, T *out, const int *done, int slot)
{
  FeAmg &A = *ctx->amg;
  const AmgBlockPlan &pl = A.plan[slot];
  const int B = pl.from;
  const T oc = (T)A.oc;
  const int *act = A.act;
  for(int l = 0; l < B; ++l) {
    const FeAmgLevel &h = A.lv[l];
    AmgLev<T, TV> L = amgView<T, TV>(h);
    const T *b = l == 0 ? in : L.b;
    T *bn = pick<T>(A.lv[l + 1].b64, A.lv[l + 1].b32);
    const dim3 gp(ctx->amgXcd ? (h.padTiles + 7) & ~7 : h.padTiles);
    const bool pad = h.padTiles > 0 && ctx->amgPadTiles;
    const int2 *tp = pad ? h.probPad : h.probTile;
    int Gr = std::max(1, (h.n + kAmgTB - 1) / kAmgTB);
    if(ctx->amgXcd) Gr = (Gr + 7) & ~7;
    const bool preD = ctx->amgPreRows > 0 && h.n >= ctx->amgPreRows && (ctx->amgLanes == 2 || ctx->amgLanes == 4);
    if(pad && preD) {
      hipLaunchKernelGGL((amgScaleRows<T, TV, kAmgTB>), dim3(Gr), dim3(kAmgTB), 0, ctx->stream, L, b, ctx->amgXcd, done,
                         (const int2 *)h.probUp, act);
      if(ctx->amgLanes == 2)
        hipLaunchKernelGGL((amgDownP<T, TV, 2, kAmgTB, true>), gp, dim3(kAmgTB), 0, ctx->stream, L, b, bn, h.prow,
                           h.prg, h.pagg, h.padTiles, ctx->amgXcd, done, tp, act);
      else
        hipLaunchKernelGGL((amgDownP<T, TV, 4, kAmgTB, true>), gp, dim3(kAmgTB), 0, ctx->stream, L, b, bn, h.prow,
                           h.prg, h.pagg, h.padTiles, ctx->amgXcd, done, tp, act);
    } else if(pad && ctx->amgLanes == 2)
      hipLaunchKernelGGL((amgDownP<T, TV, 2, kAmgTB>), gp, dim3(kAmgTB), 0, ctx->stream, L, b, bn, h.prow, h.prg,
                         h.pagg, h.padTiles, ctx->amgXcd, done, tp, act);
    else if(pad && ctx->amgLanes == 8)
      hipLaunchKernelGGL((amgDownP<T, TV, 8, kAmgTB>), gp, dim3(kAmgTB), 0, ctx->stream, L, b, bn, h.prow, h.prg,
                         h.pagg, h.padTiles, ctx->amgXcd, done, tp, act);
    else if(pad)
      hipLaunchKernelGGL((amgDownP<T, TV, kAmgLPR, kAmgTB>), gp, dim3(kAmgTB), 0, ctx->stream, L, b, bn, h.prow,
                         h.prg, h.pagg, h.padTiles, ctx->amgXcd, done, tp, act);
    else
      hipLaunchKernelGGL((amgDown<T, TV, kAmgLPR, kAmgTB>), dim3(h.nTiles), dim3(kAmgTB), 0, ctx->stream, L, b, bn,
                         h.tileRow, h.tileAgg, done, tp, act);
  }
  hipLaunchKernelGGL((amgBlock<T, TV>), dim3(1), dim3(kAmgBlockTB), 0, ctx->stream,
                     (const AmgLev<T, TV> *)A.desc[slot], pl, (const T *)pick<T>(A.cinv64, A.cinv32), oc, in, out,
                     done);
  for(int l = B - 1; l >= 0; --l) {
    const FeAmgLevel &h = A.lv[l];
    AmgLev<T, TV> L = amgView<T, TV>(h);
    const T *b = l == 0 ? in : L.b;
    T *x = l == 0 ? out : L.x;
    int G = std::max(1, (h.n + kAmgTB - 1) / kAmgTB);
    if(ctx->amgXcd) G = (G + 7) & ~7;
    const T *e = pick<T>(A.lv[l + 1].x64, A.lv[l + 1].x32);
    // profiling: events around the level-0 prolongation kernel (the dominant
    // kernel of the batched multigrid solves; its row pass stays outside)
    const bool pre = ctx->amgPreRows > 0 && h.n >= ctx->amgPreRows && (ctx->amgLanes == 2 || ctx->amgLanes == 4);
    if(pre)
      hipLaunchKernelGGL((amgUpStart<T, TV, kAmgTB>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, e, oc, ctx->amgXcd, done,
                         (const int2 *)h.probUp, act);
    if(l == 0 && ctx->profiling)
      if(int rc = feTimerMark(ctx, ctx->spmvTimer)) return rc;
    if(pre && ctx->amgLanes == 2) {
      hipLaunchKernelGGL((amgUp<T, TV, 2, kAmgTB, true>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, b, e, oc, x,
                         ctx->amgXcd, done, (const int2 *)h.probUp, act);
    } else if(pre) {
      hipLaunchKernelGGL((amgUp<T, TV, 4, kAmgTB, true>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, b, e, oc, x,
                         ctx->amgXcd, done, (const int2 *)h.probUp, act);
    } else if(ctx->amgLanes == 2)
      hipLaunchKernelGGL((amgUp<T, TV, 2, kAmgTB>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, b, e, oc, x, ctx->amgXcd,
                         done, (const int2 *)h.probUp, act);
    else if(ctx->amgLanes == 8)
      hipLaunchKernelGGL((amgUp<T, TV, 8, kAmgTB>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, b, e, oc, x, ctx->amgXcd,
                         done, (const int2 *)h.probUp, act);
    else
      hipLaunchKernelGGL((amgUp<T, TV, kAmgLPR, kAmgTB>), dim3(G), dim3(kAmgTB), 0, ctx->stream, L, b, e, oc, x,
                         ctx->amgXcd, done, (const int2 *)h.probUp, act);
    if(l == 0 && ctx->profiling)
      if(int rc = feTimerMark(ctx, ctx->spmvTimer)) return rc;
  }
  return FE_STATUS_OK;
}

}  // namespace

// out = M^-1 in (one V-cycle).  FP64 vectors read FP64 or FP32 level values
// (values32); FP32 vectors read FP32 values.  Skipped once *done is set.
int feAmgApply(fe_hip_ctx *ctx, const double *in, double *out, const int *done, bool values32)
{
  if(values32) return amgApplyT<double, float>(ctx, in, out, done, 1);
  return amgApplyT<double, double>(ctx, in, out, done, 0);
}

int feAmgApply(fe_hip_ctx *ctx, const float *in, float *out, const int *done, bool)
{
  return amgApplyT<float, float>(ctx, in, out, done, 2);
}

int feAmgPlanFrom(const fe_hip_ctx *ctx, int slot)
{
  return ctx->amg ? ctx->amg->plan[slot].from : -1;
}

int feAmgLevels(const fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap)
{
  if(!ctx->amg) {
    *levels = 0;
    return FE_STATUS_OK;
  }
  const FeAmg &A = *ctx->amg;
  *levels = (int)A.lv.size();
  for(int l = 0; l < (int)A.lv.size() && l < cap; ++l) {
    rows[l] = A.lv[l].n;
    nnz[l] = A.lv[l].nnz;
  }
  return FE_STATUS_OK;
}

extern "C" {

int fe_hip_amg_levels(fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap)
{
  if(!ctx || !levels || (cap > 0 && (!rows || !nnz))) return FE_STATUS_POINTER_ERROR;
  return feAmgLevels(ctx, levels, rows, nnz, cap);
}

int fe_hip_amg_apply(fe_hip_ctx *ctx, const double *r, double *z)
{
  if(!ctx || !r || !z) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if(int rc = feAmgSetup(ctx)) return rc;
  const bool nested = feAmgNestedUsable(ctx);
  if(!nested)
    if(int rc = feAmgNumeric(ctx)) return rc;
  const int n = ctx->n;
  double *d = nullptr;
  FE_HIP_TRY(ctx, hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
  int rc = FE_STATUS_OK;
  if(hipMemcpyAsync(d, r, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    ctx->setError("fe_hip_amg_apply: copy of r failed", __FILE__, __LINE__);
    rc = FE_STATUS_ERROR;
  }
  if(rc == FE_STATUS_OK)
    rc = nested ? feAmgPcgApply(ctx, (const double *)d, d + n)
                : feAmgApply(ctx, (const double *)d, d + n, nullptr, ctx->pcValues32 != 0);
  if(rc == FE_STATUS_OK && (hipGetLastError() != hipSuccess ||
                            hipMemcpyAsync(z, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                            hipStreamSynchronize(ctx->stream) != hipSuccess)) {
    ctx->setError("fe_hip_amg_apply: V-cycle failed", __FILE__, __LINE__);
    rc = FE_STATUS_ERROR;
  }
  (void)hipFree(d);
  if(rc == FE_STATUS_OK && !nested) rc = feAmgCheck(ctx);
  return rc;
}

}  // extern "C"
