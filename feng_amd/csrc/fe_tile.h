// Row-tile building blocks of the pipelined Jacobi-PCG kernels (single
// system in fe_krylov.hip, segmented batches in fe_batch.hip).  A block owns
// tiles of TB rows; in the product phase each group of LPR lanes takes LPR
// rows of the tile with their gathers interleaved, the row sums go through
// LDS, and in the update phase thread t updates row t of the tile.
#pragma once

#include "fe_hip_internal.h"

namespace fe_tile {

// First row tile of this block: when the grid is a multiple of 8 the 8 XCDs
// (block b runs on XCD b % 8) take 8 contiguous ranges of tiles, so gathers
// of neighbouring rows stay in one XCD's L2; tiles then advance by gridDim.
__device__ __forceinline__ int tileStart()
{
  const int G = gridDim.x;
  if(G & 7) return blockIdx.x;
  return (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
}

template <typename T>
struct TileVecs {
  T *x, *r, *w, *s, *z, *p, *dinv, *m[2];
};

template <typename T>
struct TileRow {
  T d, r, w, s, z, p, x;
};

template <typename T>
__device__ __forceinline__ void tileLoadRow(const TileVecs<T> &v, int row, int n, TileRow<T> &o)
{
  if(row < n) {
    o.d = v.dinv[row]; o.r = v.r[row]; o.w = v.w[row]; o.s = v.s[row];
    o.z = v.z[row]; o.p = v.p[row]; o.x = v.x[row];
  } else {
    o.d = o.r = o.w = o.s = o.z = o.p = o.x = (T)0;
  }
}

// Sums three per-thread values over the block; thread 0 stores them into the
// three partial arrays at blockIdx.x.  One barrier.
template <int TB>
__device__ __forceinline__ void tileStore3(double a, double b, double c, double *__restrict__ pa,
                                           double *__restrict__ pb, double *__restrict__ pc)
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
  if(threadIdx.x == 0) {
    double s0 = 0., s1 = 0., s2 = 0.;
#pragma unroll
    for(int i = 0; i < TB / 64; ++i) {
      s0 += lds[0][i];
      s1 += lds[1][i];
      s2 += lds[2][i];
    }
    pa[blockIdx.x] = s0;
    pb[blockIdx.x] = s1;
    pc[blockIdx.x] = s2;
  }
}

// y[row] = sum_k val[k] f(ja[k]) for the LPR rows base + u RPB + g of a
// tile (u < LPR), gathers of the LPR rows interleaved; results into ldsY.
template <typename T, int LPR, int TB, typename F, typename TV = T>
__device__ __forceinline__ void tileProductF(const int *__restrict__ ja, const TV *__restrict__ val,
                                             const int (&kb)[LPR], const int (&ke)[LPR], T *ldsY, F f)
{
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR, g = threadIdx.x / LPR;
  // the LPR rows of a lane group advance together: per step each lane loads
  // the column ids and values of KS entries of every row, then gathers
  constexpr int KS = LPR >= 16 ? 1 : 16 / LPR;
  T acc[LPR];
  int k0 = 0;
#pragma unroll
  for(int u = 0; u < LPR; ++u) acc[u] = (T)0;
  for(;;) {
    int j[LPR][KS];
    T v[LPR][KS];
    bool ok[LPR][KS];
    bool any = false;
#pragma unroll
    for(int u = 0; u < LPR; ++u)
#pragma unroll
      for(int s = 0; s < KS; ++s) {
        const int kk = kb[u] + k0 + s * LPR;
        ok[u][s] = kk < ke[u];
        any |= ok[u][s];
        j[u][s] = ok[u][s] ? ja[kk] : 0;
        v[u][s] = ok[u][s] ? (T)val[kk] : (T)0;
      }
    if(!any) break;
#pragma unroll
    for(int u = 0; u < LPR; ++u)
#pragma unroll
      for(int s = 0; s < KS; ++s)
        if(ok[u][s]) acc[u] += v[u][s] * f(j[u][s]);
    k0 += KS * LPR;
  }
#pragma unroll
  for(int u = 0; u < LPR; ++u) {
#pragma unroll
    for(int q = LPR / 2; q > 0; q >>= 1) acc[u] += __shfl_xor(acc[u], q, LPR);
    if(sub == u) ldsY[u * RPB + g] = acc[u];
  }
}

// y[row] = sum_k val[k] src[ja[k]] (times scale[ja[k]] when SCALE).
template <typename T, int LPR, int TB, bool SCALE>
__device__ __forceinline__ void tileProduct(int n, int base, const int *__restrict__ ia,
                                            const int *__restrict__ ja, const T *__restrict__ val,
                                            const T *__restrict__ src, const T *__restrict__ scale,
                                            const int (&kb)[LPR], const int (&ke)[LPR], T *ldsY)
{
  tileProductF<T, LPR, TB>(ja, val, kb, ke, ldsY, [&](int j) { return SCALE ? scale[j] * src[j] : src[j]; });
  (void)n;
  (void)base;
  (void)ia;
}

template <int LPR, int TB>
__device__ __forceinline__ void tileBounds(int n, int base, const int *__restrict__ ia, int (&kb)[LPR],
                                           int (&ke)[LPR])
{
  constexpr int RPB = TB / LPR;
  const int sub = threadIdx.x % LPR, g = threadIdx.x / LPR;
#pragma unroll
  for(int u = 0; u < LPR; ++u) {
    const int row = base + u * RPB + g;
    if(row < n) {
      kb[u] = ia[row] + sub;
      ke[u] = ia[row + 1];
    } else {
      kb[u] = 0;
      ke[u] = 0;
    }
  }
}


}  // namespace fe_tile
