// Row-gather assembly of the fused scalar forms (DIFFUSION, TRANSIENT_MASS,
// SOURCE): one launch for the whole matrix and right-hand side instead of one
// launch per color, with the summation order of the colored scatter.
//
// The reference scatters element matrices color by color
// (feLinearSystemMklPardiso.cpp:501-749, element loop per color of
// feCncGeo::colorElements, feCncGeo.cpp:735-794): within a color no two
// elements share a vertex, so each CSR value receives at most one element
// contribution per color, and its final value is the sum of its element
// contributions in color order.  Here the owner of CSR row i (one lane)
// walks the elements incident to DOF i in that same order (color-sorted
// element order) and, for each, computes row a (the local index of DOF i)
// of the element matrix and the entry a of the element residual, adding
// them into registers; the row is then added into the CSR once.  No two
// lanes write the same value: no atomics, no colors as launches, each value
// and RHS entry read-modified-written once, bitwise repeatable.
//
// Incidences: for the rows of a 64-row slice (one wavefront), the packed
// (element << 4 | local index) of every element incident to each row, in
// color-sorted element order, stored slot-major (slot k of the slice's 64
// lanes contiguous) so a wavefront reads them coalesced; -1 pads rows with
// fewer incidences than the slice's longest.  Each incidence carries, in a
// second word, the position within the row of every DOF of the element (its
// CSR slot, SB bits each; all ones: a DOF that is not an unknown).  The row
// is accumulated in LDS, one column of MAXR slots per lane.  Built once per
// CSR pattern on the host.
#include "fe_element.h"

#include <algorithm>
#include <numeric>
#include <cmath>
#include <vector>

namespace {

constexpr int kGatherSlice = 64;
constexpr int kGatherTB = 256;

struct GatherArgs {
  int n, nSlices, what, nq;
  const int *ia, *sliceOff, *conn, *adr;  // adr: E x NF, mesh element order
  const int2 *inc;                        // (element << 4 | local index, packed CSR slots)
  const int *order;                       // slice of each (block, wavefront) slot
  const double *xyz, *tab, *u;
  FeFn coeff, src, mass;
  int hasDiff, hasSrc, hasMass;
  int nsol;
  double cn[fe_hip_ctx::kMaxPast + 1];
  const double *past[fe_hip_ctx::kMaxPast];
  double *val, *rhs;
  float *val32;  // non-null: also write the FP32 rounding of every updated value
  int fresh;     // the values are zero and this sweep is their only writer: stored, not read and added
  int xcd;
};

// Slot of this block: with xcd set and a grid that is a multiple of 8, each
// of the 8 XCDs (block b runs on XCD b % 8) takes one contiguous range of
// slots.  Slots map to slices in Morton order of the slices' element
// centroids, so the blocks an XCD runs together cover a compact region of the
// mesh and the elements their rows share stay in that XCD's L2.
__device__ __forceinline__ int gatherBlock(int xcd)
{
  const int G = gridDim.x;
  if(!xcd || (G & 7)) return blockIdx.x;
  return (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
}

// N consecutive ints of one element in the widest aligned loads: 16-byte
// loads for 4 (element rows of 16 bytes), 8-byte loads for even N (rows of
// 8k bytes), single ints otherwise.
template <int N>
__device__ __forceinline__ void gatherInts(const int *__restrict__ p, int (&v)[N])
{
  if constexpr(N == 4) {
    const int4 w = *(const int4 *)p;
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
    v[3] = w.w;
  } else if constexpr(N % 2 == 0) {
#pragma unroll
    for(int i = 0; i < N / 2; ++i) {
      const int2 w = ((const int2 *)p)[i];
      v[2 * i] = w.x;
      v[2 * i + 1] = w.y;
    }
  } else {
#pragma unroll
    for(int j = 0; j < N; ++j) v[j] = p[j];
  }
}

// One lane per CSR row.  TC: the arithmetic type of the element rows and of
// their accumulation (double, or float for FP32 element matrices whose sums
// are added to the FP64 values widened).
template <typename TC, int NF, int DIM, int MAXR>
__global__ __launch_bounds__(kGatherTB) void asmGather(GatherArgs a)
{
  constexpr int SB = MAXR <= 16 ? 4 : 5;  // bits per packed slot
  constexpr int SMASK = (1 << SB) - 1;
  static_assert(NF * SB <= 32, "packed slots of one incidence");
  __shared__ FeRefTablesFor<NF> T;
  __shared__ TC acc[MAXR * kGatherTB];  // slot s of this lane's row at acc[s * kGatherTB + threadIdx.x]
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(T) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  constexpr int NV = DIM + 1;
  constexpr int WPB = kGatherTB / kGatherSlice;
  const int slot = gatherBlock(a.xcd) * WPB + (int)(threadIdx.x / kGatherSlice);
  const int lane = threadIdx.x % kGatherSlice;
  if(slot >= a.nSlices) return;
  const int slice = a.order[slot];
  const int row = slice * kGatherSlice + lane;
  int k0 = 0, len = 0;
  if(row < a.n) {
    k0 = a.ia[row];
    len = a.ia[row + 1] - k0;
  }
#pragma unroll
  for(int s = 0; s < MAXR; ++s) acc[s * kGatherTB + threadIdx.x] = (TC)0;
  double racc = 0.;
  const bool doMat = (a.what & 1) && (a.hasDiff || a.hasMass), doRes = (a.what & 2);
  const int off = a.sliceOff[slice], W = (a.sliceOff[slice + 1] - off) / kGatherSlice;
  const int nq = a.nq;
  for(int k = 0; k < W; ++k) {
    // the incidence words stream through once: non-temporal, so they do not
    // displace the element data the neighbouring rows reuse from L2
    const unsigned long long pk =
        __builtin_nontemporal_load((const unsigned long long *)&a.inc[off + k * kGatherSlice + lane]);
    const int2 pw = make_int2((int)(unsigned)pk, (int)(unsigned)(pk >> 32));
    const int p = pw.x;
    if(p < 0) continue;
    const int e = p >> 4, ra = p & 15;
    int cv[NV], ad[NF];
    gatherInts<NV>(a.conn + (size_t)e * NV, cv);
    gatherInts<NF>(a.adr + (size_t)e * NF, ad);
    double c[NV][3];
#pragma unroll
    for(int j = 0; j < NV; ++j)
#pragma unroll
      for(int d = 0; d < 3; ++d) c[j][d] = a.xyz[3 * (size_t)cv[j] + d];
    double J, drdx[3], drdy[3], drdz[3];
    feGeometry<DIM>(c, J, drdx, drdy, drdz);
    double ul[NF];
#pragma unroll
    for(int b = 0; b < NF; ++b) ul[b] = doRes ? a.u[ad[b]] : 0.;
    TC rowv[NF];
#pragma unroll
    for(int b = 0; b < NF; ++b) rowv[b] = (TC)0;
    double be = 0.;
    if(a.hasDiff) {
      for(int q = 0; q < nq; ++q) {
        double x[3] = {0., 0., 0.};
#pragma unroll
        for(int j = 0; j < NV; ++j)
#pragma unroll
          for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[q][j];
        const double kc = feEvalFn(a.coeff, x, e, nq, q);
        double gp[NF][DIM], ga[DIM];
#pragma unroll
        for(int b = 0; b < NF; ++b) {
          if(DIM == 2) {
            gp[b][0] = T.dLr[q][b] * drdx[0] + T.dLs[q][b] * drdx[1];
            gp[b][1 % DIM] = T.dLr[q][b] * drdy[0] + T.dLs[q][b] * drdy[1];
          } else {
            gp[b][0] = T.dLr[q][b] * drdx[0] + T.dLs[q][b] * drdx[1] + T.dLt[q][b] * drdx[2];
            gp[b][1 % DIM] = T.dLr[q][b] * drdy[0] + T.dLs[q][b] * drdy[1] + T.dLt[q][b] * drdy[2];
            gp[b][2 % DIM] = T.dLr[q][b] * drdz[0] + T.dLs[q][b] * drdz[1] + T.dLt[q][b] * drdz[2];
          }
        }
        if(DIM == 2) {
          ga[0] = T.dLr[q][ra] * drdx[0] + T.dLs[q][ra] * drdx[1];
          ga[1 % DIM] = T.dLr[q][ra] * drdy[0] + T.dLs[q][ra] * drdy[1];
        } else {
          ga[0] = T.dLr[q][ra] * drdx[0] + T.dLs[q][ra] * drdx[1] + T.dLt[q][ra] * drdx[2];
          ga[1 % DIM] = T.dLr[q][ra] * drdy[0] + T.dLs[q][ra] * drdy[1] + T.dLt[q][ra] * drdy[2];
          ga[2 % DIM] = T.dLr[q][ra] * drdz[0] + T.dLs[q][ra] * drdz[1] + T.dLt[q][ra] * drdz[2];
        }
        const double jw = J * T.w[q];
        if(doMat) {
          const TC kt = (TC)kc, jt = (TC)jw;
#pragma unroll
          for(int b = 0; b < NF; ++b)
#pragma unroll
            for(int d = 0; d < DIM; ++d) rowv[b] += kt * (TC)ga[d] * (TC)gp[b][d] * jt;
        }
        if(doRes) {
          double gu[DIM];
#pragma unroll
          for(int d = 0; d < DIM; ++d) gu[d] = 0.;
#pragma unroll
          for(int b = 0; b < NF; ++b)
#pragma unroll
            for(int d = 0; d < DIM; ++d) gu[d] += ul[b] * gp[b][d];
#pragma unroll
          for(int d = 0; d < DIM; ++d) be -= kc * ga[d] * gu[d] * jw;
        }
      }
    }
    if(a.hasMass) {
      double ud[NF];
      if(doRes) {
#pragma unroll
        for(int b = 0; b < NF; ++b) {
          double v = a.cn[0] * ul[b];
          for(int j = 1; j < a.nsol; ++j) v += a.cn[j] * a.past[j - 1][ad[b]];
          ud[b] = v;
        }
      }
      const double c0 = a.cn[0];
      for(int q = 0; q < nq; ++q) {
        double x[3] = {0., 0., 0.};
#pragma unroll
        for(int j = 0; j < NV; ++j)
#pragma unroll
          for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[q][j];
        const double m = feEvalFn(a.mass, x, e, nq, q);
        const double jw = J * T.w[q];
        const double La = T.L[q][ra];
        if(doMat) {
#pragma unroll
          for(int b = 0; b < NF; ++b) rowv[b] += (TC)(La * m * c0 * T.L[q][b] * jw);
        }
        if(doRes) {
          double uq = 0.;
#pragma unroll
          for(int b = 0; b < NF; ++b) uq += T.L[q][b] * ud[b];
          be -= La * m * uq * jw;
        }
      }
    }
    if(doRes && a.hasSrc) {
      for(int q = 0; q < nq; ++q) {
        double x[3] = {0., 0., 0.};
#pragma unroll
        for(int j = 0; j < NV; ++j)
#pragma unroll
          for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[q][j];
        const double S = feEvalFn(a.src, x, e, nq, q);
        be -= T.L[q][ra] * S * (J * T.w[q]);
      }
    }
    racc += be;
    if(doMat) {
      const unsigned sl = (unsigned)pw.y;
#pragma unroll
      for(int b = 0; b < NF; ++b) {
        const int s = (sl >> (SB * b)) & SMASK;
        if(s != SMASK) acc[s * kGatherTB + threadIdx.x] += rowv[b];
      }
    }
  }
  if(row >= a.n) return;
  if(doRes) a.rhs[row] += racc;
  if(doMat) {
    for(int s = 0; s < len; ++s) {
      const double v = (a.fresh ? 0. : a.val[k0 + s]) + (double)acc[s * kGatherTB + threadIdx.x];
      a.val[k0 + s] = v;
      if(a.val32) a.val32[k0 + s] = (float)v;
    }
  }
}

template <typename TC, int NF, int DIM, int MAXR>
void gatherLaunch(fe_hip_ctx *ctx, const GatherArgs &a)
{
  constexpr int WPB = kGatherTB / kGatherSlice;
  int G = (a.nSlices + WPB - 1) / WPB;
  if(a.xcd) G = (G + 7) & ~7;
  hipLaunchKernelGGL((asmGather<TC, NF, DIM, MAXR>), dim3(G), dim3(kGatherTB), 0, ctx->stream, a);
}

template <typename TC, int NF, int DIM>
void gatherLaunchR(fe_hip_ctx *ctx, const GatherArgs &a, int maxR)
{
  // row capacity: a slot index below all ones (the "not an unknown" mark)
  if(maxR < 8) gatherLaunch<TC, NF, DIM, 8>(ctx, a);
  else if(maxR < 16) gatherLaunch<TC, NF, DIM, 16>(ctx, a);
  else if(maxR <= 24) gatherLaunch<TC, NF, DIM, 24>(ctx, a);
  else gatherLaunch<TC, NF, DIM, 31>(ctx, a);
}

}  // namespace

int feGatherPrepare(fe_hip_ctx *ctx, int field, bool *ok)
{
  *ok = false;
  FeGatherField &g = ctx->gather[field];
  if(g.built) {
    *ok = g.usable;
    return FE_STATUS_OK;
  }
  g.built = true;
  const FeField &F = ctx->fields[field];
  const int nf = F.nloc, E = ctx->E, n = ctx->n;
  if(!(F.kind == FE_TRI_P1 || F.kind == FE_TRI_P2 || F.kind == FE_TET_P1) || E >= (1 << 27)) return FE_STATUS_OK;
  std::vector<int> ia(n + 1), adr((size_t)E * nf);
  FE_HIP_TRY(ctx, hipMemcpy(ia.data(), ctx->d_ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  std::vector<int> ja(ia[n]);
  FE_HIP_TRY(ctx, hipMemcpy(ja.data(), ctx->d_ja, sizeof(int) * ja.size(), hipMemcpyDeviceToHost));
  FE_HIP_TRY(ctx, hipMemcpy(adr.data(), F.adr, sizeof(int) * adr.size(), hipMemcpyDeviceToHost));
  int maxR = 0;
  for(int i = 0; i < n; ++i) maxR = std::max(maxR, ia[i + 1] - ia[i]);
  if(maxR > kGatherMaxRow) return FE_STATUS_OK;  // the colored sweep keeps such patterns
  const int sb = maxR < 16 ? 4 : 5, smask = (1 << sb) - 1;
  // incidences per row in color-sorted element order
  std::vector<int> cnt(n, 0);
  for(int ee = 0; ee < E; ++ee) {
    const int e = ctx->h_perm[ee];
    for(int b = 0; b < nf; ++b) {
      const int d = adr[(size_t)e * nf + b];
      if(d < n) cnt[d]++;
    }
  }
  const int nSlices = (n + kGatherSlice - 1) / kGatherSlice;
  std::vector<int> off(nSlices + 1, 0);
  for(int s = 0; s < nSlices; ++s) {
    int w = 0;
    for(int i = s * kGatherSlice; i < std::min(n, (s + 1) * kGatherSlice); ++i) w = std::max(w, cnt[i]);
    const int64_t next = (int64_t)off[s] + (int64_t)w * kGatherSlice;
    FE_REQUIRE(ctx, next < ((int64_t)1 << 31), "gather incidences overflow int32");
    off[s + 1] = (int)next;
  }
  std::vector<int2> inc(std::max(1, off[nSlices]), make_int2(-1, 0));
  std::vector<int> fill(n, 0);
  for(int ee = 0; ee < E; ++ee) {
    const int e = ctx->h_perm[ee];
    for(int b = 0; b < nf; ++b) {
      const int d = adr[(size_t)e * nf + b];
      if(d >= n) continue;
      // CSR slot of every DOF of the element in row d
      unsigned slots = 0;
      const int *rb = ja.data() + ia[d], *re = ja.data() + ia[d + 1];
      for(int c = 0; c < nf; ++c) {
        const int col = adr[(size_t)e * nf + c];
        int s = smask;
        if(col < n) {
          const int *it = std::lower_bound(rb, re, col);
          FE_REQUIRE(ctx, it != re && *it == col, "gather assembly: an element column is missing from the pattern");
          s = (int)(it - rb);
        }
        slots |= (unsigned)s << (sb * c);
      }
      inc[(size_t)off[d / kGatherSlice] + (size_t)fill[d]++ * kGatherSlice + d % kGatherSlice] =
          make_int2(e << 4 | b, (int)slots);
    }
  }
  // slice order: Morton code of the mean centroid of the elements incident to
  // the slice's rows, so the slices that run together on one XCD share
  // elements and their rows' elements stay in that XCD's L2
  std::vector<int> order(nSlices);
  {
    const int nv = ctx->nvpe;
    std::vector<int> conn((size_t)E * nv);
    std::vector<double> xyz((size_t)ctx->V * 3);
    FE_HIP_TRY(ctx, hipMemcpy(conn.data(), ctx->d_conn, sizeof(int) * conn.size(), hipMemcpyDeviceToHost));
    FE_HIP_TRY(ctx, hipMemcpy(xyz.data(), ctx->d_xyz, sizeof(double) * xyz.size(), hipMemcpyDeviceToHost));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for(int v = 0; v < ctx->V; ++v)
      for(int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], xyz[3 * (size_t)v + d]);
        hi[d] = std::max(hi[d], xyz[3 * (size_t)v + d]);
      }
    std::vector<double> cen((size_t)nSlices * 3, 0.);
    std::vector<int> num(nSlices, 0);
    for(int e = 0; e < E; ++e) {
      double ce[3] = {0., 0., 0.};
      for(int j = 0; j < nv; ++j)
        for(int d = 0; d < 3; ++d) ce[d] += xyz[3 * (size_t)conn[(size_t)e * nv + j] + d] / nv;
      for(int b = 0; b < nf; ++b) {
        const int d = adr[(size_t)e * nf + b];
        if(d >= n) continue;
        const int sl = d / kGatherSlice;
        for(int x = 0; x < 3; ++x) cen[3 * (size_t)sl + x] += ce[x];
        num[sl]++;
      }
    }
    std::vector<uint64_t> key(nSlices);
    for(int sl = 0; sl < nSlices; ++sl) {
      uint64_t code = 0;
      unsigned q[3];
      for(int d = 0; d < 3; ++d) {
        const double ext = hi[d] - lo[d];
        const double t = num[sl] && ext > 0. ? (cen[3 * (size_t)sl + d] / num[sl] - lo[d]) / ext : 0.;
        q[d] = (unsigned)std::min(1023., std::max(0., t * 1024.));
      }
      for(int bit = 9; bit >= 0; --bit)
        for(int d = 0; d < 3; ++d) code = code << 1 | ((q[d] >> bit) & 1u);
      key[sl] = code;
    }
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
  }
  if(feAlloc(ctx, &g.d_sliceOff, off.size()) || feAlloc(ctx, &g.d_inc, inc.size()) ||
     feAlloc(ctx, &g.d_order, order.size()))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(g.d_sliceOff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(g.d_inc, inc.data(), sizeof(int2) * inc.size(), hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(g.d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  g.nSlices = nSlices;
  g.maxR = maxR;
  g.incBytes = (int64_t)inc.size() * 8;
  g.usable = true;
  *ok = true;
  return FE_STATUS_OK;
}

int feGatherLaunch(fe_hip_ctx *ctx, const FeGatherSweep &s)
{
  const FeGatherField &g = ctx->gather.at(s.field);
  const FeField &F = ctx->fields[s.field];
  GatherArgs a{};
  a.n = ctx->n;
  a.nSlices = g.nSlices;
  a.what = s.what;
  a.nq = s.nq;
  a.ia = ctx->d_ia;
  a.sliceOff = g.d_sliceOff;
  a.inc = g.d_inc;
  a.order = g.d_order;
  a.conn = ctx->d_conn;
  a.adr = F.adr;
  a.xyz = ctx->d_xyz;
  a.tab = s.tab;
  a.u = ctx->d_u;
  a.hasDiff = s.hasDiff;
  a.hasSrc = s.hasSrc;
  a.hasMass = s.hasMass;
  a.coeff = s.coeff;
  a.src = s.src;
  a.mass = s.mass;
  a.nsol = ctx->nsol;
  for(int j = 0; j <= fe_hip_ctx::kMaxPast; ++j) a.cn[j] = ctx->bdf[j];
  for(int j = 0; j < fe_hip_ctx::kMaxPast; ++j) a.past[j] = ctx->d_past[j];
  a.val = ctx->d_val;
  a.rhs = ctx->d_rhs;
  a.val32 = s.val32;
  a.fresh = s.freshVal ? 1 : 0;
  a.xcd = ctx->gatherXcd;
  const bool f32 = s.fp32Math;
  if(F.kind == FE_TRI_P1)
    f32 ? gatherLaunchR<float, 3, 2>(ctx, a, g.maxR) : gatherLaunchR<double, 3, 2>(ctx, a, g.maxR);
  else if(F.kind == FE_TRI_P2)
    f32 ? gatherLaunchR<float, 6, 2>(ctx, a, g.maxR) : gatherLaunchR<double, 6, 2>(ctx, a, g.maxR);
  else if(F.kind == FE_TET_P1)
    f32 ? gatherLaunchR<float, 4, 3>(ctx, a, g.maxR) : gatherLaunchR<double, 4, 3>(ctx, a, g.maxR);
  else FE_REQUIRE(ctx, false, "gather assembly: unsupported field kind");
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}
