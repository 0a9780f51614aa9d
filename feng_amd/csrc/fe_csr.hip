// Device construction of the CSR pattern (feEZCompressedRowStorage,
// feCompressedRowStorage.cpp:15-133) and of the per-element scatter-position
// maps used by the atomic-free colored assembly.
//
// Pattern algorithm (GPU-native, no global sort):
//   1. per matrix-layout group, gather the element -> row / column DOF maps
//      into color-sorted element order;
//   2. count, for every unknown row r, the (group, element) pairs whose row
//      map contains r (atomic counters -- setup only), scan, fill;
//   3. one wavefront per row gathers the candidate columns of those pairs
//      plus the diagonal into LDS, bitonic-sorts them, and counts / writes the
//      unique columns below n -- exactly "push_back, sort, unique" of the
//      reference, row by row;
//   4. scan the counts into ia, second pass writes ja;
//   5. one thread per element binary-searches every (i, j) of its local
//      matrix in the sorted row and records the value slot (or -1 for
//      essential rows/columns).
#include "fe_hip_internal.h"

#include <algorithm>
#include <climits>

namespace {

constexpr int kScanItems = 1024;  // per block: 256 threads x 4

__global__ void scanBlocks(const int *__restrict__ in, int *__restrict__ out, int n,
                           long long *__restrict__ blockSums)
{
  __shared__ long long s[256];
  const int base = blockIdx.x * kScanItems + threadIdx.x * 4;
  long long v[4];
  long long t = 0;
#pragma unroll
  for(int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    t += v[k];
  }
  s[threadIdx.x] = t;
  __syncthreads();
  for(int o = 1; o < 256; o <<= 1) {
    long long add = (threadIdx.x >= o) ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  long long run = s[threadIdx.x] - t;  // exclusive prefix of this thread
#pragma unroll
  for(int k = 0; k < 4; ++k) {
    if(base + k < n) out[base + k] = (int)run;  // block-local, fixed below
    run += v[k];
  }
  if(threadIdx.x == 255) blockSums[blockIdx.x] = s[255];
}

__global__ void scanSums(long long *__restrict__ sums, int nb, long long *__restrict__ total)
{
  __shared__ long long s[1024];
  long long carry = 0;
  for(int c0 = 0; c0 < nb; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    long long v = (i < nb) ? sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for(int o = 1; o < 1024; o <<= 1) {
      long long add = (threadIdx.x >= o) ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if(i < nb) sums[i] = carry + s[threadIdx.x] - v;
    long long blockTotal = s[1023];
    __syncthreads();
    carry += blockTotal;
  }
  if(threadIdx.x == 0) *total = carry;
}

__global__ void addOffsets(int *__restrict__ out, int n, const long long *__restrict__ sums,
                           const long long *__restrict__ total)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = (int)(out[i] + sums[i / kScanItems]);
  if(i == 0) out[n] = (int)*total;
}

// Gather the row (I) and column (J) DOF maps of one layout group into
// color-sorted element order.
struct LayoutDesc {
  const int *adr[4];
  int nloc[4];
  int count;
};

__global__ void gatherLayout(int E, const int *__restrict__ perm, LayoutDesc L, int nTot,
                             int *__restrict__ out)
{
  const int ee = blockIdx.x * blockDim.x + threadIdx.x;
  if(ee >= E) return;
  const int e = perm[ee];
  int o = 0;
  for(int f = 0; f < L.count; ++f)
    for(int k = 0; k < L.nloc[f]; ++k) out[(size_t)ee * nTot + o++] = L.adr[f][(size_t)e * L.nloc[f] + k];
}

__global__ void countRowRefs(int E, int nI, const int *__restrict__ adrI, int n,
                             int *__restrict__ cnt)
{
  const int ee = blockIdx.x * blockDim.x + threadIdx.x;
  if(ee >= E) return;
  for(int i = 0; i < nI; ++i) {
    const int r = adrI[(size_t)ee * nI + i];
    if(r < n) atomicAdd(&cnt[r], 1);
  }
}

__global__ void fillRowRefs(int E, int nI, const int *__restrict__ adrI, int n, int G, int g,
                            const int *__restrict__ refPtr, int *__restrict__ cursor,
                            int *__restrict__ refs)
{
  const int ee = blockIdx.x * blockDim.x + threadIdx.x;
  if(ee >= E) return;
  for(int i = 0; i < nI; ++i) {
    const int r = adrI[(size_t)ee * nI + i];
    if(r < n) {
      const int slot = atomicAdd(&cursor[r], 1);
      refs[refPtr[r] + slot] = ee * G + g;
    }
  }
}

struct GroupCols {
  const int *adrJ[8];
  int nJ[8];
};

__device__ __forceinline__ int waveExclusiveScan(int v, int lane, int *total)
{
  int x = v;
#pragma unroll
  for(int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if(lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// One 64-thread block (a single wavefront) per unknown row.
// Extra columns of a row (the periodic (slave, master) entries,
// feCompressedRowStorage.cpp:96-108): xPtr[n + 1] / xCol, or null.
template <int CAP>
__global__ __launch_bounds__(64) void csrRows(int n, const int *__restrict__ refPtr,
                                              const int *__restrict__ refs, int G, GroupCols gc,
                                              const int *__restrict__ xPtr, const int *__restrict__ xCol,
                                              int mode, int *__restrict__ uniq,
                                              const int *__restrict__ ia, int *__restrict__ ja,
                                              int *__restrict__ errFlag)
{
  __shared__ int buf[CAP];
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int beg = refPtr[r], end = refPtr[r + 1];
  int base = 0;
  bool overflow = false;
  for(int c0 = beg; c0 < end; c0 += 64) {
    const int idx = c0 + lane;
    int g = 0, ee = 0, cnt = 0;
    if(idx < end) {
      const int ref = refs[idx];
      g = ref % G;
      ee = ref / G;
      cnt = gc.nJ[g];
    }
    int tot;
    const int off = waveExclusiveScan(cnt, lane, &tot);
    if(base + tot + 1 > CAP) {
      overflow = true;
      break;
    }
    for(int k = 0; k < cnt; ++k) {
      const int c = gc.adrJ[g][(size_t)ee * gc.nJ[g] + k];
      buf[base + off + k] = (c < n) ? c : INT_MAX;
    }
    base += tot;
  }
  const int xb = xPtr ? xPtr[r] : 0, xn = xPtr ? xPtr[r + 1] - xb : 0;
  if(overflow || base + 1 + xn > CAP) {
    if(lane == 0) atomicMax(errFlag, 1);
    return;
  }
  if(lane == 0) buf[base] = r;  // the diagonal is always stored
  for(int k = lane; k < xn; k += 64) buf[base + 1 + k] = xCol[xb + k];
  const int N = base + 1 + xn;
  int P = 1;
  while(P < N) P <<= 1;
  for(int i = N + lane; i < P; i += 64) buf[i] = INT_MAX;
  __syncthreads();
  // Bitonic sort of buf[0, P) ascending.
  for(int k = 2; k <= P; k <<= 1) {
    for(int j = k >> 1; j > 0; j >>= 1) {
      for(int i = lane; i < P; i += 64) {
        const int ixj = i ^ j;
        if(ixj > i) {
          const int a = buf[i], b = buf[ixj];
          const bool up = (i & k) == 0;
          if((a > b) == up) {
            buf[i] = b;
            buf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // Unique values below INT_MAX.
  int written = 0;
  const int rowStart = (mode == 1) ? ia[r] : 0;
  for(int c0 = 0; c0 < P; c0 += 64) {
    const int i = c0 + lane;
    int isNew = 0, v = INT_MAX;
    if(i < P) {
      v = buf[i];
      isNew = (v != INT_MAX) && (i == 0 || buf[i - 1] != v);
    }
    int tot;
    const int off = waveExclusiveScan(isNew, lane, &tot);
    if(mode == 1 && isNew) ja[rowStart + written + off] = v;
    written += tot;
  }
  if(mode == 0 && lane == 0) uniq[r] = written;
}

__global__ void findDiag(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                         int *__restrict__ diag)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= n) return;
  int lo = ia[r], hi = ia[r + 1] - 1, p = -1;
  while(lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = ja[mid];
    if(c == r) { p = mid; break; }
    if(c < r) lo = mid + 1; else hi = mid - 1;
  }
  diag[r] = p;
}

__global__ void buildPositions(int E, int nI, int nJ, const int *__restrict__ adrI,
                               const int *__restrict__ adrJ, int n, const int *__restrict__ ia,
                               const int *__restrict__ ja, int *__restrict__ pos,
                               int *__restrict__ errFlag)
{
  const int ee = blockIdx.x * blockDim.x + threadIdx.x;
  if(ee >= E) return;
  for(int i = 0; i < nI; ++i) {
    const int r = adrI[(size_t)ee * nI + i];
    for(int j = 0; j < nJ; ++j) {
      const int c = adrJ[(size_t)ee * nJ + j];
      int p = -1;
      if(r < n && c < n) {
        int lo = ia[r], hi = ia[r + 1] - 1;
        while(lo <= hi) {
          const int mid = (lo + hi) >> 1;
          const int cc = ja[mid];
          if(cc == c) { p = mid; break; }
          if(cc < c) lo = mid + 1; else hi = mid - 1;
        }
        if(p < 0) atomicMax(errFlag, 2);
      }
      pos[(size_t)(i * nJ + j) * E + ee] = p;
    }
  }
}

}  // namespace

int feExclusiveScan(fe_hip_ctx *ctx, const int *d_in, int *d_out, int n, int64_t *total)
{
  const int nb = (n + kScanItems - 1) / kScanItems;
  long long *sums = nullptr;
  if(feAlloc(ctx, &sums, (size_t)nb + 2)) return FE_STATUS_OUT_OF_MEMORY;
  long long *d_total = sums + nb + 1;
  if(nb > 0) hipLaunchKernelGGL(scanBlocks, dim3(nb), dim3(256), 0, ctx->stream, d_in, d_out, n, sums);
  hipLaunchKernelGGL(scanSums, dim3(1), dim3(1024), 0, ctx->stream, sums, nb, d_total);
  hipLaunchKernelGGL(addOffsets, dim3((n + 256) / 256), dim3(256), 0, ctx->stream, d_out, n, sums, d_total);
  FE_HIP_TRY(ctx, hipGetLastError());
  long long h = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&h, d_total, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  feFree(ctx, sums);
  FE_REQUIRE(ctx, h < INT_MAX, "CSR size exceeds int32 indexing");
  if(total) *total = h;
  return FE_STATUS_OK;
}

extern "C" int fe_hip_build_csr(fe_hip_ctx *ctx, int64_t *nnzOut)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, !ctx->csrBuilt, "CSR already built");
  FE_REQUIRE(ctx, ctx->n > 0, "set the system sizes first");
  FE_REQUIRE(ctx, ctx->colored, "no coloring");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int E = ctx->E, n = ctx->n;
  // 1. layout groups of the matrix forms
  ctx->groups.clear();
  for(FeForm &F : ctx->forms) {
    if(!F.hasMatrix) continue;
    int g = -1;
    for(size_t k = 0; k < ctx->groups.size(); ++k)
      if(ctx->groups[k].layoutI == F.layoutI && ctx->groups[k].layoutJ == F.layoutJ) g = (int)k;
    if(g < 0) {
      FeGroup G;
      G.layoutI = F.layoutI;
      G.layoutJ = F.layoutJ;
      G.nI = F.nI;
      G.nJ = F.nJ;
      ctx->groups.push_back(G);
      g = (int)ctx->groups.size() - 1;
    }
    F.group = g;
  }
  const int G = (int)ctx->groups.size();
  FE_REQUIRE(ctx, G <= 8, "too many matrix layouts");
  FE_REQUIRE(ctx, (int64_t)E * std::max(G, 1) < INT_MAX, "too many elements");
  const int TB = 256, nbE = (E + TB - 1) / TB;
  for(FeGroup &grp : ctx->groups) {
    if(feAlloc(ctx, &grp.d_adrI, (size_t)E * grp.nI) || feAlloc(ctx, &grp.d_adrJ, (size_t)E * grp.nJ))
      return FE_STATUS_OUT_OF_MEMORY;
    for(int side = 0; side < 2; ++side) {
      const std::vector<int> &lay = side == 0 ? grp.layoutI : grp.layoutJ;
      LayoutDesc L{};
      L.count = (int)lay.size();
      FE_REQUIRE(ctx, L.count <= 4, "layout too long");
      for(int f = 0; f < L.count; ++f) {
        L.adr[f] = ctx->fields[lay[f]].adr;
        L.nloc[f] = ctx->fields[lay[f]].nloc;
      }
      hipLaunchKernelGGL(gatherLayout, dim3(nbE), dim3(TB), 0, ctx->stream, E, ctx->d_perm, L,
                         side == 0 ? grp.nI : grp.nJ, side == 0 ? grp.d_adrI : grp.d_adrJ);
    }
  }
  FE_HIP_TRY(ctx, hipGetLastError());
  // 2. row -> (group, element) references
  int *cnt = nullptr, *refPtr = nullptr, *cursor = nullptr, *refs = nullptr, *err = nullptr;
  if(feAlloc(ctx, &cnt, n) || feAlloc(ctx, &refPtr, (size_t)n + 1) || feAlloc(ctx, &cursor, n) ||
     feAlloc(ctx, &err, 1))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * n, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(cursor, 0, sizeof(int) * n, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
  for(FeGroup &grp : ctx->groups)
    hipLaunchKernelGGL(countRowRefs, dim3(nbE), dim3(TB), 0, ctx->stream, E, grp.nI, grp.d_adrI, n, cnt);
  int64_t nrefs = 0;
  if(int rc = feExclusiveScan(ctx, cnt, refPtr, n, &nrefs)) return rc;
  if(feAlloc(ctx, &refs, (size_t)std::max<int64_t>(nrefs, 1))) return FE_STATUS_OUT_OF_MEMORY;
  for(int g = 0; g < G; ++g)
    hipLaunchKernelGGL(fillRowRefs, dim3(nbE), dim3(TB), 0, ctx->stream, E, ctx->groups[g].nI,
                       ctx->groups[g].d_adrI, n, G, g, refPtr, cursor, refs);
  // 3-4. per-row sort / unique
  GroupCols gc{};
  for(int g = 0; g < G; ++g) {
    gc.adrJ[g] = ctx->groups[g].d_adrJ;
    gc.nJ[g] = ctx->groups[g].nJ;
  }
  if(feAlloc(ctx, &ctx->d_ia, (size_t)n + 1)) return FE_STATUS_OUT_OF_MEMORY;
  constexpr int CAP = 2048;
  // periodic (slave, master) entries as extra columns of the slave rows
  int *xPtr = nullptr, *xCol = nullptr;
  if(!ctx->perSlave.empty()) {
    std::vector<int> hp(n + 1, 0);
    for(size_t k = 0; k < ctx->perSlave.size(); ++k) hp[ctx->perSlave[k] + 1]++;
    for(int i = 0; i < n; ++i) hp[i + 1] += hp[i];
    std::vector<int> hc(ctx->perSlave.size()), fill(hp.begin(), hp.end() - 1);
    for(size_t k = 0; k < ctx->perSlave.size(); ++k) hc[fill[ctx->perSlave[k]]++] = ctx->perMaster[k];
    if(feAlloc(ctx, &xPtr, (size_t)n + 1) || feAlloc(ctx, &xCol, hc.size())) return FE_STATUS_OUT_OF_MEMORY;
    FE_HIP_TRY(ctx, hipMemcpy(xPtr, hp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    FE_HIP_TRY(ctx, hipMemcpy(xCol, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(csrRows<CAP>, dim3(n), dim3(64), 0, ctx->stream, n, refPtr, refs, std::max(G, 1),
                     gc, (const int *)xPtr, (const int *)xCol, 0, cnt, (const int *)nullptr, (int *)nullptr, err);
  FE_HIP_TRY(ctx, hipGetLastError());
  int hErr = 0;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&hErr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, hErr == 0, "a CSR row has more than 2048 candidate columns");
  int64_t nnz = 0;
  if(int rc = feExclusiveScan(ctx, cnt, ctx->d_ia, n, &nnz)) return rc;
  ctx->nnz = nnz;
  if(feAlloc(ctx, &ctx->d_ja, (size_t)std::max<int64_t>(nnz, 1))) return FE_STATUS_OUT_OF_MEMORY;
  hipLaunchKernelGGL(csrRows<CAP>, dim3(n), dim3(64), 0, ctx->stream, n, refPtr, refs, std::max(G, 1),
                     gc, (const int *)xPtr, (const int *)xCol, 1, (int *)nullptr, ctx->d_ia, ctx->d_ja, err);
  FE_HIP_TRY(ctx, hipGetLastError());
  // diagonal positions and scatter maps
  if(feAlloc(ctx, &ctx->d_diag, n)) return FE_STATUS_OUT_OF_MEMORY;
  hipLaunchKernelGGL(findDiag, dim3((n + TB - 1) / TB), dim3(TB), 0, ctx->stream, n, ctx->d_ia, ctx->d_ja, ctx->d_diag);
  for(FeGroup &grp : ctx->groups) {
    if(feAlloc(ctx, &grp.d_pos, (size_t)E * grp.nI * grp.nJ)) return FE_STATUS_OUT_OF_MEMORY;
    hipLaunchKernelGGL(buildPositions, dim3(nbE), dim3(TB), 0, ctx->stream, E, grp.nI, grp.nJ,
                       grp.d_adrI, grp.d_adrJ, n, ctx->d_ia, ctx->d_ja, grp.d_pos, err);
  }
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipMemcpyAsync(&hErr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FE_REQUIRE(ctx, hErr == 0, "scatter position not found in the CSR pattern");
  feFree(ctx, cnt);
  feFree(ctx, refPtr);
  feFree(ctx, cursor);
  feFree(ctx, refs);
  feFree(ctx, err);
  if(xPtr) {
    feFree(ctx, xPtr);
    feFree(ctx, xCol);
  }
  if(feAlloc(ctx, &ctx->d_val, (size_t)std::max<int64_t>(nnz, 1))) return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_val, 0, sizeof(double) * nnz, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  // lanes per row of the vector SpMV: a quarter of the average row length
  // (at least 4), measured best on gfx950 for the P1/P2 stencils
  const double avg = (double)nnz / n;
  int lanes = 4;
  while(lanes < 32 && 4 * lanes < avg) lanes <<= 1;
  if(!ctx->lanesFixed) ctx->spmvLanes = lanes;
  ctx->csrBuilt = true;
  if(nnzOut) *nnzOut = nnz;
  return FE_STATUS_OK;
}

extern "C" int fe_hip_get_csr(fe_hip_ctx *ctx, int *ia, int *ja)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(ia) FE_HIP_TRY(ctx, hipMemcpyAsync(ia, ctx->d_ia, sizeof(int) * ((size_t)ctx->n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if(ja) FE_HIP_TRY(ctx, hipMemcpyAsync(ja, ctx->d_ja, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}
