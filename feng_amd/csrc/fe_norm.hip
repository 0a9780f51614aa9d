// Norms and error norms of one field on the device (feNorm::compute,
// feNorm.cpp:124-246, with computeLpNorm :323-398, computeVectorLpNorm
// :1399-1444, computeVectorLInfNorm :1445-1484, the time-derivative
// L-infinity norms :1485-1545 / :1618-1642, computeLInfNorm :1593-1617,
// computeH1SemiNorm :1643-1733, computeVectorH1SemiNorm :1794-1847,
// computeH1Norm :1848-1853, computeArea :1862-1873, the integrals
// computeIntegral :1875-1891, computeIntegralSolutionDependentFunction
// :1893-1917, computeIntegralDt :1964-1977, computeIntegralUserFunction
// :1979-1994).
//
// One thread per element (mesh order, no coloring: nothing is scattered):
// the element's contribution over its quadrature points, a block reduction
// into one partial per block, and one wavefront folding the partials in a
// fixed order, so the value is bitwise repeatable.  Exact solutions are
// coefficient callbacks (analytic kinds or quadrature-point tables) as in
// the assembly; the gradient of the exact solution is passed as its
// components, like the reference's vector callback of the H1 norms.
#include "fe_element.h"

#include <cmath>
#include <mutex>
#include <string>

namespace {

struct NormArgs {
  int E, type, nq, nsol;
  const int *conn, *adr;  // adr: E x NF, mesh element order
  const double *xyz, *tab, *u;
  FeFn ex[4];
  double cn[fe_hip_ctx::kMaxPast + 1];
  const double *past[fe_hip_ctx::kMaxPast];
  double *partials;
};

__device__ __forceinline__ bool normIsMax(int t)
{
  return t == FE_NORM_LINF || t == FE_NORM_LINF_ERROR || t == FE_NORM_LINF_UDOT_ERROR ||
         t == FE_NORM_VECTOR_LINF_ERROR || t == FE_NORM_VECTOR_LINF_UDOT_ERROR;
}

// NC = 1: scalar field with NF functions; NC = 2: the vector P2 field,
// function f = 2 a + c is psi_a e_c (NF = 12, six scalar functions).
template <int NF, int DIM, int NC>
__global__ __launch_bounds__(256) void normElements(NormArgs a)
{
  __shared__ FeRefTablesFor<NF / NC> T;
  __shared__ double red[4];
  {
    const double *src = a.tab;
    double *dst = (double *)&T;
    for(int i = threadIdx.x; i < (int)(sizeof(T) / sizeof(double)); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  constexpr int NV = DIM + 1, NS = NF / NC;
  const int t = a.type;
  const bool isMax = normIsMax(t);
  const bool udot = t == FE_NORM_LINF_UDOT_ERROR || t == FE_NORM_VECTOR_LINF_UDOT_ERROR || t == FE_NORM_INTEGRAL_DT;
  const bool integ = t == FE_NORM_INTEGRAL || t == FE_NORM_INTEGRAL_DT || t == FE_NORM_INTEGRAL_SOL_DEPENDENT_F ||
                     t == FE_NORM_INTEGRAL_F;
  const bool err = t == FE_NORM_L1_ERROR || t == FE_NORM_L2_ERROR || t == FE_NORM_LINF_ERROR || udot ||
                   t == FE_NORM_VECTOR_L2_ERROR || t == FE_NORM_VECTOR_LINF_ERROR || t == FE_NORM_H1_ERROR;
  const bool grad = t == FE_NORM_SEMI_H1 || t == FE_NORM_SEMI_H1_ERROR || t == FE_NORM_H1 || t == FE_NORM_H1_ERROR ||
                    t == FE_NORM_VECTOR_SEMI_H1_ERROR;
  const bool val = t != FE_NORM_SEMI_H1 && t != FE_NORM_SEMI_H1_ERROR && t != FE_NORM_AREA &&
                   t != FE_NORM_VECTOR_SEMI_H1_ERROR && !integ;
  // exact callbacks: value(s) first, then the gradient components
  const int g0 = (t == FE_NORM_H1_ERROR) ? 1 : 0;
  const bool gradErr = t == FE_NORM_SEMI_H1_ERROR || t == FE_NORM_H1_ERROR || t == FE_NORM_VECTOR_SEMI_H1_ERROR;
  double acc = 0.;
  for(int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.E; e += gridDim.x * blockDim.x) {
    double c[NV][3];
#pragma unroll
    for(int j = 0; j < NV; ++j) {
      const int v = a.conn[(size_t)e * NV + j];
#pragma unroll
      for(int d = 0; d < 3; ++d) c[j][d] = a.xyz[3 * (size_t)v + d];
    }
    double J, drdx[3], drdy[3], drdz[3];
    feGeometry<DIM>(c, J, drdx, drdy, drdz);
    double ul[NF];
#pragma unroll
    for(int i = 0; i < NF; ++i) {
      const int ad = a.adr[(size_t)e * NF + i];
      double v = a.u[ad];
      if(udot) {
        v *= a.cn[0];
        for(int j = 1; j < a.nsol; ++j) v += a.cn[j] * a.past[j - 1][ad];
      }
      ul[i] = v;
    }
    for(int k = 0; k < T.nq; ++k) {
      double x[3] = {0., 0., 0.};
#pragma unroll
      for(int j = 0; j < NV; ++j)
#pragma unroll
        for(int d = 0; d < 3; ++d) x[d] += c[j][d] * T.Lg[k][j];
      const double jw = J * T.w[k];
      double q = 0.;  // this point's integrand (sum norms) or value (max norms)
      if(val) {
#pragma unroll
        for(int cc = 0; cc < NC; ++cc) {
          double uh = 0.;
#pragma unroll
          for(int b = 0; b < NS; ++b) uh += T.L[k][b] * ul[NC * b + cc];
          const double d = (err ? feEvalFn(a.ex[cc], x, e, T.nq, k) : 0.) - uh;
          if(isMax) q = fmax(q, fabs(d));
          else if(t == FE_NORM_L1 || t == FE_NORM_L1_ERROR) q += fabs(d);
          else q += d * d;
        }
      }
      if(grad && NC == 2) {
        // grad(u)[m 2 + n] = d u_n / d x_m, exact components in that order
#pragma unroll
        for(int m = 0; m < 2; ++m)
#pragma unroll
          for(int cc = 0; cc < 2; ++cc) {
            double gh = 0.;
#pragma unroll
            for(int b = 0; b < NS; ++b) {
              const double gphys = m == 0 ? T.dLr[k][b] * drdx[0] + T.dLs[k][b] * drdx[1]
                                          : T.dLr[k][b] * drdy[0] + T.dLs[k][b] * drdy[1];
              gh += ul[NC * b + cc] * gphys;
            }
            const double gd = (gradErr ? feEvalFn(a.ex[m * 2 + cc], x, e, T.nq, k) : 0.) - gh;
            q += gd * gd;
          }
      } else if(grad) {
        double gh[3] = {0., 0., 0.};
#pragma unroll
        for(int b = 0; b < NS; ++b) {
          const double gr = T.dLr[k][b], gs = T.dLs[k][b], gt = DIM == 3 ? T.dLt[k][b] : 0.;
          gh[0] += ul[b] * (gr * drdx[0] + gs * drdx[1] + gt * drdx[2]);
          gh[1] += ul[b] * (gr * drdy[0] + gs * drdy[1] + gt * drdy[2]);
          if(DIM == 3) gh[2] += ul[b] * (gr * drdz[0] + gs * drdz[1] + gt * drdz[2]);
        }
#pragma unroll
        for(int d = 0; d < DIM; ++d) {
          const double gd = (gradErr ? feEvalFn(a.ex[g0 + d], x, e, T.nq, k) : 0.) - gh[d];
          q += gd * gd;
        }
      }
      if(t == FE_NORM_AREA) q = 1.;
      if(integ) {
        // signed integrands of the scalar field (u_h, du_h/dt, f(x, u_h)) or
        // of the callback alone
        double uh = 0.;
#pragma unroll
        for(int b = 0; b < NS; ++b) uh += T.L[k][b] * ul[NC * b];
        if(t == FE_NORM_INTEGRAL_F) q = feEvalFn(a.ex[0], x, e, T.nq, k);
        else if(t == FE_NORM_INTEGRAL_SOL_DEPENDENT_F) q = feEvalFnU(a.ex[0], x, e, T.nq, k, uh);
        else q = uh;
      }
      if(isMax) acc = fmax(acc, q);
      else acc += q * jw;
    }
  }
  // block reduction: wavefront shuffles, then the 4 wavefronts in order
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(acc, o, 64);
    acc = isMax ? fmax(acc, v) : acc + v;
  }
  if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if(threadIdx.x == 0) {
    double r = red[0];
    for(int w = 1; w < (int)(blockDim.x >> 6); ++w) r = isMax ? fmax(r, red[w]) : r + red[w];
    a.partials[blockIdx.x] = r;
  }
}

// One wavefront folds the G block partials in a fixed order.
__global__ __launch_bounds__(64) void normFinal(const double *__restrict__ partials, int G, int isMax,
                                                double *__restrict__ out)
{
  double r = 0.;
  for(int b = threadIdx.x; b < G; b += 64) r = isMax ? fmax(r, partials[b]) : r + partials[b];
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(r, o, 64);
    r = isMax ? fmax(r, v) : r + v;
  }
  if(threadIdx.x == 0) *out = r;
}

// Number of exact callbacks each norm takes for a field with nc components
// in dimension dim (-1: not supported on the device).
int normExactCount(int type, int nc, int dim)
{
  switch(type) {
    case FE_NORM_L1: case FE_NORM_L2: case FE_NORM_LINF: case FE_NORM_SEMI_H1: case FE_NORM_H1:
    case FE_NORM_AREA:
      return nc == 1 || type == FE_NORM_AREA ? 0 : -1;
    case FE_NORM_VECTOR_L2: return nc == 2 ? 0 : -1;
    case FE_NORM_L1_ERROR: case FE_NORM_L2_ERROR: case FE_NORM_LINF_ERROR: case FE_NORM_LINF_UDOT_ERROR:
      return nc == 1 ? 1 : -1;
    case FE_NORM_VECTOR_L2_ERROR: case FE_NORM_VECTOR_LINF_ERROR: case FE_NORM_VECTOR_LINF_UDOT_ERROR:
      return nc == 2 ? 2 : -1;
    case FE_NORM_SEMI_H1_ERROR: return nc == 1 ? dim : -1;
    case FE_NORM_VECTOR_SEMI_H1_ERROR: return nc == 2 ? 4 : -1;
    case FE_NORM_H1_ERROR: return nc == 1 ? 1 + dim : -1;
    case FE_NORM_INTEGRAL: case FE_NORM_INTEGRAL_DT: return nc == 1 ? 0 : -1;
    case FE_NORM_INTEGRAL_SOL_DEPENDENT_F: return nc == 1 ? 1 : -1;
    case FE_NORM_INTEGRAL_F: return 1;
  }
  return -1;
}

}  // namespace

namespace {

// Scratch of the allocation-free entry point: the reference tables, the
// block partials, the result and the exact-solution tables, at fixed offsets
// (doubles).  One copy per GPU, held until the process ends.
constexpr size_t kScrTab = 0;
constexpr size_t kScrTabDoubles = ((sizeof(FeRefTablesHO) + 255) / 256) * 256 / sizeof(double);
constexpr int kNormMaxBlocks = 1024;  // grid cap of normElements (one partial per block)
constexpr size_t kScrPart = kScrTab + kScrTabDoubles;
constexpr size_t kScrOut = kScrPart + kNormMaxBlocks;
constexpr size_t kScrData = kScrOut + 32;
constexpr size_t kScrDataDoubles = (size_t)1 << 20;  // exact-solution tables, 8 MiB
constexpr size_t kScrDoubles = kScrData + kScrDataDoubles;
constexpr int kScrMaxDevices = 64;

static __device__ __align__(256) double scratch[kScrDoubles];

std::mutex g_scrMutex;
double *g_scrPtr[kScrMaxDevices] = {};
hipError_t g_scrErr[kScrMaxDevices] = {};
bool g_scrDone[kScrMaxDevices] = {};

// Device buffers of one norm evaluation beyond its inputs.
struct NormBufs {
  const double *tab;  // reference tables in the layout of the field's kernel
  double *data;       // exact-solution tables, dataCap doubles
  size_t dataCap;
  double *partials;   // >= kNormMaxBlocks doubles
  double *out;        // the folded value
};

// Checks the arguments of fe_hip_norm / fe_hip_norm_noalloc; returns the
// field kind and the number of exact-table doubles the call copies.
int normCheck(fe_hip_ctx *ctx, int type, int field, int nexact, const fe_hip_fn *exact, double *value, int *kindOut,
              size_t *tabDoubles, int nq)
{
  if(!ctx || !value || (nexact > 0 && !exact)) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->d_u && ctx->E > 0, "upload the mesh and set the sizes first");
  FE_REQUIRE(ctx, field >= 0 && field < (int)ctx->fields.size(), "unknown field");
  const int kind = ctx->fields[field].kind;
  const int nc = kind == FE_TRI_VP2 ? 2 : 1;
  const int need = normExactCount(type, nc, ctx->dim);
  FE_REQUIRE(ctx, need >= 0, "norm type not available on the device for this field: " + std::to_string(type));
  FE_REQUIRE(ctx, nexact == need, "norm type " + std::to_string(type) + " takes " + std::to_string(need) +
                                      " exact callback(s)");
  if(type == FE_NORM_LINF_UDOT_ERROR || type == FE_NORM_VECTOR_LINF_UDOT_ERROR || type == FE_NORM_INTEGRAL_DT)
    FE_REQUIRE(ctx, ctx->nsol == 1 || ctx->d_past[0], "the time derivative needs the solution history");
  size_t td = 0;
  for(int i = 0; i < nexact; ++i) {
    FE_REQUIRE(ctx, exact[i].kind >= FE_FN_CONST && exact[i].kind <= FE_FN_POLY_U, "unknown function kind");
    if(exact[i].kind == FE_FN_POLY_U) {
      FE_REQUIRE(ctx, type == FE_NORM_INTEGRAL_SOL_DEPENDENT_F,
                 "FE_FN_POLY_U depends on the solution: only INTEGRAL_SOL_DEPENDENT_F takes it");
      FE_REQUIRE(ctx, exact[i].data && exact[i].count >= 1 && exact[i].count <= 8,
                 "FE_FN_POLY_U takes 1 to 8 polynomial coefficients");
      td += (size_t)exact[i].count;
    } else if(exact[i].kind == FE_FN_ELEM || exact[i].kind == FE_FN_QP) {
      const int64_t cnt = exact[i].kind == FE_FN_ELEM ? ctx->E : (int64_t)ctx->E * nq;
      FE_REQUIRE(ctx, exact[i].data && exact[i].count == cnt, "exact-solution table has the wrong size");
      td += (size_t)cnt;
    }
  }
  *kindOut = kind;
  *tabDoubles = td;
  return FE_STATUS_OK;
}

// Copies the exact tables into b.data, launches the element and folding
// kernels on the context stream and returns the norm (stream synchronized).
int normRun(fe_hip_ctx *ctx, int type, int field, int kind, int nq, int nexact, const fe_hip_fn *exact,
            const NormBufs &b, double *value)
{
  const bool isMax = type == FE_NORM_LINF || type == FE_NORM_LINF_ERROR || type == FE_NORM_LINF_UDOT_ERROR ||
                     type == FE_NORM_VECTOR_LINF_ERROR || type == FE_NORM_VECTOR_LINF_UDOT_ERROR;
  NormArgs a{};
  a.E = ctx->E;
  a.type = type;
  a.nq = nq;
  a.conn = ctx->d_conn;
  a.adr = ctx->fields[field].adr;
  a.xyz = ctx->d_xyz;
  a.tab = b.tab;
  a.u = ctx->d_u;
  a.nsol = ctx->nsol;
  for(int j = 0; j <= fe_hip_ctx::kMaxPast; ++j) a.cn[j] = ctx->bdf[j];
  for(int j = 0; j < fe_hip_ctx::kMaxPast; ++j) a.past[j] = ctx->d_past[j];
  size_t off = 0;
  for(int i = 0; i < nexact; ++i) {
    FeFn &f = a.ex[i];
    f.kind = exact[i].kind;
    f.c0 = exact[i].c0;
    f.c1 = exact[i].c1;
    f.data = nullptr;
    f.n = f.kind == FE_FN_POLY_U ? (int)exact[i].count : 0;
    if(f.kind == FE_FN_ELEM || f.kind == FE_FN_QP || f.kind == FE_FN_POLY_U) {
      f.data = b.data + off;
      FE_HIP_TRY(ctx, hipMemcpyAsync(f.data, exact[i].data, sizeof(double) * exact[i].count, hipMemcpyHostToDevice,
                                     ctx->stream));
      off += (size_t)exact[i].count;
    }
  }
  const int G = std::max(1, std::min(kNormMaxBlocks, (ctx->E + 255) / 256));
  a.partials = b.partials;
  if(kind == FE_TRI_P1) hipLaunchKernelGGL((normElements<3, 2, 1>), dim3(G), dim3(256), 0, ctx->stream, a);
  else if(kind == FE_TRI_P2) hipLaunchKernelGGL((normElements<6, 2, 1>), dim3(G), dim3(256), 0, ctx->stream, a);
  else if(kind == FE_TET_P1) hipLaunchKernelGGL((normElements<4, 3, 1>), dim3(G), dim3(256), 0, ctx->stream, a);
  else if(kind == FE_TRI_VP2) hipLaunchKernelGGL((normElements<12, 2, 2>), dim3(G), dim3(256), 0, ctx->stream, a);
  else if(kind == FE_TRI_P3) hipLaunchKernelGGL((normElements<10, 2, 1>), dim3(G), dim3(256), 0, ctx->stream, a);
  else if(kind == FE_TRI_P4) hipLaunchKernelGGL((normElements<15, 2, 1>), dim3(G), dim3(256), 0, ctx->stream, a);
  else FE_REQUIRE(ctx, false, "field kind without a device norm");
  hipLaunchKernelGGL(normFinal, dim3(1), dim3(64), 0, ctx->stream, b.partials, G, isMax ? 1 : 0, b.out);
  FE_HIP_TRY(ctx, hipGetLastError());
  double r = 0.;
  FE_HIP_TRY(ctx, hipMemcpyAsync(&r, b.out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  // tiny negative sums from negative weights (feNorm.cpp:390-393)
  const bool integ = type == FE_NORM_INTEGRAL || type == FE_NORM_INTEGRAL_DT ||
                     type == FE_NORM_INTEGRAL_SOL_DEPENDENT_F || type == FE_NORM_INTEGRAL_F;
  if(!isMax && !integ && r < 0. && fabs(r) < 1e-14) r = fabs(r);
  const bool root = type == FE_NORM_L2 || type == FE_NORM_L2_ERROR || type == FE_NORM_VECTOR_L2 ||
                    type == FE_NORM_VECTOR_L2_ERROR || type == FE_NORM_SEMI_H1 || type == FE_NORM_SEMI_H1_ERROR ||
                    type == FE_NORM_H1 || type == FE_NORM_H1_ERROR || type == FE_NORM_VECTOR_SEMI_H1_ERROR;
  *value = root ? sqrt(r) : r;
  return FE_STATUS_OK;
}

}  // namespace

void feNormScratchInit(int device)
{
  if(device < 0 || device >= kScrMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scrMutex);
  if(g_scrDone[device]) return;
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  g_scrErr[device] = e;
  g_scrPtr[device] = e == hipSuccess ? p : nullptr;
  g_scrDone[device] = true;
}

extern "C" int fe_hip_norm(fe_hip_ctx *ctx, int type, int field, int quad_degree, int nexact,
                           const fe_hip_fn *exact, double *value)
{
  if(!ctx || !value || (nexact > 0 && !exact)) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->d_u && ctx->E > 0, "upload the mesh and set the sizes first");
  FE_REQUIRE(ctx, field >= 0 && field < (int)ctx->fields.size(), "unknown field");
  // reference tables of (kind, degree), cached
  const int key = ctx->fields[field].kind * 64 + quad_degree;
  double *tab = nullptr;
  auto it = ctx->normTabs.find(key);
  if(it != ctx->normTabs.end()) tab = it->second;
  else {
    FeRefTablesHO T;
    const int nq = feBuildRefTables(ctx->fields[field].kind, quad_degree, &T);
    FE_REQUIRE(ctx, nq > 0, "unsupported quadrature degree " + std::to_string(quad_degree));
    if(int rc = feUploadRefTables(ctx, T, &tab)) return rc;
    ctx->normTabs[key] = tab;
    ctx->normNq[key] = nq;
  }
  const int nq = ctx->normNq[key];
  int kind = 0;
  size_t tabDoubles = 0;
  if(int rc = normCheck(ctx, type, field, nexact, exact, value, &kind, &tabDoubles, nq)) return rc;
  // exact tables go to one cached device buffer
  if(tabDoubles > ctx->normDataCap) {
    if(ctx->d_normData) feFree(ctx, ctx->d_normData);
    ctx->d_normData = nullptr;
    ctx->normDataCap = 0;
    if(feAlloc(ctx, &ctx->d_normData, tabDoubles)) return FE_STATUS_OUT_OF_MEMORY;
    ctx->normDataCap = tabDoubles;
  }
  NormBufs b{tab, ctx->d_normData, ctx->normDataCap, ctx->d_partials, ctx->d_partials + 3 * ctx->maxBlocks + 8};
  return normRun(ctx, type, field, kind, nq, nexact, exact, b, value);
}

extern "C" int fe_hip_norm_noalloc(fe_hip_ctx *ctx, int type, int field, int quad_degree, int nexact,
                                   const fe_hip_fn *exact, double *value)
{
  if(!ctx || !value || (nexact > 0 && !exact)) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, field >= 0 && field < (int)ctx->fields.size(), "unknown field");
  FE_REQUIRE(ctx, ctx->device >= 0 && ctx->device < kScrMaxDevices,
             "fe_hip_norm_noalloc: device id beyond the scratch table (" + std::to_string(kScrMaxDevices) + ")");
  double *scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scrMutex);
    FE_REQUIRE(ctx, g_scrDone[ctx->device] && g_scrPtr[ctx->device],
               std::string("fe_hip_norm_noalloc: no device address for the scratch: ") +
                   hipGetErrorString(g_scrErr[ctx->device]));
    scr = g_scrPtr[ctx->device];
  }
  FeRefTablesHO T;
  const int nq = feBuildRefTables(ctx->fields[field].kind, quad_degree, &T);
  FE_REQUIRE(ctx, nq > 0, "unsupported quadrature degree " + std::to_string(quad_degree));
  int kind = 0;
  size_t tabDoubles = 0;
  if(int rc = normCheck(ctx, type, field, nexact, exact, value, &kind, &tabDoubles, nq)) return rc;
  FE_REQUIRE(ctx, tabDoubles <= kScrDataDoubles,
             "fe_hip_norm_noalloc: exact-solution tables of " + std::to_string(tabDoubles) +
                 " doubles exceed its scratch of " + std::to_string(kScrDataDoubles) + " doubles");
  // the reference tables in the layout the field's kernel loads, copied
  // into the scratch (the stream is synchronized before T goes away)
  FeRefTables lo;
  const void *img = &T;
  size_t imgBytes = sizeof(FeRefTablesHO);
  if(T.nf <= 6) {
    feLowerRefTables(T, &lo);
    img = &lo;
    imgBytes = sizeof(FeRefTables);
  }
  static_assert(sizeof(FeRefTablesHO) <= kScrTabDoubles * sizeof(double), "table slot of the scratch");
  FE_HIP_TRY(ctx, hipMemcpyAsync(scr + kScrTab, img, imgBytes, hipMemcpyHostToDevice, ctx->stream));
  NormBufs b{scr + kScrTab, scr + kScrData, kScrDataDoubles, scr + kScrPart, scr + kScrOut};
  const int rc = normRun(ctx, type, field, kind, nq, nexact, exact, b, value);
  if(rc == FE_STATUS_OK) return rc;
  (void)hipStreamSynchronize(ctx->stream);  // T and lo outlive every copy from them
  return rc;
}
