// Internal definitions shared by the libfe_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <cstdio>
#include <string>
#include <vector>

#include "fe_hip.h"

#define FE_WAVE 64

// ---------------------------------------------------------------------------
// Error handling: every HIP call is checked; a failure records the message in
// the context and returns FE_STATUS_ERROR through the C ABI.
#define FE_HIP_TRY(ctx, call)                                                                   \
  do {                                                                                          \
    hipError_t err_ = (call);                                                                   \
    if(err_ != hipSuccess) {                                                                    \
      (ctx)->setError(std::string(#call) + ": " + hipGetErrorString(err_), __FILE__, __LINE__); \
      return FE_STATUS_ERROR;                                                                   \
    }                                                                                           \
  } while(0)

#define FE_REQUIRE(ctx, cond, msg)                         \
  do {                                                     \
    if(!(cond)) {                                          \
      (ctx)->setError(std::string(msg), __FILE__, __LINE__); \
      return FE_STATUS_ERROR;                              \
    }                                                      \
  } while(0)

// ---------------------------------------------------------------------------
// Reference-element tables of one (element kind, quadrature degree) pair,
// staged into LDS by the element kernels.
constexpr int kMaxQuad = 36;  // quadrature points of one rule (the degree-12 triangle rule has 33)
// Reference-element tables of one kind and quadrature rule, with room for
// NFM shape functions: the kernels of P1, P2 and the vector P2 pair load the
// 6-function layout into LDS, those of P3 / P4 triangles the 15-function one.
template <int NFM>
struct FeRefTablesT {
  int nq, nf, dim, nv;
  double w[kMaxQuad];
  double L[kMaxQuad][NFM];
  double dLr[kMaxQuad][NFM], dLs[kMaxQuad][NFM], dLt[kMaxQuad][NFM];
  double Lg[kMaxQuad][4];  // P1 geometric basis at the quadrature points
};
constexpr int kMaxNfHO = 15;
using FeRefTables = FeRefTablesT<6>;
using FeRefTablesHO = FeRefTablesT<kMaxNfHO>;
// The LDS layout of a kernel instantiated for NF shape functions.
template <int NF>
using FeRefTablesFor = FeRefTablesT<(NF > 6 ? kMaxNfHO : 6)>;

int feBuildRefTables(int kind, int degree, FeRefTablesHO *t);  // host, returns nq or <0
// Device copy of host tables in the layout their kernels load (FeRefTables
// for nf <= 6, FeRefTablesHO otherwise); *d is owned by the context.
int feUploadRefTables(fe_hip_ctx *ctx, const FeRefTablesHO &t, double **d);
// The FeRefTables (nf <= 6) form of host tables, as feUploadRefTables uploads it.
void feLowerRefTables(const FeRefTablesHO &t, FeRefTables *lo);
// Looks up, once per device, the device address of the allocation-free
// norm's static scratch (the device must be current); fe_hip_ctx_create
// calls it.  A failed lookup is reported by fe_hip_norm_noalloc.
void feNormScratchInit(int device);

struct FeFn {
  int kind;
  double c0, c1;
  double *data;  // device copy for FE_FN_ELEM / FE_FN_QP (mesh element order) / FE_FN_POLY_U
  int n;         // FE_FN_POLY_U: number of coefficients
};

struct FeField {
  int kind, nloc;
  int *adr = nullptr;  // device, E * nloc, original element order
};

struct FeForm {
  int sysType;
  std::vector<int> fields;  // indices into ctx->fields, reference order
  int degree;
  std::vector<FeFn> coeffs;
  FeRefTablesHO tab;
  double *d_tab = nullptr;  // device copy of tab
  bool hasMatrix;
  // layouts (feSysElm createElementarySystem): field ids of rows / cols
  std::vector<int> layoutI, layoutJ;
  int nI = 0, nJ = 0;
  int group = -1;  // CSR/position-map group (identical layouts share one)
  // MFMA element path of DIFFUSION: reference tensors S[K][16 NT] with
  // K = nq * (metric terms) padded to a multiple of 4
  int mfmaK = 0, mfmaNT = 0;
  double *d_sref64 = nullptr;
  float *d_sref32 = nullptr;
};

// Distinct (layoutI, layoutJ) pairs of the matrix forms: the CSR candidates
// and the scatter-position maps are built per group.
struct FeGroup {
  std::vector<int> layoutI, layoutJ;
  int nI, nJ;
  int *d_adrI = nullptr;  // E * nI, color-sorted element order
  int *d_adrJ = nullptr;  // E * nJ
  int *d_pos = nullptr;   // (nI*nJ) * E, slot-major, -1 when filtered
};

// Row-gather assembly of the fused scalar forms (fe_gather.hip): per field,
// the incidences (element << 4 | local index) of every row in color-sorted
// element order, slot-major per 64-row slice.
constexpr int kGatherMaxRow = 30;  // longest CSR row the gather kernel accumulates (5-bit slots)
struct FeGatherField {
  bool built = false, usable = false;
  int nSlices = 0, maxR = 0;
  int64_t incBytes = 0;
  int *d_sliceOff = nullptr;
  int2 *d_inc = nullptr;
  int *d_order = nullptr;  // slices in Morton order of their incident elements' centroids
};
// One fused scalar sweep handed to the gather kernel.
struct FeGatherSweep {
  int field, what, nq;
  const double *tab;
  int hasDiff, hasSrc, hasMass;
  FeFn coeff, src, mass;
  float *val32;   // non-null: write the FP32 rounding of the updated values
  bool fp32Math;  // element rows computed and summed in FP32
  bool freshVal;  // the values are zero and no earlier sweep of the call wrote them: stored, not read
};

struct FeKernelTimer {
  std::vector<hipEvent_t> ev;  // pairs
  size_t used = 0;
  double total_ms = 0.;
  int64_t launches = 0;
};
// Records the next event of t on the stream (growing the pool); after the
// stream has been synchronized, feTimerCollect adds the complete pairs to
// the totals (an open pair's first event is kept).
int feTimerMark(fe_hip_ctx *ctx, FeKernelTimer &t);
int feTimerCollect(fe_hip_ctx *ctx, FeKernelTimer &t);

// Preconditioner state: ILU(0)/IC(0) factor on A's pattern, row levels, and
// the split factors G (lower CSR, unit diagonal) / G^T (upper CSR) / D^-1.
struct FePrecond {
  bool patternReady = false, factored = false;
  std::vector<int> levelPtr;
  int *d_levelRows = nullptr, *d_iaL = nullptr, *d_jaL = nullptr, *d_iaU = nullptr, *d_jaU = nullptr;
  int *d_nShift = nullptr;
  int64_t nnzL = 0, nnzU = 0;
  double *d_lu = nullptr, *d_gL64 = nullptr, *d_gU64 = nullptr, *d_dinv64 = nullptr, *d_dval = nullptr;
  float *d_gL32 = nullptr, *d_gU32 = nullptr, *d_dinv32 = nullptr;
  int *d_tposL = nullptr;  // for each lower entry of G, its position in the upper CSR of G^T
  int *d_tposA = nullptr;  // for each entry (i, j) of A, the position of (j, i)
  int *d_rowOf = nullptr;  // row of each entry of A
  bool symmetricPattern = false;
  double *d_icL = nullptr, *d_icD = nullptr;  // parallel IC(0) iterates (two of each)
  int lastShifts = 0;
  double lastSetupMs = 0.;
  int sweeps = 0;  // GMRES/ILU(0): 0 = exact level-scheduled triangular solves, k = k Jacobi sweeps per factor
  // dependency levels of the upper triangle (row i after every j > i of its
  // row): the order of the exact backward solve; device copies of both level
  // pointers for the kernels that run several small levels in one workgroup
  std::vector<int> levelPtrU;
  int *d_levelRowsU = nullptr, *d_levelPtr = nullptr, *d_levelPtrU = nullptr;
  // launches of the exact solves: a level of many rows on its own grid, or
  // a run of consecutive small levels in one workgroup
  struct TriSeg {
    int l0, l1;
    bool run;
  };
  std::vector<TriSeg> triL, triU;
};

struct FeMpnn;   // learned-preconditioner state (fe_learned.hip)
struct FeAmg;    // aggregation multigrid hierarchy (fe_amg.hip)
struct FeAmgPcg;  // nested-ordered multigrid PCG (fe_amg_pcg.hip)
struct FeFieldSplit;  // Schur block preconditioner of saddle-point systems (fe_fieldsplit.hip)
struct FeDirect;      // dense LU of the direct solve (fe_direct.hip)

// Host copy of the aggregation hierarchy of the CSR pattern (built by
// feAmgSetup): per level the CSR pattern, the aggregate of each row (empty
// on the coarsest) and, for levels >= 1, the entries of the previous level
// (positions in its CSR) that sum into each entry (cptr / clist).
struct FeAmgHostLevel {
  int n = 0, nc = 0;
  std::vector<int> ia, ja, agg, cptr, clist;
};
struct FeBatch;  // batched-solver state (fe_batch.hip)

struct fe_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool borrowedStream = false;  // a sub-context on its parent's stream: not destroyed with it
  std::string err;
  FeFieldSplit *fieldSplit = nullptr;
  FeDirect *direct = nullptr;

  // mesh
  int dim = 0, V = 0, E = 0, nvpe = 0;
  double *d_xyz = nullptr;
  int *d_conn = nullptr;  // original element order

  // coloring: elements sorted by color
  std::vector<int> colorStart;  // ncolors + 1
  int *d_perm = nullptr;        // color-sorted position -> element id
  std::vector<int> h_perm;
  bool colored = false;

  std::vector<FeField> fields;
  std::vector<FeForm> forms;
  std::vector<FeGroup> groups;

  int n = 0, nDofs = 0;

  // CSR
  int64_t nnz = 0;
  int *d_ia = nullptr, *d_ja = nullptr, *d_diag = nullptr;
  double *d_val = nullptr;
  float *d_val32 = nullptr;
  double *d_rhs = nullptr, *d_u = nullptr, *d_du = nullptr;
  bool csrBuilt = false;
  bool matrixZeroed = false;  // values (and the FP32 copy) zeroed, no matrix assembled since
  // Time integration: u_dot = sum_{j < nsol} bdf[j] sol[j] with sol[0] = d_u
  // (the current iterate) and sol[j] = d_past[j - 1] (nDofs each, allocated
  // on first use); TRANSIENT_MASS matrices carry c0 = bdf[0].
  static constexpr int kMaxPast = 4;
  double *d_past[kMaxPast] = {};
  int nsol = 1;
  double bdf[kMaxPast + 1] = {};
  // Constraints applied to the assembled system: rows of essential vector
  // components (sorted, < n; conFlag[n] marks them) and periodic
  // (master, slave) pairs of unknowns, in the reference's map order.
  std::vector<int> conRows;
  int *d_conFlag = nullptr;
  std::vector<int> perMaster, perSlave;
  int *d_perSlave = nullptr, *d_perMaster = nullptr;  // slave rows and the master each one ends with
  int nPerRows = 0;
  // norms: reference tables per (kind, degree) and a buffer for exact-solution tables
  std::map<int, double *> normTabs;
  std::map<int, int> normNq;
  double *d_normData = nullptr;
  size_t normDataCap = 0;
  // launch shapes of the Krylov iteration kernels (fe_hip_set_tuning)
  int spmvLanes = 4;
  int spmvUnroll = 1;
  int iterBlock = 512;
  int iterGrid = 1024;
  bool lanesFixed = false;
  int asmKernel = 2;     // 0: colored scalar FP64 loop (reference order), 1: colored MFMA GEMM form,
                         // 2: row gather (one launch, colored summation order) where the sweep allows it
  int gatherXcd = 1;     // XCD-aware slice order in the gather assembly
  int asmFp32Math = 0;   // FE_PREC_FP32_IR gather: 0 = FP32 values rounded from the FP64 rows,
                         // 1 = element rows computed and summed in FP32
  std::map<int, FeGatherField> gather;
  int scalarLanes = 0;   // scalar sweeps: 1 = one lane per local row (measured no faster), 0 = one thread per element
  int fp32Values = 1;    // FP32 values: 0 = rounded FP64 values, 1 = FP32 MFMA element matrices
  int pipeKernel = 2;  // pipelined PCG: 0 Chronopoulos-Gear, 1 tiled Ghysels-Vanroose, 2 by precision
  int icSweeps = 0;    // FE_PC_IC0: 0 exact ILU(0), > 0 parallel IC(0) sweeps
  int pcValues32 = 1;  // FP64 split PCG: factor values read as FP32 (vectors and arithmetic FP64)
  int pcFp32 = 0;      // FP64 split PCG: apply G^T D^-1 G in FP64 (0) or FP32 (1; the pipelined
                       // recurrences then drift from M^-1 r: 20% more iterations measured)
  // FE_PC_AMG: l1-Jacobi factor, coarse over-correction, rows up to which a
  // level runs inside the single-workgroup kernel
  double amgOmega = 1.8, amgOverCorrect = 1.8;
  int amgBlockRows = 1024;
  int amgXcd = 1;    // XCD-aware tile order in the multigrid level kernels
  int xcdTiles = 0;  // XCD-aware tile order in the row-tile PCG kernels (measured slower at 512-row tiles)
  int amgLanes = 4;  // lanes per row in the multigrid level kernels
  int amgPadTiles = 1;  // restriction over aggregate-aligned tiles (0: tile-header kernel)
  // per-level multigrid: levels of at least this many rows store the gathered
  // vectors (l1 b, u + oc e[agg]) in a row pass first (0: never)
  int amgPreRows = 1 << 20;
  int batchAmgGlobal = 0;  // batched AMG: 0 per-problem CG scalars, 1 one CG over the union system
  FeAmg *amg = nullptr;
  std::vector<FeAmgHostLevel> amgHost;  // the hierarchy's pattern on the host (empty before the first AMG solve)
  int amgHostSweeps = 0;                // stalled coarsening: sweeps on the coarsest level
  int amgHostMaxAgg = 0;                // largest aggregate (rows) of any level
  FeAmgPcg *amgPcg = nullptr;
  int amgWcycle = 0;  // nested PCG with FP32 level values: W-cycle kernel for the coarse levels (1) or V-cycle (0)
  int amgFuse = 0;    // nested PCG: levels 1 .. coarsest in one launch with grid barriers (1) or one launch per level (0)
  int batchGraph = 1;  // batched multigrid PCG: chunks of iterations replayed from a captured graph
  int batchSkip = 1;   // batched multigrid PCG: the V-cycle skips the tiles of stopped problems
  int gmresOrth = 1;   // GMRES Arnoldi: 1 = classical Gram-Schmidt once (PETSc's default), 2 = twice (CGS2)
  int nsMfma = 1;     // Taylor-Hood Jacobian sweeps (stress + divergence + convection) on the matrix cores
  int amgKernel = 2;  // FE_PC_AMG: 1 = per-level kernels over the original numbering,
                      // 2 = nested-ordered fused PCG (fe_amg_pcg.hip)
  bool amgActive = false;  // the current CG solve applies the AMG V-cycle
  FeMpnn *mpnn = nullptr;
  FeBatch *batch = nullptr;    // Jacobi-PCG: 0 = on-the-fly D^-1 r gather, 1 = tiled (one gathered vector)

  // reductions
  double *d_partials = nullptr;
  unsigned *d_counter = nullptr;
  int maxBlocks = 0;

  // Krylov workspace (allocated on first solve)
  void *d_work = nullptr;
  size_t workBytes = 0;
  void *d_state = nullptr;
  void *h_state = nullptr;  // pinned mirror
  void *h_iterParams = nullptr;  // pinned: iteration parameters of graph replays
  int graphIters = 0;  // iterations per captured graph in the split/multigrid PCG (0: direct launches; graphs
                       // measured no faster here: the iteration kernels, not the host, set the pace)
  std::map<std::string, void *> graphs;  // instantiated iteration graphs (hipGraphExec_t)

  FePrecond pc;

  // measurement
  bool profiling = false;
  FeKernelTimer spmvTimer;
  FeKernelTimer iterTimer;  // whole iterations of the nested multigrid PCG (all their launches)
  double lastAsmMs = 0.;
  hipEvent_t asmEv[2] = {nullptr, nullptr};  // bracket the last fe_hip_assemble, read lazily
  bool asmPending = false;

  std::vector<void *> owned;  // device allocations freed at destroy

  void setError(const std::string &m, const char *file, int line)
  {
    char buf[64];
    snprintf(buf, sizeof(buf), " (%s:%d)", file, line);
    err = m + buf;
  }
};

// Allocation helper: device buffer owned by the context.
template <typename T>
int feAlloc(fe_hip_ctx *ctx, T **p, size_t count)
{
  void *q = nullptr;
  size_t bytes = count * sizeof(T);
  if(bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&q, bytes);
  if(e != hipSuccess) {
    ctx->setError(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e),
                  __FILE__, __LINE__);
    return FE_STATUS_OUT_OF_MEMORY;
  }
  ctx->owned.push_back(q);
  *p = (T *)q;
  return FE_STATUS_OK;
}
void feFree(fe_hip_ctx *ctx, void *p);

// Exclusive scan of n int32 counts into out (n+1 entries, out[n] = total).
// Returns the total through *total (host), or an error when it overflows int32.
int feExclusiveScan(fe_hip_ctx *ctx, const int *d_in, int *d_out, int n, int64_t *total);

// Gather assembly (fe_gather.hip): incidences of a field (built once per
// pattern; *ok = false when the field or pattern does not fit the kernel),
// and one fused scalar sweep.
int feGatherPrepare(fe_hip_ctx *ctx, int field, bool *ok);
int feGatherLaunch(fe_hip_ctx *ctx, const FeGatherSweep &s);

// Preconditioner set-up (fe_precond.hip)
int fePrecondSetupPattern(fe_hip_ctx *ctx);
// sweeps = 0: exact level-scheduled ILU(0); > 0: parallel IC(0) sweeps.
int fePrecondFactor(fe_hip_ctx *ctx, int sweeps = 0);
// Parallel IC(0) sweeps that feed the learned correction.
constexpr int kLearnedIcSweeps = 3;

// Learned correction of the split factor (fe_learned.hip)
int feLearnedCorrect(fe_hip_ctx *ctx);
void feMpnnRelease(fe_hip_ctx *ctx);

// Aggregation multigrid (fe_amg.hip): hierarchy of the CSR pattern (once),
// numeric set-up per matrix, one V-cycle out = M^-1 in (skipped once *done).
int feAmgSetup(fe_hip_ctx *ctx);
int feAmgNumeric(fe_hip_ctx *ctx);
int feAmgCheck(fe_hip_ctx *ctx);
int feAmgApply(fe_hip_ctx *ctx, const double *in, double *out, const int *done, bool values32);
int feAmgApply(fe_hip_ctx *ctx, const float *in, float *out, const int *done, bool values32);
int feAmgLevels(const fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap);
// Block-diagonal systems of independent problems (problem p: rows
// [rows[p], rows[p + 1]) of level 0): each tile of the level kernels gets the
// range of problems it touches, and with act (device, nprob flags, nonzero:
// the problem has stopped) the V-cycle's level launches skip the tiles whose
// problems have all stopped.  act = nullptr detaches (every tile runs).
int feAmgProblemTiles(fe_hip_ctx *ctx, int nprob, const int *rows, const int *act);
// First level of the single-workgroup kernel for a (vector, value) type slot
// (0 f64/f64, 1 f64/f32, 2 f32/f32): part of the launch sequence a graph captures.
int feAmgPlanFrom(const fe_hip_ctx *ctx, int slot);
void feAmgRelease(fe_hip_ctx *ctx);
// Nested-ordered multigrid PCG (fe_amg_pcg.hip): the FE_PC_AMG hierarchy with
// the rows of every level ordered so that each aggregate is a contiguous
// range, and standard PCG whose reductions are folded into the next kernel.
// Solves A dx = r with FP64 refinement over inner solves of vector type
// float (precision FE_PREC_FP32_IR) or double; dx accumulates into x64.
struct FeAmgPcgResult {
  int iters = 0, outer = 0, converged = 0, reason = 0, breakdowns = 0;
  double relres = 0.;
};
// FP64 residual of the system: r = rhs - A x (original CSR), *out = ||r||_2^2
// (stream synchronized).
int feResidualNorm2(fe_hip_ctx *ctx, const double *x, double *r, double *out);
int feSolveDirect(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double *du, fe_hip_solve_stats *stats);
void feDirectRelease(fe_hip_ctx *ctx);
int feAmgPcgSolve(fe_hip_ctx *ctx, const fe_hip_solver_opts *o, double bnorm, double target, double *x64,
                  double *r64, FeAmgPcgResult *res);
int feAmgPcgLevels(const fe_hip_ctx *ctx, int *levels, int *rows, int64_t *nnz, int cap);
// The nested-ordered PCG takes the hierarchy: two levels at least, a dense
// coarsest level, aggregates that fit a tile (after feAmgSetup).
bool feAmgNestedUsable(const fe_hip_ctx *ctx);
int feAmgPcgWLevel(const fe_hip_ctx *ctx);  // first level of the W-cycle kernel, 0 if none
int feAmgPcgWDense(const fe_hip_ctx *ctx);  // its dense level, 0 if none
int feAmgPcgFuseGrid(const fe_hip_ctx *ctx);  // workgroups of the fused coarse-level launch, 0 if not used
// One V-cycle z = M^-1 r through the nested-ordered kernels (device arrays of
// n doubles, original numbering), FP64 vectors, values per pcValues32:
// feAmgPcgApply = feAmgPcgPrepare (pattern once, numeric set-up of the
// current values) + feAmgPcgVcycle (the V-cycle alone).
int feAmgPcgApply(fe_hip_ctx *ctx, const double *r, double *z);
int feAmgPcgPrepare(fe_hip_ctx *ctx);
int feAmgPcgVcycle(fe_hip_ctx *ctx, const double *r, double *z);

// Block upper-triangular Schur preconditioner of a saddle-point system
// (fe_fieldsplit.hip): set-up of the current values, then z = P^-1 r on
// device vectors of n doubles.
int feFieldSplitSetup(fe_hip_ctx *ctx);
int feFieldSplitApply(fe_hip_ctx *ctx, const double *r, double *z);
void feFieldSplitRelease(fe_hip_ctx *ctx);
void feAmgPcgRelease(fe_hip_ctx *ctx);
void feBatchRelease(fe_hip_ctx *ctx);

// GMRES (fe_gmres.hip)
int feSolveGmres(fe_hip_ctx *ctx, const fe_hip_solver_opts *opts, double *du, fe_hip_solve_stats *stats);

// Krylov entry (fe_krylov.hip)
int feSolve(fe_hip_ctx *ctx, const fe_hip_solver_opts *opts, double *du,
            fe_hip_solve_stats *stats);

// ---------------------------------------------------------------------------
// Device helpers
__device__ __forceinline__ double feWaveSum(double v)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over k = kb, kb + STRIDE, ... < ke of val[k] * f(ja[k]).  The column
// indices and values of KS entries are loaded before any gathered value is
// used, so a lane keeps KS independent gathers in flight instead of one
// dependent pair per entry; the accumulation order is that of the plain loop.
template <int KS, int STRIDE, typename T, typename F>
__device__ __forceinline__ T feRowDot(const int *__restrict__ ja, const T *__restrict__ val, int kb, int ke, F f)
{
  T acc = (T)0;
  for(int k = kb; k < ke; k += KS * STRIDE) {
    int j[KS];
    T v[KS];
    bool ok[KS];
#pragma unroll
    for(int s = 0; s < KS; ++s) {
      const int kk = k + s * STRIDE;
      ok[s] = kk < ke;
      j[s] = ok[s] ? ja[kk] : 0;
      v[s] = ok[s] ? val[kk] : (T)0;
    }
#pragma unroll
    for(int s = 0; s < KS; ++s)
      if(ok[s]) acc += v[s] * f(j[s]);
  }
  return acc;
}

// Entries a lane of an LPR-lane row group keeps in flight: 16 per lane.
template <int LPR>
struct FeKs {
  static constexpr int value = LPR >= 16 ? 1 : 16 / LPR;
};

// Block sum: result valid in thread 0. lds must hold blockDim/64 doubles.
__device__ __forceinline__ double feBlockSum(double v, double *lds)
{
  v = feWaveSum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if(lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.;
  if(threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for(int i = 0; i < nw; ++i) s += lds[i];
  }
  return s;
}

// Cross-block reductions never use per-block device-scope fences: on this
// multi-XCD part an agent-scope release is an L2 write-back (buffer_wbl2),
// which costs tens of microseconds when every block issues one.  Instead each
// block stores its partial with a plain store, and the partials are folded
// after the kernel boundary (which is the only release/acquire needed): by
// the next kernel of the iteration, or by feFinalizeSums.

// Sum of G partials (fixed order per thread, then block tree): every thread
// of the block gets the total.  lds must hold blockDim/64 + 1 doubles.
__device__ __forceinline__ double feFoldPartials(const double *__restrict__ part, int G, double *lds)
{
  double s = 0.;
  for(int b = threadIdx.x; b < G; b += blockDim.x) s += part[b];
  s = feWaveSum(s);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if(lane == 0) lds[wid] = s;
  __syncthreads();
  if(threadIdx.x == 0) {
    double t = 0.;
    const int nw = (blockDim.x + 63) >> 6;
    for(int i = 0; i < nw; ++i) t += lds[i];
    lds[nw] = t;
  }
  __syncthreads();
  return lds[(blockDim.x + 63) >> 6];
}

// Block partial store: thread 0 writes the block sum of v into part[blockIdx].
__device__ __forceinline__ void feStoreBlockSum(double v, double *__restrict__ part, double *lds)
{
  const double b = feBlockSum(v, lds);
  if(threadIdx.x == 0) part[blockIdx.x] = b;
  __syncthreads();
}
