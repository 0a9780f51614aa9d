// Context lifetime, uploads, fields, forms and solution I/O of libfe_hip.so.
#include "fe_jacobian_check.h"
#include "fe_hip_internal.h"
#include "fe_lagrange.h"

#include <algorithm>
#include <cmath>
#include <cstring>

// ---------------------------------------------------------------------------
// Quadrature rules (data of feQuadratureTri.cpp:788-920 and
// feQuadratureTet.cpp:96-110) and reference bases (feSpace_2D.cpp:118-134,
// :536-553, :727-741; feSpace_3D.cpp:104-242 with n = 1).
namespace {

struct QuadBuilder {
  double r[kMaxQuad], s[kMaxQuad], t[kMaxQuad], w[kMaxQuad];
  int n = 0;
  void pt(double a, double b, double wt) { r[n] = a; s[n] = b; t[n] = 0.; w[n] = wt; ++n; }
  void orbit3(double a, double wt) { double b = 1. - 2. * a; pt(a, a, wt); pt(a, b, wt); pt(b, a, wt); }
  // the orbit of (a, a, b) given b, a = (1 - b) / 2
  void orbit3b(double b, double wt) { const double a = (1. - b) / 2.; pt(a, a, wt); pt(a, b, wt); pt(b, a, wt); }
  void orbit3r(double a, double b, double c, double wt) { pt(a, b, wt); pt(c, a, wt); pt(b, c, wt); }
  void orbit6(double a, double b, double wt)
  {
    double c = 1. - a - b;
    pt(a, b, wt); pt(b, a, wt); pt(a, c, wt); pt(c, a, wt); pt(b, c, wt); pt(c, b, wt);
  }
};

bool triRule(int degree, QuadBuilder &q)
{
  switch(degree) {
    case 0:
    case 1: q.pt(1. / 3., 1. / 3., 0.5); return true;
    case 2: q.orbit3(1. / 6., 1. / 6.); return true;
    case 3: q.pt(1. / 3., 1. / 3., -0.28125); q.orbit3(0.2, 25. / 96.); return true;
    case 4:
      q.orbit3(0.091576213509770743460, 0.054975871827660933819);
      q.orbit3(0.44594849091596488632, 0.11169079483900573285);
      return true;
    case 5:
      q.pt(1. / 3., 1. / 3., 0.1125);
      q.orbit3(0.10128650732345633880, 0.062969590272413576298);
      q.orbit3(0.47014206410511508977, 0.066197076394253090369);
      return true;
    case 6:
      q.orbit3(0.063089014491502228340, 0.025422453185103408460);
      q.orbit3(0.24928674517091042129, 0.058393137863189683013);
      q.orbit6(0.053145049844816947353, 0.31035245103378440542, 0.041425537809186787597);
      return true;
    case 7:
      q.orbit3r(0.062382265094402118174, 0.067517867073916085443,
                1. - 0.062382265094402118174 - 0.067517867073916085443, 0.026517028157436251429);
      q.orbit3r(0.055225456656926611737, 0.32150249385198182267,
                1. - 0.055225456656926611737 - 0.32150249385198182267, 0.043881408714446055037);
      q.orbit3r(0.034324302945097146470, 0.66094919618673565761, 0.30472650086816719592,
                0.028775042784981585738);
      q.orbit3r(0.51584233435359177926, 0.27771616697639178257, 0.20644149867001643817,
                0.067493187009802774463);
      return true;
    case 8:
      q.pt(1. / 3., 1. / 3., 0.0721578038388935841255455552445323);
      q.orbit3(0.170569307751760206622293501491464, 0.0516086852673591251408957751460645);
      q.orbit3(0.0505472283170309754584235505965989, 0.0162292488115990401554629641708902);
      q.orbit3(0.459292588292723156028815514494169, 0.0475458171336423123969480521942921);
      q.orbit6(0.008394777409957605337213834539296, 0.263112829634638113421785786284643,
               0.0136151570872174971324223450369544);
      return true;
    case 9:
      q.pt(1. / 3., 1. / 3., 0.0485678981413994169096209912536443);
      q.orbit3b(0.020634961602524744433, 0.0156673501135695352684274156436046);
      q.orbit3b(0.12582081701412672546, 0.0389137705023871396583696781497019);
      q.orbit3(0.188203535619032730240961280467335, 0.0398238694636051265164458871320226);
      q.orbit3(0.0447295133944527098651065899662763, 0.0127888378293490156308393992794999);
      q.orbit6(0.0368384120547362836348175987833851, 0.2219629891607656956751025276931919,
               0.0216417696886446886446886446886446);
      return true;
    case 10:
      q.pt(1. / 3., 1. / 3., 0.0454089951913767900476432975500142);
      q.orbit3b(0.028844733232685245264984935583748, 0.0183629788782333523585030359456832);
      q.orbit3(0.109481575485037054795458631340522, 0.0226605297177639673913028223692986);
      q.orbit6(0.141707219414879954756683250476361, 0.307939838764120950165155022930631,
               0.0363789584227100543021575883096803);
      q.orbit6(0.025003534762686386073988481007746, 0.246672560639902693917276465411176,
               0.0141636212655287424183685307910495);
      q.orbit6(0.0095408154002994575801528096228873, 0.0668032510122002657735402127620247,
               4.71083348186641172996373548344341E-03);
      return true;
    case 11:
      q.orbit6(0.0, 0.141129718717363295960826061941652, 3.68119189165027713212944752369032E-03);
      q.pt(1. / 3., 1. / 3., 0.0439886505811161193990465846607278);
      q.orbit3(0.0259891409282873952600324854988407, 4.37215577686801152475821439991262E-03);
      q.orbit3(0.0942875026479224956305697762754049, 0.0190407859969674687575121697178070);
      q.orbit3b(0.010726449965572372516734795387128, 9.42772402806564602923839129555767E-03);
      q.orbit3(0.207343382614511333452934024112966, 0.0360798487723697630620149942932315);
      q.orbit3b(0.122184388599015809877869236727746, 0.0346645693527679499208828254519072);
      q.orbit6(0.0448416775891304433090523914688007, 0.2772206675282791551488214673424523,
               0.0205281577146442833208261574536469);
      return true;
    case 12:
      q.orbit3b(2.35652204523900E-02, 1.28655332202275E-02);
      q.orbit3b(1.20551215411079E-01, 2.18462722690190E-02);
      q.orbit3(2.71210385012116E-01, 3.14291121089425E-02);
      q.orbit3(1.27576145541586E-01, 1.73980564653545E-02);
      q.orbit3(2.13173504532100E-02, 3.08313052577950E-03);
      q.orbit6(1.15343494534698E-01, 2.75713269685514E-01, 2.01857788831905E-02);
      q.orbit6(2.28383322222570E-02, 2.81325580989940E-01, 1.11783866011515E-02);
      q.orbit6(2.57340505483300E-02, 1.16251915907597E-01, 8.65811555432950E-03);
      return true;
  }
  return false;
}

bool tetRule(int degree, QuadBuilder &q)
{
  if(degree <= 1) {
    q.r[0] = q.s[0] = q.t[0] = 0.25;
    q.w[0] = 0.166666666666667;
    q.n = 1;
    return true;
  }
  if(degree == 2) {
    const double a = 0.138196601125, b = 0.585410196625;
    const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    for(int i = 0; i < 4; ++i) {
      q.r[i] = P[i][0]; q.s[i] = P[i][1]; q.t[i] = P[i][2]; q.w[i] = 0.0416666666666667;
    }
    q.n = 4;
    return true;
  }
  return false;
}

}  // namespace

int feBuildRefTables(int kind, int degree, FeRefTablesHO *T)
{
  memset(T, 0, sizeof(*T));
  QuadBuilder q;
  const bool tet = (kind == FE_TET_P1);
  if(!(tet ? tetRule(degree, q) : triRule(degree, q))) return -1;
  T->nq = q.n;
  T->dim = tet ? 3 : 2;
  T->nv = tet ? 4 : 3;
  // vector P2 uses the scalar P2 tables
  T->nf = (kind == FE_TRI_P1) ? 3 : tet ? 4 : kind == FE_TRI_P3 ? 10 : kind == FE_TRI_P4 ? 15 : 6;
  for(int k = 0; k < q.n; ++k) {
    const double r = q.r[k], s = q.s[k], t = q.t[k];
    T->w[k] = q.w[k];
    if(kind == FE_TRI_P3 || kind == FE_TRI_P4) {
      for(int i = 0; i < T->nf; ++i) feTriShape(kind, i, r, s, T->L[k][i], T->dLr[k][i], T->dLs[k][i]);
    } else if(kind == FE_TRI_P1) {
      const double L[3] = {1. - r - s, r, s}, dr[3] = {-1., 1., 0.}, ds[3] = {-1., 0., 1.};
      for(int i = 0; i < 3; ++i) { T->L[k][i] = L[i]; T->dLr[k][i] = dr[i]; T->dLs[k][i] = ds[i]; }
    } else if(tet) {
      const double L[4] = {1. - r - s - t, r, s, t};
      const double dr[4] = {-1., 1., 0., 0.}, ds[4] = {-1., 0., 1., 0.}, dt[4] = {-1., 0., 0., 1.};
      for(int i = 0; i < 4; ++i) {
        T->L[k][i] = L[i]; T->dLr[k][i] = dr[i]; T->dLs[k][i] = ds[i]; T->dLt[k][i] = dt[i];
      }
    } else {
      const double L[6] = {(1. - r - s) * (1. - 2. * r - 2. * s), r * (2. * r - 1.), s * (2. * s - 1.),
                           4. * r * (1. - r - s), 4. * r * s, 4. * s * (1. - r - s)};
      const double dr[6] = {4. * (r + s) - 3., 4. * r - 1., 0., 4. * (1. - 2. * r - s), 4. * s, -4. * s};
      const double ds[6] = {4. * (r + s) - 3., 0., 4. * s - 1., -4. * r, 4. * r, 4. * (1. - r - 2. * s)};
      for(int i = 0; i < 6; ++i) { T->L[k][i] = L[i]; T->dLr[k][i] = dr[i]; T->dLs[k][i] = ds[i]; }
    }
    if(tet) {
      T->Lg[k][0] = 1. - r - s - t; T->Lg[k][1] = r; T->Lg[k][2] = s; T->Lg[k][3] = t;
    } else {
      T->Lg[k][0] = 1. - r - s; T->Lg[k][1] = r; T->Lg[k][2] = s;
    }
  }
  return q.n;
}

void feLowerRefTables(const FeRefTablesHO &t, FeRefTables *lo)
{
  memset(lo, 0, sizeof(*lo));
  lo->nq = t.nq; lo->nf = t.nf; lo->dim = t.dim; lo->nv = t.nv;
  for(int k = 0; k < kMaxQuad; ++k) {
    lo->w[k] = t.w[k];
    for(int i = 0; i < 6; ++i) {
      lo->L[k][i] = t.L[k][i]; lo->dLr[k][i] = t.dLr[k][i]; lo->dLs[k][i] = t.dLs[k][i]; lo->dLt[k][i] = t.dLt[k][i];
    }
    for(int i = 0; i < 4; ++i) lo->Lg[k][i] = t.Lg[k][i];
  }
}

int feUploadRefTables(fe_hip_ctx *ctx, const FeRefTablesHO &t, double **d)
{
  if(t.nf > 6) {
    if(feAlloc(ctx, d, sizeof(FeRefTablesHO) / sizeof(double) + 1)) return FE_STATUS_OUT_OF_MEMORY;
    FE_HIP_TRY(ctx, hipMemcpy(*d, &t, sizeof(FeRefTablesHO), hipMemcpyHostToDevice));
    return FE_STATUS_OK;
  }
  FeRefTables lo;
  feLowerRefTables(t, &lo);
  if(feAlloc(ctx, d, sizeof(FeRefTables) / sizeof(double) + 1)) return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(*d, &lo, sizeof(FeRefTables), hipMemcpyHostToDevice));
  return FE_STATUS_OK;
}

void feFree(fe_hip_ctx *ctx, void *p)
{
  if(!p) return;
  auto it = std::find(ctx->owned.begin(), ctx->owned.end(), p);
  if(it != ctx->owned.end()) ctx->owned.erase(it);
  (void)hipFree(p);
}

static int nlocOf(int kind)
{
  switch(kind) {
    case FE_TRI_P1: return 3;
    case FE_TRI_P2: return 6;
    case FE_TRI_VP2: return 12;
    case FE_TET_P1: return 4;
    case FE_TRI_P3: return 10;
    case FE_TRI_P4: return 15;
  }
  return 0;
}

extern "C" {

int fe_hip_ctx_create(int device, fe_hip_ctx **out)
{
  if(!out) return FE_STATUS_POINTER_ERROR;
  *out = nullptr;
  int count = 0;
  if(hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
    return FE_STATUS_ERROR;
  fe_hip_ctx *ctx = new fe_hip_ctx;
  ctx->device = device;
  if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FE_STATUS_ERROR;
  }
  feNormScratchInit(device);  // allocation-free norm: device address of its scratch on this device
  ctx->maxBlocks = 65536;
  if(feAlloc(ctx, &ctx->d_partials, (size_t)ctx->maxBlocks * 4) ||
     feAlloc(ctx, &ctx->d_counter, 64)) {
    for(void *p : ctx->owned) (void)hipFree(p);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FE_STATUS_OUT_OF_MEMORY;
  }
  if(hipMemset(ctx->d_counter, 0, 64 * sizeof(unsigned)) != hipSuccess) {
    for(void *p : ctx->owned) (void)hipFree(p);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FE_STATUS_ERROR;
  }
  *out = ctx;
  return FE_STATUS_OK;
}

int fe_hip_ctx_destroy(fe_hip_ctx *ctx)
{
  if(!ctx) return FE_STATUS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for(hipEvent_t e : ctx->spmvTimer.ev) (void)hipEventDestroy(e);
  for(hipEvent_t e : ctx->iterTimer.ev) (void)hipEventDestroy(e);
  for(hipEvent_t e : ctx->asmEv)
    if(e) (void)hipEventDestroy(e);
  feMpnnRelease(ctx);
  feBatchRelease(ctx);
  feAmgRelease(ctx);
  feAmgPcgRelease(ctx);
  feFieldSplitRelease(ctx);
  feDirectRelease(ctx);
  for(void *p : ctx->owned) (void)hipFree(p);
  for(auto &g : ctx->graphs) (void)hipGraphExecDestroy((hipGraphExec_t)g.second);
  ctx->graphs.clear();
  if(ctx->h_state) (void)hipHostFree(ctx->h_state);
  if(!ctx->borrowedStream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FE_STATUS_OK;
}

const char *fe_hip_last_error(const fe_hip_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fe_hip_upload_mesh(fe_hip_ctx *ctx, int dim, int V, const double *xyz, int E, int nvpe,
                       const int *conn)
{
  if(!ctx || !xyz || !conn) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, (dim == 2 && nvpe == 3) || (dim == 3 && nvpe == 4), "unsupported element type");
  FE_REQUIRE(ctx, V > 0 && E > 0, "empty mesh");
  for(size_t i = 0; i < (size_t)E * nvpe; ++i)
    FE_REQUIRE(ctx, conn[i] >= 0 && conn[i] < V, "connectivity out of range");
  {
    double det = 0.;
    const int bad = feFirstNonPositiveJacobian(dim, xyz, E, nvpe, conn, &det);
    FE_REQUIRE(ctx, bad < 0, "Negative or zero jacobian = " + std::to_string(det) + " on element " +
                                 std::to_string(bad) + " (feCncGeo::computeJacobians)");
  }
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  ctx->dim = dim; ctx->V = V; ctx->E = E; ctx->nvpe = nvpe;
  if(feAlloc(ctx, &ctx->d_xyz, (size_t)3 * V) || feAlloc(ctx, &ctx->d_conn, (size_t)E * nvpe))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(ctx->d_xyz, xyz, sizeof(double) * 3 * V, hipMemcpyHostToDevice));
  FE_HIP_TRY(ctx, hipMemcpy(ctx->d_conn, conn, sizeof(int) * (size_t)E * nvpe, hipMemcpyHostToDevice));
  // Default coloring (first-fit over vertex adjacency); replaced by fe_hip_set_coloring.
  std::vector<int> colorOf(E);
  std::vector<uint64_t> used((size_t)V * 4, 0);
  for(int e = 0; e < E; ++e) {
    uint64_t forb[4] = {0, 0, 0, 0};
    for(int j = 0; j < nvpe; ++j)
      for(int w = 0; w < 4; ++w) forb[w] |= used[(size_t)conn[(size_t)e * nvpe + j] * 4 + w];
    int c = -1;
    for(int w = 0; w < 4 && c < 0; ++w)
      if(~forb[w]) c = 64 * w + __builtin_ctzll(~forb[w]);
    FE_REQUIRE(ctx, c >= 0, "more than 256 colors");
    colorOf[e] = c;
    for(int j = 0; j < nvpe; ++j) used[(size_t)conn[(size_t)e * nvpe + j] * 4 + c / 64] |= 1ull << (c % 64);
  }
  return fe_hip_set_coloring(ctx, colorOf.data());
}

int fe_hip_set_coloring(fe_hip_ctx *ctx, const int *colorOf)
{
  if(!ctx || !colorOf) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->E > 0, "upload the mesh first");
  FE_REQUIRE(ctx, !ctx->csrBuilt, "coloring must be set before fe_hip_build_csr");
  int nc = 0;
  for(int e = 0; e < ctx->E; ++e) {
    FE_REQUIRE(ctx, colorOf[e] >= 0, "negative color");
    nc = std::max(nc, colorOf[e] + 1);
  }
  ctx->colorStart.assign(nc + 1, 0);
  for(int e = 0; e < ctx->E; ++e) ctx->colorStart[colorOf[e] + 1]++;
  for(int c = 0; c < nc; ++c) ctx->colorStart[c + 1] += ctx->colorStart[c];
  std::vector<int> pos(ctx->colorStart.begin(), ctx->colorStart.end() - 1);
  ctx->h_perm.assign(ctx->E, 0);
  for(int e = 0; e < ctx->E; ++e) ctx->h_perm[pos[colorOf[e]]++] = e;
  if(!ctx->d_perm && feAlloc(ctx, &ctx->d_perm, ctx->E)) return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(ctx->d_perm, ctx->h_perm.data(), sizeof(int) * ctx->E, hipMemcpyHostToDevice));
  ctx->colored = true;
  return FE_STATUS_OK;
}

int fe_hip_add_field(fe_hip_ctx *ctx, int kind, const int *adr, int *field)
{
  if(!ctx || !adr) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->E > 0, "upload the mesh first");
  const int nl = nlocOf(kind);
  FE_REQUIRE(ctx, nl > 0, "unknown field kind");
  FE_REQUIRE(ctx, (ctx->dim == 3) == (kind == FE_TET_P1), "field kind does not match the mesh");
  FeField f;
  f.kind = kind;
  f.nloc = nl;
  if(feAlloc(ctx, &f.adr, (size_t)ctx->E * nl)) return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemcpy(f.adr, adr, sizeof(int) * (size_t)ctx->E * nl, hipMemcpyHostToDevice));
  ctx->fields.push_back(f);
  if(field) *field = (int)ctx->fields.size() - 1;
  return FE_STATUS_OK;
}

int fe_hip_set_sizes(fe_hip_ctx *ctx, int nInc, int nDofs)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, nInc > 0 && nDofs >= nInc, "invalid system sizes (no unknowns)");
  FE_REQUIRE(ctx, !ctx->csrBuilt, "sizes are fixed once the CSR is built");
  FE_REQUIRE(ctx, !ctx->d_past[0], "sizes are fixed once a solution history exists");
  ctx->n = nInc;
  ctx->nDofs = nDofs;
  if(feAlloc(ctx, &ctx->d_u, nDofs) || feAlloc(ctx, &ctx->d_rhs, nInc) || feAlloc(ctx, &ctx->d_du, nInc))
    return FE_STATUS_OUT_OF_MEMORY;
  FE_HIP_TRY(ctx, hipMemset(ctx->d_u, 0, sizeof(double) * nDofs));
  FE_HIP_TRY(ctx, hipMemset(ctx->d_rhs, 0, sizeof(double) * nInc));
  FE_HIP_TRY(ctx, hipMemset(ctx->d_du, 0, sizeof(double) * nInc));
  return FE_STATUS_OK;
}

int fe_hip_add_form(fe_hip_ctx *ctx, int sysType, int nfields, const int *fields, int quad_degree,
                    int ncoeff, const fe_hip_fn *coeffs, int *form)
{
  if(!ctx || !fields || (ncoeff > 0 && !coeffs)) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, !ctx->csrBuilt, "forms must be added before fe_hip_build_csr");
  FE_REQUIRE(ctx, nfields >= 1 && nfields <= 2 && ncoeff >= 0 && ncoeff <= 4, "bad form arity");
  for(int i = 0; i < nfields; ++i)
    FE_REQUIRE(ctx, fields[i] >= 0 && fields[i] < (int)ctx->fields.size(), "unknown field");
  FeForm F;
  F.sysType = sysType;
  F.fields.assign(fields, fields + nfields);
  F.degree = quad_degree;
  const int k0 = ctx->fields[fields[0]].kind;
  // Layouts as set by each feSysElm::createElementarySystem.
  switch(sysType) {
    case FE_SYS_DIFFUSION:
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "DIFFUSION takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "DIFFUSION takes one coefficient");
      break;
    case FE_SYS_NONLINEAR_DIFFUSION:
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "NONLINEAR_DIFFUSION takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 2, "NONLINEAR_DIFFUSION takes (conductivity k(u), its derivative dk/du)");
      break;
    case FE_SYS_MASS:
      // feSysElm_Mass (coefficient) or feSysElm_MassPower (coefficient,
      // exponent as an FE_FN_CONST), which reuses the MASS id (feSysElm.h:408-409)
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "MASS takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1 || (ncoeff == 2 && coeffs[1].kind == FE_FN_CONST),
                 "MASS takes a coefficient and optionally a constant exponent");
      break;
    case FE_SYS_ADVECTION:
      // velocity given by its dim components
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "ADVECTION takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == ctx->dim, "ADVECTION takes one coefficient per velocity component");
      break;
    case FE_SYS_TRANSIENT_MASS:
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "TRANSIENT_MASS takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "TRANSIENT_MASS takes one coefficient");
      break;
    case FE_SYS_SOURCE:
      FE_REQUIRE(ctx, nfields == 1 && k0 != FE_TRI_VP2, "SOURCE takes one scalar field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = false;
      FE_REQUIRE(ctx, ncoeff == 1, "SOURCE takes one function");
      break;
    case FE_SYS_VECTOR_SOURCE:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "VECTOR_SOURCE takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = false;
      FE_REQUIRE(ctx, ncoeff == 2, "VECTOR_SOURCE takes one function per component");
      break;
    case FE_SYS_VECTOR_CONVECTIVE_ACCELERATION:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "VECTOR_CONVECTIVE_ACCELERATION takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "VECTOR_CONVECTIVE_ACCELERATION takes one coefficient");
      break;
    case FE_SYS_TRANSIENT_VECTOR_MASS:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "TRANSIENT_VECTOR_MASS takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "TRANSIENT_VECTOR_MASS takes one coefficient");
      break;
    case FE_SYS_VECTOR_MASS:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "VECTOR_MASS takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "VECTOR_MASS takes one coefficient");
      break;
    case FE_SYS_GRAD_SOURCE:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "GRAD_SOURCE takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = false;
      FE_REQUIRE(ctx, ncoeff == 1, "GRAD_SOURCE takes one scalar function");
      break;
    case FE_SYS_VECTOR_DIFFUSION:
      FE_REQUIRE(ctx, nfields == 1 && k0 == FE_TRI_VP2, "VECTOR_DIFFUSION takes one vector field");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 2, "VECTOR_DIFFUSION takes (coefficient, diffusivity)");
      break;
    case FE_SYS_MIXED_MASS:
    case FE_SYS_MIXED_MASS_POWER:
    case FE_SYS_MIXED_GRADGRAD:
      // fields {v, u}: test functions of v, unknown u, one scalar kind
      // (feSysElm.cpp:283-293, :352-362, :886-897)
      FE_REQUIRE(ctx, nfields == 2 && k0 != FE_TRI_VP2 && ctx->fields[fields[1]].kind == k0 &&
                          fields[0] != fields[1],
                 "mixed scalar forms take {test field, trial field} of one scalar kind");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[1]}; F.hasMatrix = true;
      if(sysType == FE_SYS_MIXED_MASS_POWER)
        FE_REQUIRE(ctx, ncoeff == 2 && coeffs[1].kind == FE_FN_CONST,
                   "MIXED_MASS_POWER takes a coefficient and a constant exponent");
      else
        FE_REQUIRE(ctx, ncoeff == 1, "this mixed form takes one coefficient");
      break;
    case FE_SYS_MIXED_GRADIENT:
      // fields {u, p}: vector test functions, scalar unknown (feSysElm_MixedGradient)
      FE_REQUIRE(ctx, nfields == 2 && k0 == FE_TRI_VP2 && ctx->fields[fields[1]].kind == FE_TRI_P1,
                 "MIXED_GRADIENT takes {vector P2 u, P1 p}");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[1]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "MIXED_GRADIENT takes one coefficient");
      break;
    case FE_SYS_MIXED_DIVERGENCE:
      // fields {p, u}: scalar test functions, vector unknown (feVectorSysElm.cpp:669-683)
      FE_REQUIRE(ctx, nfields == 2 && ctx->fields[fields[0]].kind == FE_TRI_P1 &&
                        ctx->fields[fields[1]].kind == FE_TRI_VP2, "MIXED_DIVERGENCE takes {P1 p, vector P2 u}");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[1]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 1, "MIXED_DIVERGENCE takes one coefficient");
      break;
    case FE_SYS_DIV_NEWTONIAN_STRESS:
      // fields {u, p} (feVectorSysElm.cpp:1434-1450)
      FE_REQUIRE(ctx, nfields == 2 && k0 == FE_TRI_VP2 && ctx->fields[fields[1]].kind == FE_TRI_P1,
                 "DIV_NEWTONIAN_STRESS takes {vector P2 u, P1 p}");
      F.layoutI = {fields[0]}; F.layoutJ = {fields[0], fields[1]}; F.hasMatrix = true;
      FE_REQUIRE(ctx, ncoeff == 2, "DIV_NEWTONIAN_STRESS takes (coefficient, viscosity)");
      break;
    default:
      FE_REQUIRE(ctx, false, "weak form not supported on the device: " + std::to_string(sysType));
  }
  // Forms touching a vector P2 field use the P2 tables (the pressure basis is
  // the P1 geometric basis, also tabulated).
  int tabKind = k0;
  for(int f : F.fields)
    if(ctx->fields[f].kind == FE_TRI_VP2) tabKind = FE_TRI_VP2;
  int nq = feBuildRefTables(tabKind, quad_degree, &F.tab);
  FE_REQUIRE(ctx, nq > 0, "unsupported quadrature degree " + std::to_string(quad_degree));
  for(int c = 0; c < ncoeff; ++c) {
    FeFn fn;
    fn.kind = coeffs[c].kind;
    fn.c0 = coeffs[c].c0;
    fn.c1 = coeffs[c].c1;
    fn.data = nullptr;
    fn.n = 0;
    FE_REQUIRE(ctx, fn.kind >= FE_FN_CONST && fn.kind <= FE_FN_POLY_U, "unknown function kind");
    if(fn.kind == FE_FN_POLY_U) {
      // a function of the solution: only the nonlinear diffusion evaluates it
      FE_REQUIRE(ctx, sysType == FE_SYS_NONLINEAR_DIFFUSION, "FE_FN_POLY_U is a coefficient of NONLINEAR_DIFFUSION");
      FE_REQUIRE(ctx, coeffs[c].data && coeffs[c].count >= 1 && coeffs[c].count <= 8,
                 "FE_FN_POLY_U takes 1 to 8 polynomial coefficients");
      fn.n = (int)coeffs[c].count;
      if(feAlloc(ctx, &fn.data, (size_t)fn.n)) return FE_STATUS_OUT_OF_MEMORY;
      FE_HIP_TRY(ctx, hipMemcpy(fn.data, coeffs[c].data, sizeof(double) * fn.n, hipMemcpyHostToDevice));
    } else if(fn.kind == FE_FN_ELEM || fn.kind == FE_FN_QP) {
      const int64_t need = (fn.kind == FE_FN_ELEM) ? ctx->E : (int64_t)ctx->E * nq;
      FE_REQUIRE(ctx, coeffs[c].data && coeffs[c].count == need, "coefficient table has the wrong size");
      if(feAlloc(ctx, &fn.data, (size_t)need)) return FE_STATUS_OUT_OF_MEMORY;
      FE_HIP_TRY(ctx, hipMemcpy(fn.data, coeffs[c].data, sizeof(double) * need, hipMemcpyHostToDevice));
    }
    F.coeffs.push_back(fn);
  }
  if(sysType == FE_SYS_DIFFUSION && F.tab.nf <= 15) {
    // Reference tensors of the GEMM form Ae_ij = sum_{q,t} c_{q,t}(e) S_{q,t,ij}:
    // t runs over the unique metric terms (a <= b) of M_ab = sum_d dr_a/dx_d
    // dr_b/dx_d, S_{q,(a,b),ij} = w_q (d_a phi_i d_b phi_j + d_b phi_i d_a phi_j)
    // (a single product when a == b), c = k(x_q) J M_ab.
    const int dim = F.tab.dim, nf = F.tab.nf, nq2 = F.tab.nq;
    const int nt = dim == 2 ? 3 : 6;
    int K = nq2 * nt;
    K = (K + 3) & ~3;
    const int NE = nf * nf, NT = (NE + 15) / 16;
    std::vector<double> S((size_t)K * 16 * NT, 0.);
    const int ta[6] = {0, 0, 1, 0, 0, 1}, tb[6] = {0, 1, 1, 0, 0, 1};
    const int ta3[6] = {0, 0, 0, 1, 1, 2}, tb3[6] = {0, 1, 2, 1, 2, 2};
    for(int q = 0; q < nq2; ++q)
      for(int t = 0; t < nt; ++t) {
        const int a = dim == 2 ? ta[t] : ta3[t], b = dim == 2 ? tb[t] : tb3[t];
        for(int i = 0; i < nf; ++i)
          for(int j = 0; j < nf; ++j) {
            auto d = [&](int c, int f) { return c == 0 ? F.tab.dLr[q][f] : (c == 1 ? F.tab.dLs[q][f] : F.tab.dLt[q][f]); };
            double v = d(a, i) * d(b, j);
            if(a != b) v += d(b, i) * d(a, j);
            S[(size_t)(q * nt + t) * 16 * NT + i * nf + j] = F.tab.w[q] * v;
          }
      }
    F.mfmaK = K;
    F.mfmaNT = NT;
    if(feAlloc(ctx, &F.d_sref64, S.size()) || feAlloc(ctx, &F.d_sref32, S.size())) return FE_STATUS_OUT_OF_MEMORY;
    std::vector<float> S32(S.begin(), S.end());
    FE_HIP_TRY(ctx, hipMemcpy(F.d_sref64, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice));
    FE_HIP_TRY(ctx, hipMemcpy(F.d_sref32, S32.data(), sizeof(float) * S.size(), hipMemcpyHostToDevice));
  }
  if(int rc = feUploadRefTables(ctx, F.tab, &F.d_tab)) return rc;
  for(int f : F.layoutI) F.nI += ctx->fields[f].nloc;
  for(int f : F.layoutJ) F.nJ += ctx->fields[f].nloc;
  ctx->forms.push_back(F);
  if(form) *form = (int)ctx->forms.size() - 1;
  return FE_STATUS_OK;
}

int fe_hip_set_solution(fe_hip_ctx *ctx, const double *u)
{
  if(!ctx || !u) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->d_u, "set the sizes first");
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_u, u, sizeof(double) * ctx->nDofs, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_get_solution(fe_hip_ctx *ctx, double *u)
{
  if(!ctx || !u) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->d_u, "set the sizes first");
  FE_HIP_TRY(ctx, hipMemcpyAsync(u, ctx->d_u, sizeof(double) * ctx->nDofs, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

// ---------------------------------------------------------------------------
// Solution history of the BDF schemes (feSolutionContainer.cpp:82-99,291-305).
static int fePastReady(fe_hip_ctx *ctx)
{
  FE_REQUIRE(ctx, ctx->d_u, "set the sizes first");
  for(int j = 0; j < fe_hip_ctx::kMaxPast; ++j) {
    if(ctx->d_past[j]) continue;
    if(feAlloc(ctx, &ctx->d_past[j], (size_t)ctx->nDofs)) return FE_STATUS_OUT_OF_MEMORY;
    FE_HIP_TRY(ctx, hipMemsetAsync(ctx->d_past[j], 0, sizeof(double) * ctx->nDofs, ctx->stream));
  }
  return FE_STATUS_OK;
}

int fe_hip_set_bdf(fe_hip_ctx *ctx, int nsol, const double *cn)
{
  if(!ctx || !cn) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, nsol >= 1 && nsol <= fe_hip_ctx::kMaxPast + 1, "nsol must be in [1, 5]");
  if(nsol > 1 && fePastReady(ctx)) return FE_STATUS_ERROR;
  ctx->nsol = nsol;
  for(int j = 0; j <= fe_hip_ctx::kMaxPast; ++j) ctx->bdf[j] = j < nsol ? cn[j] : 0.;
  return FE_STATUS_OK;
}

int fe_hip_rotate_solutions(fe_hip_ctx *ctx)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  if(fePastReady(ctx)) return FE_STATUS_ERROR;
  // The oldest buffer is recycled as sol[1]; the others move down one slot.
  double *last = ctx->d_past[fe_hip_ctx::kMaxPast - 1];
  for(int j = fe_hip_ctx::kMaxPast - 1; j > 0; --j) ctx->d_past[j] = ctx->d_past[j - 1];
  ctx->d_past[0] = last;
  FE_HIP_TRY(ctx, hipMemcpyAsync(last, ctx->d_u, sizeof(double) * ctx->nDofs, hipMemcpyDeviceToDevice,
                                 ctx->stream));  // stream-ordered before the next sweep
  return FE_STATUS_OK;
}

int fe_hip_set_past_solution(fe_hip_ctx *ctx, int j, const double *u)
{
  if(!ctx || !u) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, j >= 1 && j <= fe_hip_ctx::kMaxPast, "past solution index must be in [1, 4]");
  if(fePastReady(ctx)) return FE_STATUS_ERROR;
  FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_past[j - 1], u, sizeof(double) * ctx->nDofs, hipMemcpyHostToDevice,
                                 ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_get_past_solution(fe_hip_ctx *ctx, int j, double *u)
{
  if(!ctx || !u) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, j >= 1 && j <= fe_hip_ctx::kMaxPast, "past solution index must be in [1, 4]");
  if(fePastReady(ctx)) return FE_STATUS_ERROR;
  FE_HIP_TRY(ctx, hipMemcpyAsync(u, ctx->d_past[j - 1], sizeof(double) * ctx->nDofs, hipMemcpyDeviceToHost,
                                 ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_set_essential(fe_hip_ctx *ctx, const double *values)
{
  if(!ctx || !values) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->d_u, "set the sizes first");
  const int ne = ctx->nDofs - ctx->n;
  if(ne > 0) {
    FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_u + ctx->n, values, sizeof(double) * ne, hipMemcpyHostToDevice,
                                   ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return FE_STATUS_OK;
}

int fe_hip_get_system(fe_hip_ctx *ctx, double *val, double *rhs)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(val) FE_HIP_TRY(ctx, hipMemcpyAsync(val, ctx->d_val, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
  if(rhs) FE_HIP_TRY(ctx, hipMemcpyAsync(rhs, ctx->d_rhs, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_get_system32(fe_hip_ctx *ctx, float *val32)
{
  if(!ctx || !val32) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  FE_REQUIRE(ctx, ctx->d_val32, "no FP32 values: assemble the matrix with FE_PREC_FP32_IR first");
  FE_HIP_TRY(ctx, hipMemcpyAsync(val32, ctx->d_val32, sizeof(float) * ctx->nnz, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_set_system(fe_hip_ctx *ctx, const double *val, const double *rhs)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  if(val) FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_val, val, sizeof(double) * ctx->nnz, hipMemcpyHostToDevice, ctx->stream));
  if(val) ctx->matrixZeroed = false;  // the next assembly adds to these values
  if(rhs) FE_HIP_TRY(ctx, hipMemcpyAsync(ctx->d_rhs, rhs, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_set_tuning(fe_hip_ctx *ctx, const char *key, int value)
{
  if(!ctx || !key) return FE_STATUS_POINTER_ERROR;
  const std::string k(key);
  if(k == "spmv_lanes") {
    FE_REQUIRE(ctx, value == 1 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32,
               "spmv_lanes must be 1, 2, 4, 8, 16 or 32");
    ctx->spmvLanes = value;
    ctx->lanesFixed = true;
  } else if(k == "spmv_unroll") {
    FE_REQUIRE(ctx, value == 1 || value == 2, "spmv_unroll must be 1 or 2");
    ctx->spmvUnroll = value;
  } else if(k == "iter_block") {
    FE_REQUIRE(ctx, value == 256 || value == 512 || value == 1024, "iter_block must be 256, 512 or 1024");
    ctx->iterBlock = value;
  } else if(k == "scalar_lanes") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "scalar_lanes must be 0 or 1");
    ctx->scalarLanes = value;
  } else if(k == "asm_kernel") {
    FE_REQUIRE(ctx, value >= 0 && value <= 2, "asm_kernel must be 0 (colored scalar), 1 (colored MFMA) or 2 (gather)");
    ctx->asmKernel = value;
  } else if(k == "amg_kernel") {
    FE_REQUIRE(ctx, value == 1 || value == 2, "amg_kernel must be 1 (per-level kernels) or 2 (nested-ordered PCG)");
    ctx->amgKernel = value;
  } else if(k == "amg_pre_rows") {
    FE_REQUIRE(ctx, value >= 0, "amg_pre_rows must be >= 0 (0: no row passes)");
    ctx->amgPreRows = value;
  } else if(k == "batch_skip") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "batch_skip must be 0 (every tile) or 1 (skip stopped problems)");
    ctx->batchSkip = value;
  } else if(k == "batch_graph") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "batch_graph must be 0 (direct launches) or 1 (captured graph)");
    ctx->batchGraph = value;
  } else if(k == "ns_mfma") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "ns_mfma must be 0 (scalar sweep) or 1 (matrix cores)");
    ctx->nsMfma = value;
  } else if(k == "amg_fuse") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "amg_fuse must be 0 (one launch per level) or 1 (levels 1 .. coarsest in one launch)");
    ctx->amgFuse = value;
  } else if(k == "amg_wcycle") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "amg_wcycle must be 0 (V-cycle) or 1 (W-cycle kernel on the coarse levels)");
    ctx->amgWcycle = value;
  } else if(k == "gather_xcd") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "gather_xcd must be 0 or 1");
    ctx->gatherXcd = value;
  } else if(k == "asm_fp32_math") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "asm_fp32_math must be 0 (rounded FP64 rows) or 1 (FP32 rows)");
    ctx->asmFp32Math = value;
  } else if(k == "fp32_values") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "fp32_values must be 0 (rounded FP64) or 1 (FP32 MFMA)");
    ctx->fp32Values = value;
  } else if(k == "pipe_kernel") {
    FE_REQUIRE(ctx, value >= 0 && value <= 2,
               "pipe_kernel must be 0 (Chronopoulos-Gear), 1 (tiled Ghysels-Vanroose) or 2 (by precision)");
    ctx->pipeKernel = value;
  } else if(k == "pc_values32") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "pc_values32 must be 0 or 1");
    ctx->pcValues32 = value;
  } else if(k == "pc_fp32") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "pc_fp32 must be 0 or 1");
    ctx->pcFp32 = value;
  } else if(k == "gmres_orth") {
    FE_REQUIRE(ctx, value == 1 || value == 2, "gmres_orth must be 1 (classical Gram-Schmidt) or 2 (CGS2)");
    ctx->gmresOrth = value;
  } else if(k == "ilu_sweeps") {
    FE_REQUIRE(ctx, value >= 0 && value <= 64,
               "ilu_sweeps must be 0 (exact triangular solves) or a sweep count in [1, 64]");
    ctx->pc.sweeps = value;
  } else if(k == "ic0_sweeps") {
    FE_REQUIRE(ctx, value >= 0 && value <= 64, "ic0_sweeps must be 0 (exact ILU(0)) or 1..64");
    ctx->icSweeps = value;
  } else if(k == "amg_omega_x100") {
    FE_REQUIRE(ctx, value > 0 && value < 200, "amg_omega_x100 must be in (0, 200)");
    ctx->amgOmega = value / 100.;
  } else if(k == "amg_oc_x100") {
    FE_REQUIRE(ctx, value > 0 && value < 200, "amg_oc_x100 must be in (0, 200)");
    ctx->amgOverCorrect = value / 100.;
  } else if(k == "xcd_tiles") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "xcd_tiles must be 0 or 1");
    ctx->xcdTiles = value;
  } else if(k == "graph_iters") {
    FE_REQUIRE(ctx, value >= 0 && value <= 32, "graph_iters must be in [0, 32]");
    ctx->graphIters = value;
  } else if(k == "batch_amg_global") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "batch_amg_global must be 0 or 1");
    ctx->batchAmgGlobal = value;
  } else if(k == "amg_pad_tiles") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "amg_pad_tiles must be 0 or 1");
    ctx->amgPadTiles = value;
  } else if(k == "amg_lanes") {
    FE_REQUIRE(ctx, value == 2 || value == 4 || value == 8, "amg_lanes must be 2, 4 or 8");
    ctx->amgLanes = value;
  } else if(k == "amg_xcd") {
    FE_REQUIRE(ctx, value == 0 || value == 1, "amg_xcd must be 0 or 1");
    ctx->amgXcd = value;
  } else if(k == "amg_block_rows") {
    FE_REQUIRE(ctx, value >= 0 && value <= (1 << 20), "amg_block_rows must be in [0, 2^20]");
    ctx->amgBlockRows = value;
  } else if(k == "iter_grid") {
    FE_REQUIRE(ctx, value >= 8 && value <= 4096 && value % 8 == 0, "iter_grid must be a multiple of 8 in [8, 4096]");
    ctx->iterGrid = value;
  } else {
    FE_REQUIRE(ctx, false, "unknown tuning key " + k);
  }
  return FE_STATUS_OK;
}

int fe_hip_update_coeff(fe_hip_ctx *ctx, int form, int index, const double *data, int64_t count)
{
  if(!ctx || !data) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, form >= 0 && form < (int)ctx->forms.size(), "unknown form");
  FeForm &F = ctx->forms[form];
  FE_REQUIRE(ctx, index >= 0 && index < (int)F.coeffs.size(), "unknown coefficient");
  FeFn &fn = F.coeffs[index];
  if(fn.kind == FE_FN_POLY_U) {
    FE_REQUIRE(ctx, count == fn.n, "a polynomial coefficient keeps its number of coefficients");
    FE_HIP_TRY(ctx, hipMemcpyAsync(fn.data, data, sizeof(double) * fn.n, hipMemcpyHostToDevice, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FE_STATUS_OK;
  }
  if(fn.kind != FE_FN_ELEM && fn.kind != FE_FN_QP) {
    FE_REQUIRE(ctx, count == 1, "an analytic coefficient takes one value (its c0)");
    fn.c0 = data[0];
    return FE_STATUS_OK;
  }
  const int64_t need = (fn.kind == FE_FN_ELEM) ? ctx->E : (int64_t)ctx->E * F.tab.nq;
  FE_REQUIRE(ctx, count == need, "coefficient table has the wrong size");
  FE_HIP_TRY(ctx, hipMemcpyAsync(fn.data, data, sizeof(double) * need, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_get_tuning(fe_hip_ctx *ctx, const char *key, int *value)
{
  if(!ctx || !key || !value) return FE_STATUS_POINTER_ERROR;
  const std::string k(key);
  if(k == "spmv_lanes") *value = ctx->spmvLanes;
  else if(k == "spmv_unroll") *value = ctx->spmvUnroll;
  else if(k == "iter_block") *value = ctx->iterBlock;
  else if(k == "iter_grid") *value = ctx->iterGrid;
  else if(k == "asm_kernel") *value = ctx->asmKernel;
  else if(k == "gather_xcd") *value = ctx->gatherXcd;
  else if(k == "amg_kernel") *value = ctx->amgKernel;
  else if(k == "amg_wcycle") *value = ctx->amgWcycle;
  else if(k == "amg_fuse") *value = ctx->amgFuse;
  else if(k == "ns_mfma") *value = ctx->nsMfma;
  else if(k == "batch_graph") *value = ctx->batchGraph;
  else if(k == "batch_skip") *value = ctx->batchSkip;
  else if(k == "amg_pre_rows") *value = ctx->amgPreRows;
  else if(k == "amg_fuse_grid") *value = feAmgPcgFuseGrid(ctx);
  else if(k == "amg_wcycle_level") *value = feAmgPcgWLevel(ctx);
  else if(k == "amg_wcycle_dense") *value = feAmgPcgWDense(ctx);
  else if(k == "asm_fp32_math") *value = ctx->asmFp32Math;
  else if(k == "scalar_lanes") *value = ctx->scalarLanes;
  else if(k == "fp32_values") *value = ctx->fp32Values;
  else if(k == "pipe_kernel") *value = ctx->pipeKernel;
  else if(k == "ic0_sweeps") *value = ctx->icSweeps;
  else if(k == "ilu_sweeps") *value = ctx->pc.sweeps;
  else if(k == "gmres_orth") *value = ctx->gmresOrth;
  else if(k == "pc_fp32") *value = ctx->pcFp32;
  else if(k == "pc_values32") *value = ctx->pcValues32;
  else if(k == "amg_omega_x100") *value = (int)std::lround(ctx->amgOmega * 100.);
  else if(k == "amg_oc_x100") *value = (int)std::lround(ctx->amgOverCorrect * 100.);
  else if(k == "amg_block_rows") *value = ctx->amgBlockRows;
  else if(k == "amg_xcd") *value = ctx->amgXcd;
  else if(k == "xcd_tiles") *value = ctx->xcdTiles;
  else if(k == "amg_lanes") *value = ctx->amgLanes;
  else if(k == "amg_pad_tiles") *value = ctx->amgPadTiles;
  else if(k == "batch_amg_global") *value = ctx->batchAmgGlobal;
  else if(k == "graph_iters") *value = ctx->graphIters;
  else FE_REQUIRE(ctx, false, "unknown tuning key " + k);
  return FE_STATUS_OK;
}

extern "C++" {
int feTimerMark(fe_hip_ctx *ctx, FeKernelTimer &t)
{
  if(t.used + 1 > t.ev.size())
    for(int q = 0; q < 64; ++q) {
      hipEvent_t e;
      FE_HIP_TRY(ctx, hipEventCreate(&e));
      t.ev.push_back(e);
    }
  FE_HIP_TRY(ctx, hipEventRecord(t.ev[t.used++], ctx->stream));
  return FE_STATUS_OK;
}

int feTimerCollect(fe_hip_ctx *ctx, FeKernelTimer &t)
{
  size_t q = 0;
  for(; q + 1 < t.used; q += 2) {
    float ms = 0.f;
    FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, t.ev[q], t.ev[q + 1]));
    t.total_ms += ms;
    t.launches++;
  }
  if(q < t.used) std::swap(t.ev[0], t.ev[q]);
  t.used -= q;
  return FE_STATUS_OK;
}
}  // extern "C++"

int fe_hip_set_profiling(fe_hip_ctx *ctx, int on)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  ctx->profiling = on != 0;
  return FE_STATUS_OK;
}

int fe_hip_kernel_stats(fe_hip_ctx *ctx, int which, double *avg_ms, int64_t *launches)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  if(which == 0) {
    if(avg_ms) *avg_ms = ctx->spmvTimer.launches ? ctx->spmvTimer.total_ms / ctx->spmvTimer.launches : 0.;
    if(launches) *launches = ctx->spmvTimer.launches;
  } else if(which == 2) {
    if(avg_ms) *avg_ms = ctx->iterTimer.launches ? ctx->iterTimer.total_ms / ctx->iterTimer.launches : 0.;
    if(launches) *launches = ctx->iterTimer.launches;
  } else if(which == 1) {
    if(ctx->asmPending) {
      float ms = 0.f;
      FE_HIP_TRY(ctx, hipEventSynchronize(ctx->asmEv[1]));
      FE_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->asmEv[0], ctx->asmEv[1]));
      ctx->lastAsmMs = ms;
      ctx->asmPending = false;
    }
    if(avg_ms) *avg_ms = ctx->lastAsmMs;
    if(launches) *launches = 1;
  } else {
    return FE_STATUS_RANGE_ERROR;
  }
  return FE_STATUS_OK;
}

void fe_hip_default_opts(fe_hip_solver_opts *o)
{
  if(!o) return;
  o->ksp = FE_KSP_CG;
  o->pc = FE_PC_JACOBI;
  o->precision = FE_PREC_FP64;
  o->restart = 50;
  o->max_iter = 10000;
  o->rtol = 1e-8;
  o->atol = 1e-14;
  o->dtol = 1e6;
  o->inner_rtol = 1e-3;
  o->check_every = 32;
}

int fe_hip_solve(fe_hip_ctx *ctx, const fe_hip_solver_opts *opts, double *du, fe_hip_solve_stats *stats)
{
  if(!ctx || !opts) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR first");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return feSolve(ctx, opts, du, stats);
}

}  // extern "C"
