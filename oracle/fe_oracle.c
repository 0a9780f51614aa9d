/* CPU oracle for the feNG assemble-and-solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm, used by tests/ (as the
 * parity checker), by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The product never links or calls it.
 *
 * Parity pins (tests/test_oracle_*.py): every function below is checked
 * against vectors dumped from the reference itself (oracle/_ref, fixtures in
 * tests/golden/) and/or against closed forms:
 *   - or_tri_edges, or_number, or_csr: bit-exact vs reference fixtures.
 *   - or_quad_*: exact integration of monomials (quadratureTri.cpp closed form).
 *   - or_assemble_diffusion: <=1e-12 vs reference fixtures; P1 textbook stiffness.
 *   - or_assemble_ns: <=1e-12 vs reference fixtures (Taylor-Hood, Kovasznay).
 *   - solvers: residual ||b - A x|| checked directly, CG vs dense solve on
 *     tiny systems; IC(0)/ILU(0) reproduce the exact factor on a tridiagonal
 *     matrix (no fill), where they equal the complete Cholesky / LU factor.
 *   - or_l2_error_sinsin: reference convergence table scalarFE_diffusion.output.
 *   - or_element_mass / or_assemble_transient_mass: closed-form P1 mass
 *     matrices (|T|/12 (1 + d_ij) triangles, |T|/20 (1 + d_ij) tets), the P2
 *     row sums (0 at vertices, |T|/3 at edges), and, through the BDF oracle
 *     (fe_oracle_time.py), the reference's scalarFEunsteady_linearODE tables.
 */
#include "fe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Edges: feMeshRead.cpp:1412-1456.  An edge is keyed on its (min, max) vertex
 * pair; the tag counter starts at 1 and increments on each new insertion; the
 * element stores +tag for the first encounter and -tag afterwards.  Edge k of a
 * triangle joins local vertices (k, k+1 mod 3). */
typedef struct {
  int64_t *keys;
  int *vals;
  int64_t cap;
} ihash;

static void ih_init(ihash *h, int64_t n)
{
  int64_t cap = 16;
  while(cap < 2 * n) cap <<= 1;
  h->cap = cap;
  h->keys = (int64_t *)malloc(sizeof(int64_t) * cap);
  h->vals = (int *)malloc(sizeof(int) * cap);
  for(int64_t i = 0; i < cap; ++i) h->keys[i] = -1;
}
static void ih_free(ihash *h)
{
  free(h->keys);
  free(h->vals);
}
/* returns pointer to value slot; *found set when key existed */
static int *ih_get(ihash *h, int64_t key, int *found)
{
  uint64_t x = (uint64_t)key * 0x9E3779B97F4A7C15ull;
  int64_t i = (int64_t)(x >> 17) & (h->cap - 1);
  while(h->keys[i] != -1 && h->keys[i] != key) i = (i + 1) & (h->cap - 1);
  *found = (h->keys[i] == key);
  if(!*found) h->keys[i] = key;
  return &h->vals[i];
}

int or_tri_edges(int E, const int *conn, int *edges)
{
  ihash h;
  ih_init(&h, (int64_t)E * 2 + 16);
  int tag = 1;
  for(int e = 0; e < E; ++e) {
    for(int k = 0; k < 3; ++k) {
      int a = conn[3 * e + k], b = conn[3 * e + (k + 1) % 3];
      int lo = a < b ? a : b, hi = a < b ? b : a;
      int found;
      int *slot = ih_get(&h, ((int64_t)lo << 32) | (int64_t)hi, &found);
      if(!found) {
        *slot = tag;
        edges[3 * e + k] = tag++;
      } else {
        edges[3 * e + k] = -*slot;
      }
    }
  }
  ih_free(&h);
  return tag - 1;
}

/* ------------------------------------------------------------------------ */
/* Coloring: feCncGeo.cpp:735-794 (algorithm 1).  Each pass walks the
 * uncolored elements in order; an element still unmarked when visited joins
 * the current color and marks its uncolored vertex-neighbours for later
 * passes. */
int or_color_elements(int E, int nvpe, const int *conn, int V, int *colorOf)
{
  int *cnt = (int *)calloc((size_t)V + 1, sizeof(int));
  for(int i = 0; i < E * nvpe; ++i) cnt[conn[i] + 1]++;
  for(int v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
  int *lst = (int *)malloc(sizeof(int) * (size_t)E * nvpe);
  int *pos = (int *)malloc(sizeof(int) * (size_t)V);
  memcpy(pos, cnt, sizeof(int) * V);
  for(int e = 0; e < E; ++e)
    for(int j = 0; j < nvpe; ++j) lst[pos[conn[e * nvpe + j]]++] = e;
  for(int e = 0; e < E; ++e) colorOf[e] = -1;
  int ncolor = 0, remaining = E;
  while(remaining > 0) {
    for(int k = 0; k < E; ++k) {
      if(colorOf[k] != -1) continue;
      for(int i = 0; i < nvpe; ++i) {
        int s = conn[k * nvpe + i];
        for(int j = cnt[s]; j < cnt[s + 1]; ++j) {
          int m = lst[j];
          if(m != k && colorOf[m] < 0) colorOf[m] = -2;
        }
      }
    }
    for(int k = 0; k < E; ++k) {
      if(colorOf[k] == -1) {
        colorOf[k] = ncolor;
        remaining--;
      } else if(colorOf[k] == -2) {
        colorOf[k] = -1;
      }
    }
    ncolor++;
  }
  free(cnt);
  free(lst);
  free(pos);
  return ncolor;
}

/* ------------------------------------------------------------------------ */
/* Numbering: feNumber.cpp:370-483 (numberUnknowns / numberEssential) driven by
 * feMetaNumber (feNumber.cpp:677-843).  Per field, in order: unknown vertex
 * DOFs (vertex ascending, components interleaved), unknown edge DOFs (edge
 * index ascending); then, after all fields' unknowns, per field: essential
 * vertex DOFs, essential boundary-segment DOFs in boundary element order
 * (they are the 1D "element" DOFs of the boundary space, copied to their edge,
 * feNumber.cpp:452-466), then any remaining essential edge DOFs. */
static int kind_vdofs(int kind) { return kind == OR_TRI_VP2 ? 2 : 1; }
static int kind_edofs(int kind) { return kind == OR_TRI_P2 ? 1 : (kind == OR_TRI_VP2 ? 2 : 0); }
static int kind_nloc(int kind)
{
  switch(kind) {
    case OR_TRI_P1: return 3;
    case OR_TRI_P2: return 6;
    case OR_TRI_VP2: return 12;
    case OR_TET_P1: return 4;
  }
  return 0;
}

int or_number(int V, int E, int nvpe, const int *conn, int nEdges, const int *edges,
              int nB, int bnv, const int *bconn, const int *bedge,
              int nfields, const int *kinds, const int *essential,
              int **adr_out, int *nInc, int *nDofs)
{
  const int UNK = -1, ESS = -2;
  int **vnum = (int **)malloc(sizeof(int *) * nfields);
  int **enum_ = (int **)malloc(sizeof(int *) * nfields);
  for(int f = 0; f < nfields; ++f) {
    int nv = kind_vdofs(kinds[f]), ne = kind_edofs(kinds[f]);
    vnum[f] = (int *)malloc(sizeof(int) * (size_t)V * nv);
    for(int i = 0; i < V * nv; ++i) vnum[f][i] = UNK; /* every vertex carries the field */
    enum_[f] = NULL;
    if(ne > 0) {
      enum_[f] = (int *)malloc(sizeof(int) * (size_t)nEdges * ne);
      for(int i = 0; i < nEdges * ne; ++i) enum_[f][i] = UNK;
    }
    if(essential[f]) {
      for(int b = 0; b < nB; ++b) {
        for(int j = 0; j < bnv; ++j) {
          int v = bconn[b * bnv + j];
          for(int c = 0; c < nv; ++c) vnum[f][v * nv + c] = ESS;
        }
        if(ne > 0)
          for(int c = 0; c < ne; ++c) enum_[f][bedge[b] * ne + c] = ESS;
      }
    }
  }
  (void)conn;
  (void)nvpe;
  (void)E;
  int g = 0;
  for(int f = 0; f < nfields; ++f) {
    int nv = kind_vdofs(kinds[f]), ne = kind_edofs(kinds[f]);
    for(int i = 0; i < V * nv; ++i)
      if(vnum[f][i] == UNK) vnum[f][i] = g++;
    for(int i = 0; i < nEdges * ne; ++i)
      if(enum_[f][i] == UNK) enum_[f][i] = g++;
  }
  *nInc = g;
  for(int f = 0; f < nfields; ++f) {
    int nv = kind_vdofs(kinds[f]), ne = kind_edofs(kinds[f]);
    for(int i = 0; i < V * nv; ++i)
      if(vnum[f][i] == ESS) vnum[f][i] = g++;
    if(ne > 0) {
      for(int b = 0; b < nB; ++b)
        for(int c = 0; c < ne; ++c)
          if(enum_[f][bedge[b] * ne + c] == ESS) enum_[f][bedge[b] * ne + c] = g++;
      for(int i = 0; i < nEdges * ne; ++i)
        if(enum_[f][i] == ESS) enum_[f][i] = g++;
    }
  }
  *nDofs = g;
  /* Addressing vectors: feSpace_2D.cpp:154-159 (P1), :772-786 (P2),
   * :1017-1041 (vector P2: v0x v0y v1x v1y v2x v2y e0x e0y e1x e1y e2x e2y),
   * feSpace_3D.cpp:293-345 (tet). */
  for(int f = 0; f < nfields; ++f) {
    int kind = kinds[f], nl = kind_nloc(kind);
    int *adr = adr_out[f];
    for(int e = 0; e < E; ++e) {
      int *a = adr + (size_t)e * nl;
      if(kind == OR_TRI_P1 || kind == OR_TET_P1) {
        for(int j = 0; j < nl; ++j) a[j] = vnum[f][conn[e * nvpe + j]];
      } else if(kind == OR_TRI_P2) {
        for(int j = 0; j < 3; ++j) a[j] = vnum[f][conn[e * 3 + j]];
        for(int k = 0; k < 3; ++k) a[3 + k] = enum_[f][abs(edges[e * 3 + k]) - 1];
      } else {
        for(int j = 0; j < 3; ++j)
          for(int c = 0; c < 2; ++c) a[2 * j + c] = vnum[f][conn[e * 3 + j] * 2 + c];
        for(int k = 0; k < 3; ++k)
          for(int c = 0; c < 2; ++c) a[6 + 2 * k + c] = enum_[f][(abs(edges[e * 3 + k]) - 1) * 2 + c];
      }
    }
  }
  for(int f = 0; f < nfields; ++f) {
    free(vnum[f]);
    free(enum_[f]);
  }
  free(vnum);
  free(enum_);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CSR: feCompressedRowStorage.cpp:15-133.  Row i holds i itself plus every
 * adrJ[j] < n that some element couples to a row adrI[i] = i (i < n); sorted,
 * unique, 0-based. */
static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int64_t or_csr(int n, int nforms, const int *E, const int *nI, const int *nJ,
               const int *const *adrI, const int *const *adrJ, int *ia, int *ja)
{
  int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  for(int i = 0; i < n; ++i) cnt[i + 1] = 1;
  for(int f = 0; f < nforms; ++f)
    for(int e = 0; e < E[f]; ++e)
      for(int i = 0; i < nI[f]; ++i) {
        int r = adrI[f][(size_t)e * nI[f] + i];
        if(r >= n) continue;
        for(int j = 0; j < nJ[f]; ++j)
          if(adrJ[f][(size_t)e * nJ[f] + j] < n) cnt[r + 1]++;
      }
  for(int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  int *buf = (int *)malloc(sizeof(int) * (size_t)(cnt[n] > 0 ? cnt[n] : 1));
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for(int i = 0; i < n; ++i) {
    pos[i] = cnt[i];
    buf[pos[i]++] = i;
  }
  for(int f = 0; f < nforms; ++f)
    for(int e = 0; e < E[f]; ++e)
      for(int i = 0; i < nI[f]; ++i) {
        int r = adrI[f][(size_t)e * nI[f] + i];
        if(r >= n) continue;
        for(int j = 0; j < nJ[f]; ++j) {
          int c = adrJ[f][(size_t)e * nJ[f] + j];
          if(c < n) buf[pos[r]++] = c;
        }
      }
  int64_t nnz = 0;
  for(int i = 0; i < n; ++i) {
    int64_t b = cnt[i], len = cnt[i + 1] - cnt[i];
    qsort(buf + b, (size_t)len, sizeof(int), cmp_int);
    int64_t u = 0;
    for(int64_t k = 0; k < len; ++k)
      if(k == 0 || buf[b + k] != buf[b + k - 1]) buf[b + u++] = buf[b + k];
    if(ia) ia[i] = (int)nnz;
    if(ja)
      for(int64_t k = 0; k < u; ++k) ja[nnz + k] = buf[b + k];
    nnz += u;
  }
  if(ia) ia[n] = (int)nnz;
  free(cnt);
  free(buf);
  free(pos);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Quadrature tables (data, as in feQuadratureTri.cpp:788-920 -- MFEM's
 * symmetric triangle rules -- and the Solin tet tables of
 * feQuadratureTet.cpp:96-110). */
static int tri_pt(double *r, double *s, double *w, int off, double a, double b, double wt)
{
  r[off] = a;
  s[off] = b;
  w[off] = wt;
  return off + 1;
}
static int tri_p3(double *r, double *s, double *w, int off, double a, double wt)
{
  double b = 1. - 2. * a;
  off = tri_pt(r, s, w, off, a, a, wt);
  off = tri_pt(r, s, w, off, a, b, wt);
  return tri_pt(r, s, w, off, b, a, wt);
}
/* the orbit of (a, a, b) given b, a = (1 - b) / 2 (AddTriPoints3b) */
static int tri_p3b(double *r, double *s, double *w, int off, double b, double wt)
{
  double a = (1. - b) / 2.;
  off = tri_pt(r, s, w, off, a, a, wt);
  off = tri_pt(r, s, w, off, a, b, wt);
  return tri_pt(r, s, w, off, b, a, wt);
}
static int tri_p3r(double *r, double *s, double *w, int off, double a, double b, double c,
                   double wt)
{
  off = tri_pt(r, s, w, off, a, b, wt);
  off = tri_pt(r, s, w, off, c, a, wt);
  return tri_pt(r, s, w, off, b, c, wt);
}
static int tri_p6(double *r, double *s, double *w, int off, double a, double b, double wt)
{
  double c = 1. - a - b;
  off = tri_pt(r, s, w, off, a, b, wt);
  off = tri_pt(r, s, w, off, b, a, wt);
  off = tri_pt(r, s, w, off, a, c, wt);
  off = tri_pt(r, s, w, off, c, a, wt);
  off = tri_pt(r, s, w, off, b, c, wt);
  return tri_pt(r, s, w, off, c, b, wt);
}

int or_quad_tri(int degree, double *r, double *s, double *w)
{
  int o = 0;
  switch(degree) {
    case 0:
    case 1: o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., 0.5); break;
    case 2: o = tri_p3(r, s, w, 0, 1. / 6., 1. / 6.); break;
    case 3:
      o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., -0.28125);
      o = tri_p3(r, s, w, o, 0.2, 25. / 96.);
      break;
    case 4:
      o = tri_p3(r, s, w, 0, 0.091576213509770743460, 0.054975871827660933819);
      o = tri_p3(r, s, w, o, 0.44594849091596488632, 0.11169079483900573285);
      break;
    case 5:
      o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., 0.1125);
      o = tri_p3(r, s, w, o, 0.10128650732345633880, 0.062969590272413576298);
      o = tri_p3(r, s, w, o, 0.47014206410511508977, 0.066197076394253090369);
      break;
    case 6:
      o = tri_p3(r, s, w, 0, 0.063089014491502228340, 0.025422453185103408460);
      o = tri_p3(r, s, w, o, 0.24928674517091042129, 0.058393137863189683013);
      o = tri_p6(r, s, w, o, 0.053145049844816947353, 0.31035245103378440542,
                 0.041425537809186787597);
      break;
    case 7:
      o = tri_p3r(r, s, w, 0, 0.062382265094402118174, 0.067517867073916085443,
                  1. - 0.062382265094402118174 - 0.067517867073916085443,
                  0.026517028157436251429);
      o = tri_p3r(r, s, w, o, 0.055225456656926611737, 0.32150249385198182267,
                  1. - 0.055225456656926611737 - 0.32150249385198182267,
                  0.043881408714446055037);
      o = tri_p3r(r, s, w, o, 0.034324302945097146470, 0.66094919618673565761,
                  0.30472650086816719592, 0.028775042784981585738);
      o = tri_p3r(r, s, w, o, 0.51584233435359177926, 0.27771616697639178257,
                  0.20644149867001643817, 0.067493187009802774463);
      break;
    case 8:
      o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., 0.0721578038388935841255455552445323);
      o = tri_p3(r, s, w, o, 0.170569307751760206622293501491464,
                 0.0516086852673591251408957751460645);
      o = tri_p3(r, s, w, o, 0.0505472283170309754584235505965989,
                 0.0162292488115990401554629641708902);
      o = tri_p3(r, s, w, o, 0.459292588292723156028815514494169,
                 0.0475458171336423123969480521942921);
      o = tri_p6(r, s, w, o, 0.008394777409957605337213834539296,
                 0.263112829634638113421785786284643, 0.0136151570872174971324223450369544);
      break;
    case 9:
      o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., 0.0485678981413994169096209912536443);
      o = tri_p3b(r, s, w, o, 0.020634961602524744433, 0.0156673501135695352684274156436046);
      o = tri_p3b(r, s, w, o, 0.12582081701412672546, 0.0389137705023871396583696781497019);
      o = tri_p3(r, s, w, o, 0.188203535619032730240961280467335, 0.0398238694636051265164458871320226);
      o = tri_p3(r, s, w, o, 0.0447295133944527098651065899662763, 0.0127888378293490156308393992794999);
      o = tri_p6(r, s, w, o, 0.0368384120547362836348175987833851, 0.2219629891607656956751025276931919,
                 0.0216417696886446886446886446886446);
      break;
    case 10:
      o = tri_pt(r, s, w, 0, 1. / 3., 1. / 3., 0.0454089951913767900476432975500142);
      o = tri_p3b(r, s, w, o, 0.028844733232685245264984935583748, 0.0183629788782333523585030359456832);
      o = tri_p3(r, s, w, o, 0.109481575485037054795458631340522, 0.0226605297177639673913028223692986);
      o = tri_p6(r, s, w, o, 0.141707219414879954756683250476361, 0.307939838764120950165155022930631,
                 0.0363789584227100543021575883096803);
      o = tri_p6(r, s, w, o, 0.025003534762686386073988481007746, 0.246672560639902693917276465411176,
                 0.0141636212655287424183685307910495);
      o = tri_p6(r, s, w, o, 0.0095408154002994575801528096228873, 0.0668032510122002657735402127620247,
                 4.71083348186641172996373548344341E-03);
      break;
    case 11:
      o = tri_p6(r, s, w, 0, 0.0, 0.141129718717363295960826061941652, 3.68119189165027713212944752369032E-03);
      o = tri_pt(r, s, w, o, 1. / 3., 1. / 3., 0.0439886505811161193990465846607278);
      o = tri_p3(r, s, w, o, 0.0259891409282873952600324854988407, 4.37215577686801152475821439991262E-03);
      o = tri_p3(r, s, w, o, 0.0942875026479224956305697762754049, 0.0190407859969674687575121697178070);
      o = tri_p3b(r, s, w, o, 0.010726449965572372516734795387128, 9.42772402806564602923839129555767E-03);
      o = tri_p3(r, s, w, o, 0.207343382614511333452934024112966, 0.0360798487723697630620149942932315);
      o = tri_p3b(r, s, w, o, 0.122184388599015809877869236727746, 0.0346645693527679499208828254519072);
      o = tri_p6(r, s, w, o, 0.0448416775891304433090523914688007, 0.2772206675282791551488214673424523,
                 0.0205281577146442833208261574536469);
      break;
    case 12:
      o = tri_p3b(r, s, w, 0, 2.35652204523900E-02, 1.28655332202275E-02);
      o = tri_p3b(r, s, w, o, 1.20551215411079E-01, 2.18462722690190E-02);
      o = tri_p3(r, s, w, o, 2.71210385012116E-01, 3.14291121089425E-02);
      o = tri_p3(r, s, w, o, 1.27576145541586E-01, 1.73980564653545E-02);
      o = tri_p3(r, s, w, o, 2.13173504532100E-02, 3.08313052577950E-03);
      o = tri_p6(r, s, w, o, 1.15343494534698E-01, 2.75713269685514E-01, 2.01857788831905E-02);
      o = tri_p6(r, s, w, o, 2.28383322222570E-02, 2.81325580989940E-01, 1.11783866011515E-02);
      o = tri_p6(r, s, w, o, 2.57340505483300E-02, 1.16251915907597E-01, 8.65811555432950E-03);
      break;
    default: return -1;
  }
  return o;
}

int or_quad_tet(int degree, double *r, double *s, double *t, double *w)
{
  if(degree <= 1) {
    r[0] = s[0] = t[0] = 0.25;
    w[0] = 0.166666666666667;
    return 1;
  }
  if(degree == 2) {
    const double a = 0.138196601125, b = 0.585410196625;
    const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    for(int q = 0; q < 4; ++q) {
      r[q] = P[q][0];
      s[q] = P[q][1];
      t[q] = P[q][2];
      w[q] = 0.0416666666666667;
    }
    return 4;
  }
  return -1;
}

/* ------------------------------------------------------------------------ */
/* Reference-element bases (feSpace_2D.cpp:118-134, :536-553, :727-741;
 * feSpace_3D.cpp:104-242 for n = 1). */
/* Cubic and quartic Lagrange triangles on equispaced nodes, written in the
 * barycentric coordinates l = (1 - r - s, r, s) with their gradients g[i]
 * with respect to (l0, l1, l2); node order: vertices, each edge k from
 * vertex k to vertex k+1, element nodes (feSpace_2D.cpp:1315-1327,
 * :1700-1715). */
static void set3(double *g, double a, double b, double c)
{
  g[0] = a;
  g[1] = b;
  g[2] = c;
}

/* P3: vertex 1/2 l(3l-1)(3l-2); edge node nearer a: 9/2 la lb (3 la - 1);
 * centroid 27 l0 l1 l2. */
static int cubic_tri(const double *l, double *L, double (*g)[3])
{
  for(int a = 0; a < 3; ++a) {
    const double x = l[a];
    L[a] = 0.5 * x * (3. * x - 1.) * (3. * x - 2.);
    double d[3] = {0., 0., 0.};
    d[a] = 0.5 * (27. * x * x - 18. * x + 2.);
    set3(g[a], d[0], d[1], d[2]);
  }
  for(int k = 0; k < 3; ++k) {
    const int v0 = k, v1 = (k + 1) % 3;
    for(int h = 0; h < 2; ++h) {
      const int a = h == 0 ? v0 : v1, b = h == 0 ? v1 : v0; /* a: the nearer vertex */
      const int i = 3 + 2 * k + h;
      L[i] = 4.5 * l[a] * l[b] * (3. * l[a] - 1.);
      double d[3] = {0., 0., 0.};
      d[a] = 4.5 * l[b] * (6. * l[a] - 1.);
      d[b] = 4.5 * l[a] * (3. * l[a] - 1.);
      set3(g[i], d[0], d[1], d[2]);
    }
  }
  L[9] = 27. * l[0] * l[1] * l[2];
  set3(g[9], 27. * l[1] * l[2], 27. * l[0] * l[2], 27. * l[0] * l[1]);
  return 10;
}

/* P4: vertex l(4l-1)(4l-2)(4l-3)/6; edge nodes at 1/4 and 3/4,
 * 8/3 la lb (4 la - 1)(4 la - 2) (a the nearer vertex), midpoint
 * 4 la lb (4 la - 1)(4 lb - 1); element nodes 32 l0 l1 l2 (4 lm - 1) with m
 * the coordinate equal to 1/2 at the node ((1/4, 1/4): m = 0, (1/2, 1/4):
 * m = 1, (1/4, 1/2): m = 2). */
static int quartic_tri(const double *l, double *L, double (*g)[3])
{
  for(int a = 0; a < 3; ++a) {
    const double x = l[a];
    L[a] = x * (4. * x - 1.) * (4. * x - 2.) * (4. * x - 3.) / 6.;
    double d[3] = {0., 0., 0.};
    d[a] = (256. * x * x * x - 288. * x * x + 88. * x - 6.) / 6.;
    set3(g[a], d[0], d[1], d[2]);
  }
  for(int k = 0; k < 3; ++k) {
    const int v0 = k, v1 = (k + 1) % 3;
    for(int h = 0; h < 3; ++h) {
      const int i = 3 + 3 * k + h;
      double d[3] = {0., 0., 0.};
      if(h == 1) {
        const double pa = l[v0] * (4. * l[v0] - 1.), pb = l[v1] * (4. * l[v1] - 1.);
        L[i] = 4. * pa * pb;
        d[v0] = 4. * (8. * l[v0] - 1.) * pb;
        d[v1] = 4. * pa * (8. * l[v1] - 1.);
      } else {
        const int a = h == 0 ? v0 : v1, b = h == 0 ? v1 : v0;
        const double x = l[a];
        const double hx = x * (4. * x - 1.) * (4. * x - 2.), dh = 48. * x * x - 24. * x + 2.;
        L[i] = 8. / 3. * l[b] * hx;
        d[a] = 8. / 3. * l[b] * dh;
        d[b] = 8. / 3. * hx;
      }
      set3(g[i], d[0], d[1], d[2]);
    }
  }
  for(int m = 0; m < 3; ++m) {
    const int i = 12 + m, o1 = (m + 1) % 3, o2 = (m + 2) % 3;
    const double pm = l[m] * (4. * l[m] - 1.);
    L[i] = 32. * pm * l[o1] * l[o2];
    double d[3];
    d[m] = 32. * (8. * l[m] - 1.) * l[o1] * l[o2];
    d[o1] = 32. * pm * l[o2];
    d[o2] = 32. * pm * l[o1];
    set3(g[i], d[0], d[1], d[2]);
  }
  return 15;
}

static int basis_eval(int kind, double r, double s, double t, double *L, double *dr,
                      double *ds, double *dt)
{
  if(kind == OR_TRI_P1) {
    L[0] = 1. - r - s; L[1] = r; L[2] = s;
    dr[0] = -1.; dr[1] = 1.; dr[2] = 0.;
    ds[0] = -1.; ds[1] = 0.; ds[2] = 1.;
    return 3;
  }
  if(kind == OR_TRI_P2 || kind == OR_TRI_VP2) {
    L[0] = (1. - r - s) * (1. - 2. * r - 2. * s);
    L[1] = r * (2. * r - 1.);
    L[2] = s * (2. * s - 1.);
    L[3] = 4. * r * (1. - r - s);
    L[4] = 4. * r * s;
    L[5] = 4. * s * (1. - r - s);
    dr[0] = 4. * (r + s) - 3.; dr[1] = 4. * r - 1.; dr[2] = 0.;
    dr[3] = 4. * (1. - 2. * r - s); dr[4] = 4. * s; dr[5] = -4. * s;
    ds[0] = 4. * (r + s) - 3.; ds[1] = 0.; ds[2] = 4. * s - 1.;
    ds[3] = -4. * r; ds[4] = 4. * r; ds[5] = 4. * (1. - r - 2. * s);
    return 6;
  }
  if(kind == OR_TRI_P3 || kind == OR_TRI_P4) {
    const double l[3] = {1. - r - s, r, s};
    double g[15][3];
    const int nf = kind == OR_TRI_P3 ? cubic_tri(l, L, g) : quartic_tri(l, L, g);
    for(int i = 0; i < nf; ++i) {
      dr[i] = g[i][1] - g[i][0]; /* l1 = r, l0 = 1 - r - s */
      ds[i] = g[i][2] - g[i][0];
    }
    return nf;
  }
  if(kind == OR_TET_P1) {
    L[0] = 1. - r - s - t; L[1] = r; L[2] = s; L[3] = t;
    dr[0] = -1.; dr[1] = 1.; dr[2] = 0.; dr[3] = 0.;
    ds[0] = -1.; ds[1] = 0.; ds[2] = 1.; ds[3] = 0.;
    dt[0] = -1.; dt[1] = 0.; dt[2] = 0.; dt[3] = 1.;
    return 4;
  }
  return 0;
}

int or_basis(int kind, double r, double s, double t, double *L, double *dr, double *ds, double *dt)
{
  double z[15];
  return basis_eval(kind, r, s, t, L, dr, ds, dt ? dt : z);
}

/* Geometry of a straight simplex: feCncGeo.cpp:318-340 (2D detJ), :380-396
 * (3D det3x3), :651-692 (inverse Jacobian, constant for P1 geometry). */
typedef struct {
  double J;
  double drdx[3], drdy[3], drdz[3];
} or_geo;

static int geo_tri(const double *c /* 3 vertices x 3 */, or_geo *g)
{
  double dxdr[2] = {-c[0] + c[3], -c[1] + c[4]};
  double dxds[2] = {-c[0] + c[6], -c[1] + c[7]};
  double J = dxdr[0] * dxds[1] - dxdr[1] * dxds[0];
  g->J = J;
  g->drdx[0] = dxds[1] / J;
  g->drdx[1] = -dxdr[1] / J;
  g->drdy[0] = -dxds[0] / J;
  g->drdy[1] = dxdr[0] / J;
  return J > 0 ? 0 : -1;
}

static int geo_tet(const double *c /* 4 vertices x 3 */, or_geo *g)
{
  double m[3][3];
  for(int i = 0; i < 3; ++i) {
    m[i][0] = -c[i] + c[3 + i];
    m[i][1] = -c[i] + c[6 + i];
    m[i][2] = -c[i] + c[9 + i];
  }
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  g->J = det;
  double ud = 1. / det, inv[3][3];
  inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * ud;
  inv[1][0] = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]) * ud;
  inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * ud;
  inv[0][1] = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]) * ud;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * ud;
  inv[2][1] = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]) * ud;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * ud;
  inv[1][2] = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]) * ud;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * ud;
  for(int i = 0; i < 3; ++i) {
    g->drdx[i] = inv[i][0];
    g->drdy[i] = inv[i][1];
    g->drdz[i] = inv[i][2];
  }
  return det > 0 ? 0 : -1;
}

/* Coefficient callbacks evaluated at physical quadrature points
 * (feSysElm.cpp:540-542: pos interpolated from the geometry). */
static double kov_lambda(double Re) { return Re / 2. - sqrt(Re * Re / 4. + 4. * M_PI * M_PI); }

static double fn_eval(or_fn f, const double *x, int e, int nq, int q)
{
  switch(f.kind) {
    case OR_FN_CONST: return f.c0;
    case OR_FN_SINSIN: return f.c0 * sin(M_PI * x[0]) * sin(M_PI * x[1]);
    case OR_FN_BLOCKS: {
      int B = (int)f.c0;
      int bi = (int)floor(x[0] * B), bj = (int)floor(x[1] * B);
      if(bi >= B) bi = B - 1;
      if(bj >= B) bj = B - 1;
      return exp(0.5 * (double)(((7 * bi + 13 * bj) % 11) - 5));
    }
    case OR_FN_SMOOTH3:
      return exp(f.c0 * sin(2. * M_PI * x[0]) * sin(2. * M_PI * x[1]) * sin(2. * M_PI * x[2]));
    case OR_FN_ELEM: return f.data[e];
    case OR_FN_QP: return f.data[(size_t)e * nq + q];
    case OR_FN_KOV_U: {
      double l = kov_lambda(f.c0);
      if(f.c1 == 0.) return 1. - exp(l * x[0]) * cos(2. * M_PI * x[1]);
      return l / (2. * M_PI) * exp(l * x[0]) * sin(2. * M_PI * x[1]);
    }
    case OR_FN_KOV_P: {
      double l = kov_lambda(f.c0);
      return 0.5 * (1. - exp(2. * l * x[0]));
    }
  }
  return 0.;
}

/* ------------------------------------------------------------------------ */
/* Scalar diffusion + source on one element: feSysElm.cpp:530-555 (Ae, loop
 * order q, i, j, d), :557-583 (Be from grad u_h), :16-27 (source). */
typedef struct {
  int nf, nq, dim, nv;
  double qr[36], qs[36], qt[36], qw[36];
  double L[36][15], dLr[36][15], dLs[36][15], dLt[36][15];
  double Lg[36][4]; /* P1 geometry basis at the quadrature points */
} or_ref;

static int ref_setup(int kind, int quad_degree, or_ref *R)
{
  R->dim = (kind == OR_TET_P1) ? 3 : 2;
  R->nv = R->dim + 1;
  if(R->dim == 2)
    R->nq = or_quad_tri(quad_degree, R->qr, R->qs, R->qw);
  else
    R->nq = or_quad_tet(quad_degree, R->qr, R->qs, R->qt, R->qw);
  if(R->nq <= 0) return -1;
  for(int q = 0; q < R->nq; ++q) {
    double t = R->dim == 3 ? R->qt[q] : 0.;
    R->nf = basis_eval(kind == OR_TRI_VP2 ? OR_TRI_P2 : kind, R->qr[q], R->qs[q], t, R->L[q],
                       R->dLr[q], R->dLs[q], R->dLt[q]);
    double d1[4], d2[4], d3[4];
    basis_eval(R->dim == 3 ? OR_TET_P1 : OR_TRI_P1, R->qr[q], R->qs[q], t, R->Lg[q], d1, d2, d3);
  }
  return 0;
}

static int elem_geo(const or_ref *R, const double *c, or_geo *g)
{
  return R->dim == 2 ? geo_tri(c, g) : geo_tet(c, g);
}

static void qp_pos(const or_ref *R, const double *c, int q, double *x)
{
  x[0] = x[1] = x[2] = 0.;
  for(int j = 0; j < R->nv; ++j)
    for(int d = 0; d < 3; ++d) x[d] += c[3 * j + d] * R->Lg[q][j];
}

static void phys_grad(const or_ref *R, const or_geo *g, int q, double *gp /* nf*dim */)
{
  for(int i = 0; i < R->nf; ++i) {
    if(R->dim == 2) {
      gp[i * 2 + 0] = R->dLr[q][i] * g->drdx[0] + R->dLs[q][i] * g->drdx[1];
      gp[i * 2 + 1] = R->dLr[q][i] * g->drdy[0] + R->dLs[q][i] * g->drdy[1];
    } else {
      gp[i * 3 + 0] = R->dLr[q][i] * g->drdx[0] + R->dLs[q][i] * g->drdx[1] + R->dLt[q][i] * g->drdx[2];
      gp[i * 3 + 1] = R->dLr[q][i] * g->drdy[0] + R->dLs[q][i] * g->drdy[1] + R->dLt[q][i] * g->drdy[2];
      gp[i * 3 + 2] = R->dLr[q][i] * g->drdz[0] + R->dLs[q][i] * g->drdz[1] + R->dLt[q][i] * g->drdz[2];
    }
  }
}

static void diffusion_elem(const or_ref *R, const double *c, int e, or_fn coeff, or_fn source,
                           const double *ul, double *Ae, double *Be)
{
  or_geo g;
  elem_geo(R, c, &g);
  const int nf = R->nf, dim = R->dim;
  double gp[45];
  for(int k = 0; k < R->nq; ++k) {
    double x[3];
    qp_pos(R, c, k, x);
    double kc = fn_eval(coeff, x, e, R->nq, k);
    phys_grad(R, &g, k, gp);
    if(Ae)
      for(int i = 0; i < nf; ++i)
        for(int j = 0; j < nf; ++j)
          for(int d = 0; d < dim; ++d)
            Ae[i * nf + j] += kc * gp[i * dim + d] * gp[j * dim + d] * g.J * R->qw[k];
    if(Be) {
      double gu[3] = {0., 0., 0.};
      for(int i = 0; i < nf; ++i)
        for(int d = 0; d < dim; ++d) gu[d] += ul[i] * gp[i * dim + d];
      for(int i = 0; i < nf; ++i)
        for(int d = 0; d < dim; ++d) Be[i] -= kc * gp[i * dim + d] * gu[d] * g.J * R->qw[k];
    }
  }
  if(Be) {
    for(int k = 0; k < R->nq; ++k) {
      double x[3];
      qp_pos(R, c, k, x);
      double S = fn_eval(source, x, e, R->nq, k);
      for(int i = 0; i < nf; ++i) Be[i] -= R->L[k][i] * S * g.J * R->qw[k];
    }
  }
}

int or_element_diffusion(int kind, const double *coords, int quad_degree, or_fn coeff,
                         int elem, double *Ae)
{
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  memset(Ae, 0, sizeof(double) * R.nf * R.nf);
  or_fn zero = {OR_FN_CONST, 0., 0., NULL};
  diffusion_elem(&R, coords, elem, coeff, zero, NULL, Ae, NULL);
  return R.nf;
}

static int find_pos(const int *ia, const int *ja, int row, int col)
{
  for(int k = ia[row]; k < ia[row + 1]; ++k)
    if(ja[k] == col) return k;
  return -1;
}

/* Assembly: the diffusion form's matrix + residual, then the source form's
 * residual, each scattered with the Pardiso semantics
 * (feLinearSystemMklPardiso.cpp:561-660 matrix, :729-738 RHS). */
int or_assemble_diffusion(int kind, int V, const double *xyz, int E, const int *conn,
                          const int *adr, int n, const int *ia, const int *ja,
                          int quad_degree, or_fn coeff, or_fn source, const double *u,
                          const int *elem_order, double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv;
  double c[12], Ae[225], Be[15], ul[15];
  or_fn zero = {OR_FN_CONST, 0., 0., NULL};
  for(int pass = 0; pass < 2; ++pass) {
    for(int ee = 0; ee < E; ++ee) {
      int e = elem_order ? elem_order[ee] : ee;
      for(int j = 0; j < nv; ++j)
        for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
      const int *a = adr + (size_t)e * nf;
      for(int i = 0; i < nf; ++i) ul[i] = u[a[i]];
      memset(Ae, 0, sizeof(Ae));
      memset(Be, 0, sizeof(Be));
      or_geo g;
      if(elem_geo(&R, c, &g)) return -2 - e;
      if(pass == 0)
        diffusion_elem(&R, c, e, coeff, zero, ul, Ae, Be);
      else {
        or_fn nocoef = {OR_FN_CONST, 0., 0., NULL};
        diffusion_elem(&R, c, e, nocoef, source, ul, NULL, Be);
      }
      for(int i = 0; i < nf; ++i) {
        if(a[i] >= n) continue;
        rhs[a[i]] += Be[i];
        if(pass == 0)
          for(int j = 0; j < nf; ++j) {
            if(a[j] >= n) continue;
            int p = find_pos(ia, ja, a[i], a[j]);
            if(p < 0) return -1;
            val[p] += Ae[i * nf + j];
          }
      }
    }
  }
  return 0;
}

/* Transient mass on one element, feSysElm.cpp:426-458: Ae_ij += phi_i m c0
 * phi_j J w (loop order q, i, j); Be_i -= phi_i m udot_h J w with udot_h the
 * interpolation of the nodal time derivative. */
static void transient_mass_elem(const or_ref *R, const double *c, int e, or_fn coeff, double c0,
                                const double *udl, double *Ae, double *Be)
{
  or_geo g;
  elem_geo(R, c, &g);
  const int nf = R->nf;
  for(int k = 0; k < R->nq; ++k) {
    double x[3];
    qp_pos(R, c, k, x);
    const double m = fn_eval(coeff, x, e, R->nq, k);
    if(Ae)
      for(int i = 0; i < nf; ++i)
        for(int j = 0; j < nf; ++j) Ae[i * nf + j] += R->L[k][i] * m * c0 * R->L[k][j] * g.J * R->qw[k];
    if(Be) {
      double ud = 0.;
      for(int j = 0; j < nf; ++j) ud += R->L[k][j] * udl[j];
      for(int i = 0; i < nf; ++i) Be[i] -= R->L[k][i] * m * ud * g.J * R->qw[k];
    }
  }
}

/* Mass (feSysElm.cpp:186-225, p = 1) and mass power (feSysElm.cpp:260-278,
 * Be_i -= c u_h^p phi_i J w) on a whole mesh; the matrix is the derivative
 * c p u_h^(p-1) phi_i phi_j J w (the reference takes it by finite
 * differences of the residual for the power form).  val may be NULL. */
int or_assemble_mass_power(int kind, int V, const double *xyz, int E, const int *conn, const int *adr, int n,
                           const int *ia, const int *ja, int quad_degree, or_fn coeff, double p, const double *u,
                           double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv;
  double c[12], Ae[225], Be[15], ul[15];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < nv; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
    const int *a = adr + (size_t)e * nf;
    for(int i = 0; i < nf; ++i) ul[i] = u[a[i]];
    memset(Ae, 0, sizeof(Ae));
    memset(Be, 0, sizeof(Be));
    or_geo g;
    if(elem_geo(&R, c, &g)) return -2 - e;
    for(int k = 0; k < R.nq; ++k) {
      double x[3];
      qp_pos(&R, c, k, x);
      const double cr = fn_eval(coeff, x, e, R.nq, k);
      double uq = 0.;
      for(int j = 0; j < nf; ++j) uq += R.L[k][j] * ul[j];
      const double d = p == 1. ? cr : cr * p * pow(uq, p - 1.);
      const double up = p == 1. ? uq : pow(uq, p);
      for(int i = 0; i < nf; ++i) {
        for(int j = 0; j < nf; ++j) Ae[i * nf + j] += d * R.L[k][i] * R.L[k][j] * g.J * R.qw[k];
        Be[i] -= cr * up * R.L[k][i] * g.J * R.qw[k];
      }
    }
    for(int i = 0; i < nf; ++i) {
      if(a[i] >= n) continue;
      rhs[a[i]] += Be[i];
      if(val)
        for(int j = 0; j < nf; ++j) {
          if(a[j] >= n) continue;
          int q = find_pos(ia, ja, a[i], a[j]);
          if(q < 0) return -1;
          val[q] += Ae[i * nf + j];
        }
    }
  }
  return 0;
}

int or_element_mass(int kind, const double *coords, int quad_degree, or_fn coeff, int elem, double *Ae)
{
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  memset(Ae, 0, sizeof(double) * R.nf * R.nf);
  transient_mass_elem(&R, coords, elem, coeff, 1., NULL, Ae, NULL);
  return R.nf;
}

int or_assemble_transient_mass(int kind, int V, const double *xyz, int E, const int *conn,
                               const int *adr, int n, const int *ia, const int *ja, int quad_degree,
                               or_fn coeff, double c0, const double *udot, double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv;
  double c[12], Ae[225], Be[15], udl[15];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < nv; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
    const int *a = adr + (size_t)e * nf;
    for(int i = 0; i < nf; ++i) udl[i] = udot[a[i]];
    memset(Ae, 0, sizeof(Ae));
    memset(Be, 0, sizeof(Be));
    or_geo g;
    if(elem_geo(&R, c, &g)) return -2 - e;
    transient_mass_elem(&R, c, e, coeff, c0, udl, val ? Ae : NULL, Be);
    for(int i = 0; i < nf; ++i) {
      if(a[i] >= n) continue;
      rhs[a[i]] += Be[i];
      if(val)
        for(int j = 0; j < nf; ++j) {
          if(a[j] >= n) continue;
          int p = find_pos(ia, ja, a[i], a[j]);
          if(p < 0) return -1;
          val[p] += Ae[i * nf + j];
        }
    }
  }
  return 0;
}

/* Threaded assembly in the reference's OpenMP form: colors one after the
 * other, the elements of a color in parallel (they share no vertex, so no
 * two threads add into the same row; feLinearSystemPETSc.cpp:543-707 runs
 * the colored element loop under `omp parallel for`).  Each element adds its
 * diffusion and source contributions in one visit. */
int or_assemble_diffusion_omp(int kind, int V, const double *xyz, int E, const int *conn,
                              const int *adr, int n, const int *ia, const int *ja,
                              int quad_degree, or_fn coeff, or_fn source, const double *u,
                              const int *colorOf, int ncolors, int threads, double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv;
  int *start = calloc((size_t)ncolors + 1, sizeof(int)), *elems = malloc(sizeof(int) * (size_t)E);
  for(int e = 0; e < E; ++e) start[colorOf[e] + 1]++;
  for(int c = 0; c < ncolors; ++c) start[c + 1] += start[c];
  int *fill = malloc(sizeof(int) * (size_t)ncolors);
  memcpy(fill, start, sizeof(int) * (size_t)ncolors);
  for(int e = 0; e < E; ++e) elems[fill[colorOf[e]]++] = e;
  free(fill);
  int err = 0;
#pragma omp parallel num_threads(threads) reduction(| : err)
  {
    double c[12], Ae[225], Be[15], ul[15];
    for(int col = 0; col < ncolors; ++col) {
#pragma omp for schedule(static)
      for(int q = start[col]; q < start[col + 1]; ++q) {
        const int e = elems[q];
        for(int j = 0; j < nv; ++j)
          for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
        const int *a = adr + (size_t)e * nf;
        for(int i = 0; i < nf; ++i) ul[i] = u[a[i]];
        memset(Ae, 0, sizeof(Ae));
        memset(Be, 0, sizeof(Be));
        or_geo g;
        if(elem_geo(&R, c, &g)) {
          err = 1;
          continue;
        }
        diffusion_elem(&R, c, e, coeff, source, ul, Ae, Be);
        for(int i = 0; i < nf; ++i) {
          if(a[i] >= n) continue;
          rhs[a[i]] += Be[i];
          for(int j = 0; j < nf; ++j) {
            if(a[j] >= n) continue;
            const int p = find_pos(ia, ja, a[i], a[j]);
            if(p < 0) err = 1;
            else val[p] += Ae[i * nf + j];
          }
        }
      }
    }
  }
  free(start);
  free(elems);
  return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Taylor-Hood Navier-Stokes: filled in by or_ns.c-style helpers below. */

/* ------------------------------------------------------------------------ */
/* Krylov solvers. */
int or_spmv(int n, const int *ia, const int *ja, const double *val, const double *x, double *y)
{
  for(int i = 0; i < n; ++i) {
    double s = 0.;
    for(int k = ia[i]; k < ia[i + 1]; ++k) s += val[k] * x[ja[k]];
    y[i] = s;
  }
  return 0;
}

static double dot(int n, const double *a, const double *b)
{
  double s = 0.;
  for(int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Preconditioned CG (Hestenes-Stiefel) from x0 = 0 with a Jacobi or IC(0)
 * preconditioner, stopped on the true-residual-equivalent recursive residual
 * ||r||_2 <= rtol ||b||_2. */
static void ic0_apply(int n, const int *ia, const int *ja, const double *lval, const double *r,
                      double *z);

static int pcg(int n, const int *ia, const int *ja, const double *val, const double *lval,
               const double *b, double *x, double rtol, int maxit, double *relres)
{
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  double *dinv = malloc(sizeof(double) * n);
  for(int i = 0; i < n; ++i) {
    double d = 0.;
    for(int k = ia[i]; k < ia[i + 1]; ++k)
      if(ja[k] == i) d = val[k];
    dinv[i] = d != 0. ? 1. / d : 1.;
    x[i] = 0.;
    r[i] = b[i];
  }
  double bn = sqrt(dot(n, b, b));
  if(bn == 0.) bn = 1.;
  if(lval) ic0_apply(n, ia, ja, lval, r, z);
  else for(int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(n, r, z);
  int it = 0;
  double rn = sqrt(dot(n, r, r));
  while(rn > rtol * bn && it < maxit) {
    or_spmv(n, ia, ja, val, p, q);
    double alpha = rz / dot(n, p, q);
    for(int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    if(lval) ic0_apply(n, ia, ja, lval, r, z);
    else for(int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    double rz1 = dot(n, r, z);
    double beta = rz1 / rz;
    rz = rz1;
    for(int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rn = sqrt(dot(n, r, r));
    it++;
  }
  /* true residual */
  or_spmv(n, ia, ja, val, x, q);
  double tr = 0.;
  for(int i = 0; i < n; ++i) tr += (b[i] - q[i]) * (b[i] - q[i]);
  if(relres) *relres = sqrt(tr) / bn;
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}

int or_cg_jacobi(int n, const int *ia, const int *ja, const double *val, const double *b,
                 double *x, double rtol, int maxit, double *relres)
{
  return pcg(n, ia, ja, val, NULL, b, x, rtol, maxit, relres);
}

/* The same Jacobi-PCG with every loop (SpMV, dots, updates) split over
 * `threads` OpenMP threads. */
int or_cg_jacobi_omp(int n, const int *ia, const int *ja, const double *val, const double *b,
                     double *x, double rtol, int maxit, int threads, double *relres)
{
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  double *dinv = malloc(sizeof(double) * n);
  double bb = 0., rz = 0., rr = 0.;
#pragma omp parallel for num_threads(threads) reduction(+ : bb, rz, rr)
  for(int i = 0; i < n; ++i) {
    double d = 0.;
    for(int k = ia[i]; k < ia[i + 1]; ++k)
      if(ja[k] == i) d = val[k];
    dinv[i] = d != 0. ? 1. / d : 1.;
    x[i] = 0.;
    r[i] = b[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
    bb += b[i] * b[i];
    rz += r[i] * z[i];
    rr += r[i] * r[i];
  }
  double bn = sqrt(bb);
  if(bn == 0.) bn = 1.;
  int it = 0;
  while(sqrt(rr) > rtol * bn && it < maxit) {
    double pq = 0.;
#pragma omp parallel for num_threads(threads) reduction(+ : pq)
    for(int i = 0; i < n; ++i) {
      double s = 0.;
      for(int k = ia[i]; k < ia[i + 1]; ++k) s += val[k] * p[ja[k]];
      q[i] = s;
      pq += p[i] * s;
    }
    const double alpha = rz / pq;
    double rz1 = 0., rr1 = 0.;
#pragma omp parallel for num_threads(threads) reduction(+ : rz1, rr1)
    for(int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
      rz1 += r[i] * z[i];
      rr1 += r[i] * r[i];
    }
    const double beta = rz1 / rz;
    rz = rz1;
    rr = rr1;
#pragma omp parallel for num_threads(threads)
    for(int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    it++;
  }
  double tr = 0.;
#pragma omp parallel for num_threads(threads) reduction(+ : tr)
  for(int i = 0; i < n; ++i) {
    double s = 0.;
    for(int k = ia[i]; k < ia[i + 1]; ++k) s += val[k] * x[ja[k]];
    tr += (b[i] - s) * (b[i] - s);
  }
  if(relres) *relres = sqrt(tr) / bn;
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}

/* IC(0): incomplete Cholesky with the pattern of the lower triangle of A
 * (the IC(0) pattern).  lval holds L on A's pattern (upper entries zero). */
int or_ic0(int n, const int *ia, const int *ja, const double *val, double *lval)
{
  int64_t nnz = ia[n];
  for(int64_t k = 0; k < nnz; ++k) lval[k] = 0.;
  double *diag = malloc(sizeof(double) * n);
  for(int i = 0; i < n; ++i) {
    /* row i of L: for j < i in pattern, L_ij = (a_ij - sum_{k<j} L_ik L_jk) / L_jj */
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      int j = ja[p];
      if(j > i) break;
      double s = val[p];
      /* sparse dot of rows i and j of L over columns < j */
      int pi = ia[i], pj = ia[j];
      while(pi < ia[i + 1] && pj < ia[j + 1]) {
        int ci = ja[pi], cj = ja[pj];
        if(ci >= j || cj >= j) break;
        if(ci == cj) { s -= lval[pi] * lval[pj]; pi++; pj++; }
        else if(ci < cj) pi++;
        else pj++;
      }
      if(j < i) {
        lval[p] = s / diag[j];
      } else {
        if(s <= 0.) { free(diag); return -1 - i; }
        diag[i] = sqrt(s);
        lval[p] = diag[i];
      }
    }
  }
  free(diag);
  return 0;
}

static void ic0_apply(int n, const int *ia, const int *ja, const double *lval, const double *r,
                      double *z)
{
  /* L y = r */
  for(int i = 0; i < n; ++i) {
    double s = r[i], d = 1.;
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      int j = ja[p];
      if(j < i) s -= lval[p] * z[j];
      else if(j == i) d = lval[p];
    }
    z[i] = s / d;
  }
  /* L^T z = y (column sweep from the bottom) */
  for(int i = n - 1; i >= 0; --i) {
    double d = 1.;
    for(int p = ia[i]; p < ia[i + 1]; ++p)
      if(ja[p] == i) d = lval[p];
    z[i] /= d;
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      int j = ja[p];
      if(j < i) z[j] -= lval[p] * z[i];
    }
  }
}

int or_pcg_ic0(int n, const int *ia, const int *ja, const double *val, const double *lval,
               const double *b, double *x, double rtol, int maxit, double *relres)
{
  return pcg(n, ia, ja, val, lval, b, x, rtol, maxit, relres);
}

/* ILU(0) on A's pattern (IKJ variant); unit-lower L and U share luval. */
int or_ilu0(int n, const int *ia, const int *ja, const double *val, double *lu)
{
  int64_t nnz = ia[n];
  memcpy(lu, val, sizeof(double) * nnz);
  int *diag = malloc(sizeof(int) * n), *iw = malloc(sizeof(int) * n);
  for(int i = 0; i < n; ++i) {
    diag[i] = -1;
    iw[i] = -1;
    for(int p = ia[i]; p < ia[i + 1]; ++p)
      if(ja[p] == i) diag[i] = p;
  }
  for(int i = 0; i < n; ++i) {
    for(int p = ia[i]; p < ia[i + 1]; ++p) iw[ja[p]] = p;
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      int k = ja[p];
      if(k >= i) break;
      double piv = lu[diag[k]];
      if(piv == 0.) { free(diag); free(iw); return -1 - k; }
      lu[p] /= piv;
      double l = lu[p];
      for(int q = diag[k] + 1; q < ia[k + 1]; ++q) {
        int w = iw[ja[q]];
        if(w >= 0) lu[w] -= l * lu[q];
      }
    }
    for(int p = ia[i]; p < ia[i + 1]; ++p) iw[ja[p]] = -1;
    if(diag[i] < 0 || lu[diag[i]] == 0.) { free(diag); free(iw); return -1 - i; }
  }
  free(diag);
  free(iw);
  return 0;
}

static void ilu0_apply(int n, const int *ia, const int *ja, const double *lu, const double *r,
                       double *z)
{
  for(int i = 0; i < n; ++i) {
    double s = r[i];
    for(int p = ia[i]; p < ia[i + 1] && ja[p] < i; ++p) s -= lu[p] * z[ja[p]];
    z[i] = s;
  }
  for(int i = n - 1; i >= 0; --i) {
    double s = z[i], d = 1.;
    for(int p = ia[i]; p < ia[i + 1]; ++p) {
      if(ja[p] > i) s -= lu[p] * z[ja[p]];
      else if(ja[p] == i) d = lu[p];
    }
    z[i] = s / d;
  }
}

/* Right-preconditioned restarted GMRES(m) with modified Gram-Schmidt and
 * Givens rotations, x0 = 0, stopped on ||b - A x||_2 <= rtol ||b||_2. */
int or_gmres_ilu0(int n, const int *ia, const int *ja, const double *val, const double *lu,
                  const double *b, double *x, int m, double rtol, int maxit, double *relres)
{
  double *Vb = malloc(sizeof(double) * (size_t)n * (m + 1));
  double *Z = malloc(sizeof(double) * (size_t)n * m);
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
  double *gg = malloc(sizeof(double) * (m + 1)), *w = malloc(sizeof(double) * n);
  double *yv = malloc(sizeof(double) * m);
  double bn = sqrt(dot(n, b, b));
  if(bn == 0.) bn = 1.;
  for(int i = 0; i < n; ++i) x[i] = 0.;
  int it = 0;
  double res = 0.;
  for(;;) {
    or_spmv(n, ia, ja, val, x, w);
    for(int i = 0; i < n; ++i) Vb[i] = b[i] - w[i];
    double beta = sqrt(dot(n, Vb, Vb));
    res = beta / bn;
    if(res <= rtol || it >= maxit) break;
    for(int i = 0; i < n; ++i) Vb[i] /= beta;
    for(int k = 0; k <= m; ++k) gg[k] = 0.;
    gg[0] = beta;
    int j;
    for(j = 0; j < m && it < maxit; ++j, ++it) {
      ilu0_apply(n, ia, ja, lu, Vb + (size_t)j * n, Z + (size_t)j * n);
      or_spmv(n, ia, ja, val, Z + (size_t)j * n, w);
      for(int k = 0; k <= j; ++k) {
        double h = dot(n, w, Vb + (size_t)k * n);
        H[k * m + j] = h;
        for(int i = 0; i < n; ++i) w[i] -= h * Vb[(size_t)k * n + i];
      }
      double hn = sqrt(dot(n, w, w));
      H[(j + 1) * m + j] = hn;
      for(int i = 0; i < n; ++i) Vb[(size_t)(j + 1) * n + i] = hn > 0. ? w[i] / hn : 0.;
      for(int k = 0; k < j; ++k) {
        double t = cs[k] * H[k * m + j] + sn[k] * H[(k + 1) * m + j];
        H[(k + 1) * m + j] = -sn[k] * H[k * m + j] + cs[k] * H[(k + 1) * m + j];
        H[k * m + j] = t;
      }
      double a = H[j * m + j], bb = H[(j + 1) * m + j];
      double r = sqrt(a * a + bb * bb);
      cs[j] = r > 0. ? a / r : 1.;
      sn[j] = r > 0. ? bb / r : 0.;
      H[j * m + j] = r;
      H[(j + 1) * m + j] = 0.;
      gg[j + 1] = -sn[j] * gg[j];
      gg[j] = cs[j] * gg[j];
      if(fabs(gg[j + 1]) / bn <= rtol) { ++j; ++it; break; }
    }
    for(int k = j - 1; k >= 0; --k) {
      double s = gg[k];
      for(int l = k + 1; l < j; ++l) s -= H[k * m + l] * yv[l];
      yv[k] = s / H[k * m + k];
    }
    for(int k = 0; k < j; ++k)
      for(int i = 0; i < n; ++i) x[i] += yv[k] * Z[(size_t)k * n + i];
  }
  if(relres) *relres = res;
  free(Vb); free(Z); free(H); free(cs); free(sn); free(gg); free(w); free(yv);
  return it;
}

/* ------------------------------------------------------------------------ */
/* L2 error vs u = sin(pi x) sin(pi y) (feNorm.cpp:323-398: sum over elements
 * and quadrature points of (u_h - u)^2 J w, square root at the end). */
double or_l2_error_sinsin(int kind, const double *xyz, int E, const int *conn, const int *adr,
                          const double *u, int quad_degree)
{
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1.;
  double acc = 0., c[9];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < 3; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * 3 + j] + d];
    or_geo g;
    geo_tri(c, &g);
    for(int k = 0; k < R.nq; ++k) {
      double x[3];
      qp_pos(&R, c, k, x);
      double uh = 0.;
      for(int i = 0; i < R.nf; ++i) uh += u[adr[(size_t)e * R.nf + i]] * R.L[k][i];
      double ue = sin(M_PI * x[0]) * sin(M_PI * x[1]);
      acc += (uh - ue) * (uh - ue) * g.J * R.qw[k];
    }
  }
  return sqrt(acc);
}

/* L2 error against an exact solution tabulated at the quadrature points
 * (ue[e nq + q]), feNorm.cpp:323-398 with p = 2, triangles. */
double or_l2_error_qp(int kind, const double *xyz, int E, const int *conn, const int *adr, const double *u,
                      int quad_degree, const double *ue)
{
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1.;
  double acc = 0., c[9];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < 3; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * 3 + j] + d];
    or_geo g;
    geo_tri(c, &g);
    for(int k = 0; k < R.nq; ++k) {
      double uh = 0.;
      for(int i = 0; i < R.nf; ++i) uh += u[adr[(size_t)e * R.nf + i]] * R.L[k][i];
      const double d = uh - ue[(size_t)e * R.nq + k];
      acc += d * d * g.J * R.qw[k];
    }
  }
  return sqrt(acc);
}

/* Advection in conservative form (feSysElm.cpp:766-835): Ae_ij -=
 * sum_d c_d dphi_i/dx_d phi_j J w, Be_i += sum_d c_d u_h dphi_i/dx_d J w,
 * the velocity given by its components cx, cy (cz for tets).  val may be
 * NULL.  Adds into val/rhs. */
int or_assemble_advection(int kind, int V, const double *xyz, int E, const int *conn, const int *adr, int n,
                          const int *ia, const int *ja, int quad_degree, or_fn cx, or_fn cy, or_fn cz,
                          const double *u, double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv, dim = R.dim;
  double c[12], Ae[225], Be[15], ul[15], gp[45];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < nv; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
    const int *a = adr + (size_t)e * nf;
    for(int i = 0; i < nf; ++i) ul[i] = u[a[i]];
    memset(Ae, 0, sizeof(Ae));
    memset(Be, 0, sizeof(Be));
    or_geo g;
    if(elem_geo(&R, c, &g)) return -2 - e;
    for(int k = 0; k < R.nq; ++k) {
      double x[3];
      qp_pos(&R, c, k, x);
      const double cv[3] = {fn_eval(cx, x, e, R.nq, k), fn_eval(cy, x, e, R.nq, k),
                            dim == 3 ? fn_eval(cz, x, e, R.nq, k) : 0.};
      phys_grad(&R, &g, k, gp);
      double uq = 0.;
      for(int j = 0; j < nf; ++j) uq += R.L[k][j] * ul[j];
      for(int i = 0; i < nf; ++i) {
        for(int j = 0; j < nf; ++j)
          for(int d = 0; d < dim; ++d) Ae[i * nf + j] -= cv[d] * gp[i * dim + d] * R.L[k][j] * g.J * R.qw[k];
        for(int d = 0; d < dim; ++d) Be[i] += cv[d] * uq * gp[i * dim + d] * g.J * R.qw[k];
      }
    }
    for(int i = 0; i < nf; ++i) {
      if(a[i] >= n) continue;
      rhs[a[i]] += Be[i];
      if(val)
        for(int j = 0; j < nf; ++j) {
          if(a[j] >= n) continue;
          int q = find_pos(ia, ja, a[i], a[j]);
          if(q < 0) return -1;
          val[q] += Ae[i * nf + j];
        }
    }
  }
  return 0;
}

static double poly_u(int n, const double *c, double u)
{
  double v = 0.;
  for(int i = n - 1; i >= 0; --i) v = v * u + c[i];
  return v;
}

int or_assemble_nonlinear_diffusion(int kind, int V, const double *xyz, int E, const int *conn, const int *adr,
                                    int n, const int *ia, const int *ja, int quad_degree, int nk, const double *k,
                                    int ndk, const double *dk, const double *u, double *val, double *rhs)
{
  (void)V;
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv, dim = R.dim;
  double c[12], Ae[225], Be[15], ul[15], gp[45];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < nv; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
    const int *a = adr + (size_t)e * nf;
    for(int i = 0; i < nf; ++i) ul[i] = u[a[i]];
    memset(Ae, 0, sizeof(Ae));
    memset(Be, 0, sizeof(Be));
    or_geo g;
    if(elem_geo(&R, c, &g)) return -2 - e;
    for(int q = 0; q < R.nq; ++q) {
      phys_grad(&R, &g, q, gp);
      double uq = 0., gu[3] = {0., 0., 0.};
      for(int j = 0; j < nf; ++j) {
        uq += R.L[q][j] * ul[j];
        for(int d = 0; d < dim; ++d) gu[d] += ul[j] * gp[j * dim + d];
      }
      const double ku = poly_u(nk, k, uq), dku = poly_u(ndk, dk, uq), jw = g.J * R.qw[q];
      for(int i = 0; i < nf; ++i) {
        double gi_gu = 0.;
        for(int d = 0; d < dim; ++d) gi_gu += gp[i * dim + d] * gu[d];
        for(int j = 0; j < nf; ++j) {
          double gi_gj = 0.;
          for(int d = 0; d < dim; ++d) gi_gj += gp[i * dim + d] * gp[j * dim + d];
          Ae[i * nf + j] += (ku * gi_gj + dku * R.L[q][j] * gi_gu) * jw;
        }
        Be[i] -= ku * gi_gu * jw;
      }
    }
    for(int i = 0; i < nf; ++i) {
      if(a[i] >= n) continue;
      rhs[a[i]] += Be[i];
      if(val)
        for(int j = 0; j < nf; ++j) {
          if(a[j] >= n) continue;
          int p = find_pos(ia, ja, a[i], a[j]);
          if(p < 0) return -1;
          val[p] += Ae[i * nf + j];
        }
    }
  }
  return 0;
}

int or_assemble_mixed(int kind, const double *xyz, int E, const int *conn, const int *adrI, const int *adrJ, int n,
                      const int *ia, const int *ja, int quad_degree, double cm, double cp, double p, double cg,
                      const double *u, double *val, double *rhs)
{
  or_ref R;
  if(ref_setup(kind, quad_degree, &R)) return -1;
  const int nf = R.nf, nv = R.nv, dim = R.dim;
  double c[12], Ae[225], Be[15], uj[15], gp[45];
  for(int e = 0; e < E; ++e) {
    for(int j = 0; j < nv; ++j)
      for(int d = 0; d < 3; ++d) c[3 * j + d] = xyz[3 * conn[e * nv + j] + d];
    const int *ai = adrI + (size_t)e * nf, *aj = adrJ + (size_t)e * nf;
    for(int j = 0; j < nf; ++j) uj[j] = u[aj[j]];
    memset(Ae, 0, sizeof(Ae));
    memset(Be, 0, sizeof(Be));
    or_geo g;
    if(elem_geo(&R, c, &g)) return -2 - e;
    for(int q = 0; q < R.nq; ++q) {
      phys_grad(&R, &g, q, gp);
      double uq = 0., gu[3] = {0., 0., 0.};
      for(int j = 0; j < nf; ++j) {
        uq += R.L[q][j] * uj[j];
        for(int d = 0; d < dim; ++d) gu[d] += uj[j] * gp[j * dim + d];
      }
      const double jw = g.J * R.qw[q];
      const double val_q = cm * uq + (cp != 0. ? cp * pow(uq, p) : 0.);
      const double dval = cm + (cp != 0. ? cp * p * pow(uq, p - 1.) : 0.);
      for(int i = 0; i < nf; ++i) {
        double gi_gu = 0.;
        for(int d = 0; d < dim; ++d) gi_gu += gp[i * dim + d] * gu[d];
        Be[i] -= (val_q * R.L[q][i] + cg * gi_gu) * jw;
        for(int j = 0; j < nf; ++j) {
          double gi_gj = 0.;
          for(int d = 0; d < dim; ++d) gi_gj += gp[i * dim + d] * gp[j * dim + d];
          Ae[i * nf + j] += (dval * R.L[q][i] * R.L[q][j] + cg * gi_gj) * jw;
        }
      }
    }
    for(int i = 0; i < nf; ++i) {
      if(ai[i] >= n) continue;
      rhs[ai[i]] += Be[i];
      if(val)
        for(int j = 0; j < nf; ++j) {
          if(aj[j] >= n) continue;
          int pp = find_pos(ia, ja, ai[i], aj[j]);
          if(pp < 0) return -1;
          val[pp] += Ae[i * nf + j];
        }
    }
  }
  return 0;
}

/* Taylor-Hood NS forms: implemented in fe_oracle_ns.c. */
