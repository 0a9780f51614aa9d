// libfe_host.so: meshes, edges, DOF numbering, coloring and synthetic
// coefficient fields for the GPU assemble-and-solve path (C ABI in
// include/fe_mesh.h).  Pure host C++17.
#include "../fe_jacobian_check.h"
#include "fe_mesh.h"
#include "../fe_lagrange.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

struct fe_mesh {
  int dim = 2, V = 0, E = 0, nvpe = 3, nB = 0, bnv = 2, nEdges = 0;
  std::vector<double> xyz;  // 3V
  std::vector<int> conn;    // E nvpe
  std::vector<int> bconn;   // nB bnv
  std::vector<int> edges;   // 3E signed tags (2D)
  std::vector<int> edgeA;   // first vertex of each edge as its first element walks it (2D)
  std::vector<int> bedge;   // nB edge index (2D)
  // boundary physical groups of a mesh read from a file (physical-tag order)
  // and the group of each boundary element; empty: one group "Bord"
  std::vector<std::string> groups;
  std::vector<int> bgroup;
};

struct fe_numbering {
  int nfields = 0, nInc = 0, nDofs = 0, E = 0;
  std::vector<int> kinds;
  std::vector<std::vector<int>> adr;
  std::vector<double> dofXyz;
  std::vector<int> dofField, dofComp;
  // per field: DOFs of each vertex (nvd per vertex: components) and of each
  // edge (ned per edge: components for the vector kind, else positions from
  // the edge's first vertex)
  std::vector<std::vector<int>> vnum, enm;
  std::vector<int> nvd, ned;
};

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): the project's seeded stream.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next()
  {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double normal()
  {
    double u1 = uniform(), u2 = uniform();
    if(u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2. * std::log(u1)) * std::cos(2. * M_PI * u2);
  }
};

int nlocOf(int kind)
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


// Edge tags by first encounter while walking triangles in order and local
// edges (0,1), (1,2), (2,0): the rule of feMeshRead.cpp:1412-1456 (a std::set
// keyed on the endpoint pair there; a hash map here -- only the tag order
// matters).
void computeEdges(fe_mesh &m)
{
  m.edges.assign((size_t)3 * m.E, 0);
  m.edgeA.clear();
  std::unordered_map<uint64_t, int> tag;
  tag.reserve((size_t)m.E * 2);
  int next = 1;
  for(int e = 0; e < m.E; ++e) {
    for(int k = 0; k < 3; ++k) {
      uint32_t a = (uint32_t)m.conn[3 * e + k], b = (uint32_t)m.conn[3 * e + (k + 1) % 3];
      uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
      auto it = tag.find(key);
      if(it == tag.end()) {
        tag.emplace(key, next);
        m.edges[3 * e + k] = next++;
        m.edgeA.push_back((int)a);
      } else {
        m.edges[3 * e + k] = -it->second;
      }
    }
  }
  m.nEdges = next - 1;
  m.bedge.assign(m.nB, -1);
  for(int b = 0; b < m.nB; ++b) {
    uint32_t a = (uint32_t)m.bconn[2 * b], c = (uint32_t)m.bconn[2 * b + 1];
    uint64_t key = a < c ? ((uint64_t)a << 32 | c) : ((uint64_t)c << 32 | a);
    auto it = tag.find(key);
    m.bedge[b] = (it == tag.end()) ? -1 : it->second - 1;
  }
}

}  // namespace

extern "C" {

int fe_mesh_square(int N, double jitter, uint64_t seed, fe_mesh **out)
{
  if(!out || N < 1) return FE_STATUS_ERROR;
  fe_mesh *m = new(std::nothrow) fe_mesh;
  if(!m) return FE_STATUS_OUT_OF_MEMORY;
  const int n1 = N + 1;
  const double h = 1.0 / N;
  m->dim = 2;
  m->nvpe = 3;
  m->bnv = 2;
  m->V = n1 * n1;
  m->xyz.resize((size_t)3 * m->V);
  SplitMix64 rng(seed);
  for(int j = 0; j < n1; ++j) {
    for(int i = 0; i < n1; ++i) {
      double x = i * h, y = j * h;
      if(jitter > 0. && i > 0 && i < N && j > 0 && j < N) {
        x += (2. * rng.uniform() - 1.) * jitter * h;
        y += (2. * rng.uniform() - 1.) * jitter * h;
      }
      size_t v = (size_t)i + (size_t)n1 * j;
      m->xyz[3 * v] = x;
      m->xyz[3 * v + 1] = y;
      m->xyz[3 * v + 2] = 0.;
    }
  }
  m->E = 2 * N * N;
  m->conn.resize((size_t)3 * m->E);
  size_t c = 0;
  for(int j = 0; j < N; ++j) {
    for(int i = 0; i < N; ++i) {
      int v00 = i + n1 * j, v10 = v00 + 1, v01 = v00 + n1, v11 = v01 + 1;
      m->conn[c++] = v00; m->conn[c++] = v10; m->conn[c++] = v11;
      m->conn[c++] = v00; m->conn[c++] = v11; m->conn[c++] = v01;
    }
  }
  m->nB = 4 * N;
  m->bconn.resize((size_t)2 * m->nB);
  c = 0;
  for(int i = 0; i < N; ++i) { m->bconn[c++] = i; m->bconn[c++] = i + 1; }
  for(int j = 0; j < N; ++j) { m->bconn[c++] = N + n1 * j; m->bconn[c++] = N + n1 * (j + 1); }
  for(int i = N; i > 0; --i) { m->bconn[c++] = i + n1 * N; m->bconn[c++] = i - 1 + n1 * N; }
  for(int j = N; j > 0; --j) { m->bconn[c++] = n1 * j; m->bconn[c++] = n1 * (j - 1); }
  computeEdges(*m);
  *out = m;
  return FE_STATUS_OK;
}

static double tetDet(const double *x, const int *t)
{
  double a[3], b[3], d[3];
  for(int k = 0; k < 3; ++k) {
    a[k] = x[3 * t[1] + k] - x[3 * t[0] + k];
    b[k] = x[3 * t[2] + k] - x[3 * t[0] + k];
    d[k] = x[3 * t[3] + k] - x[3 * t[0] + k];
  }
  return a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
         a[2] * (b[0] * d[1] - b[1] * d[0]);
}

static void boundaryFacets(fe_mesh &m)
{
  // Faces of every tet (opposite each vertex); a face seen once is on the
  // boundary.  Boundary triangles keep the tet's face orientation (outward).
  static const int F[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  std::unordered_map<uint64_t, int> count;
  count.reserve((size_t)m.E * 3);
  auto key = [](int a, int b, int c) {
    int s[3] = {a, b, c};
    std::sort(s, s + 3);
    return ((uint64_t)s[0] << 42) ^ ((uint64_t)s[1] << 21) ^ (uint64_t)s[2];
  };
  for(int e = 0; e < m.E; ++e)
    for(int f = 0; f < 4; ++f)
      count[key(m.conn[4 * e + F[f][0]], m.conn[4 * e + F[f][1]], m.conn[4 * e + F[f][2]])]++;
  m.bconn.clear();
  for(int e = 0; e < m.E; ++e)
    for(int f = 0; f < 4; ++f) {
      int a = m.conn[4 * e + F[f][0]], b = m.conn[4 * e + F[f][1]], c = m.conn[4 * e + F[f][2]];
      if(count[key(a, b, c)] == 1) {
        m.bconn.push_back(a);
        m.bconn.push_back(b);
        m.bconn.push_back(c);
      }
    }
  m.nB = (int)(m.bconn.size() / 3);
  m.bnv = 3;
}

int fe_mesh_cube(int N, fe_mesh **out)
{
  if(!out || N < 1) return FE_STATUS_ERROR;
  fe_mesh *m = new(std::nothrow) fe_mesh;
  if(!m) return FE_STATUS_OUT_OF_MEMORY;
  const int n1 = N + 1;
  const double h = 1.0 / N;
  m->dim = 3;
  m->nvpe = 4;
  m->V = n1 * n1 * n1;
  m->xyz.resize((size_t)3 * m->V);
  for(int k = 0; k < n1; ++k)
    for(int j = 0; j < n1; ++j)
      for(int i = 0; i < n1; ++i) {
        size_t v = (size_t)i + (size_t)n1 * (j + (size_t)n1 * k);
        m->xyz[3 * v] = i * h;
        m->xyz[3 * v + 1] = j * h;
        m->xyz[3 * v + 2] = k * h;
      }
  static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int stride[3] = {1, n1, n1 * n1};
  m->E = 6 * N * N * N;
  m->conn.resize((size_t)4 * m->E);
  size_t c = 0;
  for(int k = 0; k < N; ++k)
    for(int j = 0; j < N; ++j)
      for(int i = 0; i < N; ++i) {
        int v0 = i + n1 * (j + n1 * k);
        for(int p = 0; p < 6; ++p) {
          int t[4];
          t[0] = v0;
          t[1] = t[0] + stride[P[p][0]];
          t[2] = t[1] + stride[P[p][1]];
          t[3] = t[2] + stride[P[p][2]];
          if(tetDet(m->xyz.data(), t) < 0) std::swap(t[2], t[3]);
          for(int q = 0; q < 4; ++q) m->conn[c++] = t[q];
        }
      }
  boundaryFacets(*m);
  m->nEdges = 0;
  *out = m;
  return FE_STATUS_OK;
}

int fe_mesh_from_arrays(int dim, int V, const double *xyz, int E, int nvpe, const int *conn,
                        int nB, int bnv, const int *bconn, fe_mesh **out)
{
  if(!out || !xyz || !conn || (nB > 0 && !bconn)) return FE_STATUS_POINTER_ERROR;
  if(!((dim == 2 && nvpe == 3 && (nB == 0 || bnv == 2)) ||
       (dim == 3 && nvpe == 4 && (nB == 0 || bnv == 3))))
    return FE_STATUS_FORMAT_ERROR;
  fe_mesh *m = new(std::nothrow) fe_mesh;
  if(!m) return FE_STATUS_OUT_OF_MEMORY;
  m->dim = dim;
  m->V = V;
  m->E = E;
  m->nvpe = nvpe;
  m->nB = nB;
  m->bnv = dim == 2 ? 2 : 3;
  m->xyz.assign(xyz, xyz + (size_t)3 * V);
  m->conn.assign(conn, conn + (size_t)E * nvpe);
  if(nB > 0) m->bconn.assign(bconn, bconn + (size_t)nB * m->bnv);
  for(int v : m->conn)
    if(v < 0 || v >= V) { delete m; return FE_STATUS_RANGE_ERROR; }
  for(int v : m->bconn)
    if(v < 0 || v >= V) { delete m; return FE_STATUS_RANGE_ERROR; }
  if(dim == 2) computeEdges(*m);
  *out = m;
  return FE_STATUS_OK;
}

int fe_mesh_write_msh(const fe_mesh *m, const char *path)
{
  if(!m || !path) return FE_STATUS_POINTER_ERROR;
  FILE *f = fopen(path, "w");
  if(!f) return FE_STATUS_FILE_CANNOT_BE_OPENED;
  const int bdim = m->dim - 1;
  fprintf(f, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n");
  fprintf(f, "$PhysicalNames\n2\n%d 2 \"Bord\"\n%d 1 \"Domaine\"\n$EndPhysicalNames\n", bdim,
          m->dim);
  fprintf(f, "$Entities\n");
  if(m->dim == 2) {
    fprintf(f, "0 1 1 0\n");
    fprintf(f, "1 0 0 0 1 1 0 1 2 0\n");
    fprintf(f, "1 0 0 0 1 1 0 1 1 1 1\n");
  } else {
    fprintf(f, "0 0 1 1\n");
    fprintf(f, "1 0 0 0 1 1 1 1 2 0\n");
    fprintf(f, "1 0 0 0 1 1 1 1 1 1 1\n");
  }
  fprintf(f, "$EndEntities\n");
  fprintf(f, "$Nodes\n1 %d 1 %d\n%d 1 0 %d\n", m->V, m->V, m->dim, m->V);
  for(int v = 0; v < m->V; ++v) fprintf(f, "%d\n", v + 1);
  for(int v = 0; v < m->V; ++v)
    fprintf(f, "%.17g %.17g %.17g\n", m->xyz[3 * v], m->xyz[3 * v + 1], m->xyz[3 * v + 2]);
  fprintf(f, "$EndNodes\n");
  const int tot = m->nB + m->E;
  fprintf(f, "$Elements\n2 %d 1 %d\n", tot, tot);
  fprintf(f, "%d 1 %d %d\n", bdim, m->dim == 2 ? 1 : 2, m->nB);
  int tag = 1;
  for(int b = 0; b < m->nB; ++b) {
    fprintf(f, "%d", tag++);
    for(int k = 0; k < m->bnv; ++k) fprintf(f, " %d", m->bconn[(size_t)b * m->bnv + k] + 1);
    fprintf(f, "\n");
  }
  fprintf(f, "%d 1 %d %d\n", m->dim, m->dim == 2 ? 2 : 4, m->E);
  for(int e = 0; e < m->E; ++e) {
    fprintf(f, "%d", tag++);
    for(int k = 0; k < m->nvpe; ++k) fprintf(f, " %d", m->conn[(size_t)e * m->nvpe + k] + 1);
    fprintf(f, "\n");
  }
  fprintf(f, "$EndElements\n");
  int rc = ferror(f) ? FE_STATUS_WRITE_ERROR : FE_STATUS_OK;
  fclose(f);
  return rc;
}

// Reader for .msh v4.1 ASCII files (feMeshRead.cpp:1709-2005): the domain is
// every element of the highest dimension in a physical group ("Domaine",
// "Domain", ...), the boundary every element one dimension below in a
// physical group ("Bord", or "Inlet", "Outlet", "NoSlip", ...), each
// boundary element tagged with its group (groups in physical-tag order).
// Vertex ids in $Nodes order (feMeshRead.cpp:933-958); domain elements
// ordered by geometric entity (dim, tag) then file order
// (feMeshRead.cpp:1777-1808); boundary elements in file order (their global
// element order, which drives essential edge numbering).
int fe_mesh_read_msh(const char *path, fe_mesh **out)
{
  if(!path || !out) return FE_STATUS_POINTER_ERROR;
  std::ifstream in(path);
  if(!in) return FE_STATUS_FILE_CANNOT_BE_OPENED;
  std::map<std::pair<int, int>, std::string> physName;
  std::map<std::pair<int, int>, int> entityPhys;  // (dim, entity tag) -> physical tag
  std::unordered_map<long, int> nodeId;
  std::vector<double> xyz;
  struct Block {
    int dim, tag, type;
    std::vector<long> nodes;
    int nper;
  };
  std::vector<Block> blocks;
  std::string tok;
  int maxDim = 0;
  while(in >> tok) {
    if(tok == "$PhysicalNames") {
      int n;
      in >> n;
      for(int i = 0; i < n; ++i) {
        int d, t;
        std::string name;
        in >> d >> t;
        std::getline(in, name);
        size_t a = name.find('"'), b = name.rfind('"');
        if(a == std::string::npos || b <= a) return FE_STATUS_FORMAT_ERROR;
        physName[{d, t}] = name.substr(a + 1, b - a - 1);
      }
    } else if(tok == "$Entities") {
      int np, nc, ns, nv;
      in >> np >> nc >> ns >> nv;
      int counts[4] = {np, nc, ns, nv};
      for(int d = 0; d < 4; ++d) {
        for(int i = 0; i < counts[d]; ++i) {
          int tag, nphys;
          double bb;
          in >> tag;
          int nbox = (d == 0) ? 3 : 6;
          for(int k = 0; k < nbox; ++k) in >> bb;
          in >> nphys;
          int ph = 0;
          for(int k = 0; k < nphys; ++k) {
            int p;
            in >> p;
            if(k == 0) ph = p;
          }
          if(d > 0) {
            int nbnd, bt;
            in >> nbnd;
            for(int k = 0; k < nbnd; ++k) in >> bt;
          }
          entityPhys[{d, tag}] = ph;
        }
      }
    } else if(tok == "$Nodes") {
      long nblocks, nnodes, mn, mx;
      in >> nblocks >> nnodes >> mn >> mx;
      xyz.reserve((size_t)3 * nnodes);
      for(long b = 0; b < nblocks; ++b) {
        int d, t, par;
        long nn;
        in >> d >> t >> par >> nn;
        if(par != 0) return FE_STATUS_FORMAT_ERROR;
        std::vector<long> tags(nn);
        for(long i = 0; i < nn; ++i) in >> tags[i];
        for(long i = 0; i < nn; ++i) {
          double x, y, z;
          in >> x >> y >> z;
          nodeId[tags[i]] = (int)(xyz.size() / 3);
          xyz.push_back(x);
          xyz.push_back(y);
          xyz.push_back(z);
        }
      }
    } else if(tok == "$Elements") {
      long nblocks, nel, mn, mx;
      in >> nblocks >> nel >> mn >> mx;
      for(long b = 0; b < nblocks; ++b) {
        Block B;
        long ne;
        in >> B.dim >> B.tag >> B.type >> ne;
        int nper = 0;
        switch(B.type) {
          case 15: nper = 1; break;
          case 1: nper = 2; break;
          case 2: nper = 3; break;
          case 4: nper = 4; break;
          default: return FE_STATUS_FORMAT_ERROR;
        }
        B.nper = nper;
        B.nodes.resize((size_t)ne * nper);
        for(long e = 0; e < ne; ++e) {
          long et;
          in >> et;
          for(int k = 0; k < nper; ++k) in >> B.nodes[(size_t)e * nper + k];
        }
        maxDim = std::max(maxDim, B.dim);
        blocks.push_back(std::move(B));
      }
    }
  }
  if(maxDim < 2) return FE_STATUS_FORMAT_ERROR;
  auto physOf = [&](int d, int tag) -> std::string {
    auto it = entityPhys.find({d, tag});
    if(it == entityPhys.end()) return "";
    auto jt = physName.find({d, it->second});
    return jt == physName.end() ? "" : jt->second;
  };
  fe_mesh *m = new(std::nothrow) fe_mesh;
  if(!m) return FE_STATUS_OUT_OF_MEMORY;
  m->dim = maxDim;
  m->nvpe = maxDim + 1;
  m->bnv = maxDim;
  m->V = (int)(xyz.size() / 3);
  m->xyz = std::move(xyz);
  // Domain triangles in file order (edge tags are assigned in this order),
  // then reordered by entity tag for the element numbering.
  std::vector<std::pair<int, size_t>> domBlocks;
  // boundary groups: the named physical groups of dimension maxDim - 1
  std::map<int, int> groupOfTag;  // physical tag -> group index
  for(auto &pn : physName)
    if(pn.first.first == maxDim - 1) {
      groupOfTag[pn.first.second] = (int)m->groups.size();
      m->groups.push_back(pn.second);
    }
  for(size_t b = 0; b < blocks.size(); ++b) {
    const Block &B = blocks[b];
    if(B.dim == maxDim && !physOf(B.dim, B.tag).empty()) domBlocks.push_back({B.tag, b});
    if(B.dim == maxDim - 1 && !physOf(B.dim, B.tag).empty()) {
      const int grp = groupOfTag.at(entityPhys.at({B.dim, B.tag}));
      for(long v : B.nodes) {
        auto it = nodeId.find(v);
        if(it == nodeId.end()) { delete m; return FE_STATUS_READ_ERROR; }
        m->bconn.push_back(it->second);
      }
      for(size_t e = 0; e < B.nodes.size() / (size_t)B.nper; ++e) m->bgroup.push_back(grp);
    }
  }
  m->nB = (int)(m->bconn.size() / m->bnv);
  std::vector<int> fileConn;
  std::vector<int> blockStart;
  for(auto &db : domBlocks) {
    blockStart.push_back((int)(fileConn.size() / m->nvpe));
    for(long v : blocks[db.second].nodes) {
      auto it = nodeId.find(v);
      if(it == nodeId.end()) { delete m; return FE_STATUS_READ_ERROR; }
      fileConn.push_back(it->second);
    }
  }
  m->E = (int)(fileConn.size() / m->nvpe);
  std::vector<int> fileEdges;
  if(maxDim == 2) {
    m->conn = fileConn;
    computeEdges(*m);
    fileEdges = m->edges;
  }
  std::vector<size_t> order(domBlocks.size());
  for(size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return domBlocks[a].first < domBlocks[b].first; });
  m->conn.clear();
  m->edges.clear();
  for(size_t oi : order) {
    int s = blockStart[oi];
    int cnt = (int)(blocks[domBlocks[oi].second].nodes.size() / m->nvpe);
    for(int e = s; e < s + cnt; ++e) {
      for(int k = 0; k < m->nvpe; ++k) m->conn.push_back(fileConn[(size_t)e * m->nvpe + k]);
      if(maxDim == 2)
        for(int k = 0; k < 3; ++k) m->edges.push_back(fileEdges[(size_t)3 * e + k]);
    }
  }
  *out = m;
  return FE_STATUS_OK;
}

int fe_mesh_destroy(fe_mesh *m)
{
  delete m;
  return FE_STATUS_OK;
}

int fe_mesh_info(const fe_mesh *m, int *dim, int *V, int *E, int *nvpe, int *nB, int *bnv,
                 int *nEdges)
{
  if(!m) return FE_STATUS_POINTER_ERROR;
  if(dim) *dim = m->dim;
  if(V) *V = m->V;
  if(E) *E = m->E;
  if(nvpe) *nvpe = m->nvpe;
  if(nB) *nB = m->nB;
  if(bnv) *bnv = m->bnv;
  if(nEdges) *nEdges = m->nEdges;
  return FE_STATUS_OK;
}

int fe_mesh_get(const fe_mesh *m, double *xyz, int *conn, int *bconn, int *edges, int *bedge)
{
  if(!m) return FE_STATUS_POINTER_ERROR;
  if(xyz) std::copy(m->xyz.begin(), m->xyz.end(), xyz);
  if(conn) std::copy(m->conn.begin(), m->conn.end(), conn);
  if(bconn) std::copy(m->bconn.begin(), m->bconn.end(), bconn);
  if(edges) std::copy(m->edges.begin(), m->edges.end(), edges);
  if(bedge) std::copy(m->bedge.begin(), m->bedge.end(), bedge);
  return FE_STATUS_OK;
}

// ---------------------------------------------------------------------------
// Numbering.  Codes per entity: -1 unknown, -2 essential; numbers >= 0 after
// numbering.  Order (feNumber.cpp:370-483 driven by feNumber.cpp:775-784):
//   for each field: unknown vertex DOFs (vertex asc, component inner),
//                   unknown edge DOFs (edge asc, component inner);
//   then for each field: essential vertex DOFs, essential boundary-segment
//                   DOFs in boundary element order (the 1D element DOFs of the
//                   boundary space, feNumber.cpp:452-466), leftover essential
//                   edge DOFs.
int fe_mesh_check_jacobians(const fe_mesh *m, int *elem, double *det)
{
  if(!m) return FE_STATUS_POINTER_ERROR;
  const int e = feFirstNonPositiveJacobian(m->dim, m->xyz.data(), m->E, m->nvpe, m->conn.data(), det);
  if(elem) *elem = e;
  return e < 0 ? FE_STATUS_OK : FE_STATUS_ERROR;
}

int fe_mesh_boundary_groups(const fe_mesh *m, int *ngroups, char *names, int len, int *bgroup)
{
  if(!m) return FE_STATUS_POINTER_ERROR;
  const std::vector<std::string> dflt = {"Bord"};
  const std::vector<std::string> &g = m->groups.empty() ? dflt : m->groups;
  if(ngroups) *ngroups = (int)g.size();
  if(names) {
    std::string all;
    for(size_t i = 0; i < g.size(); ++i) all += (i ? "\n" : "") + g[i];
    if((int)all.size() + 1 > len) return FE_STATUS_RANGE_ERROR;
    std::copy(all.begin(), all.end(), names);
    names[all.size()] = 0;
  }
  if(bgroup)
    for(int b = 0; b < m->nB; ++b) bgroup[b] = m->bgroup.empty() ? 0 : m->bgroup[b];
  return FE_STATUS_OK;
}

int fe_numbering_create(const fe_mesh *m, int nfields, const int *kinds, const int *essential,
                        fe_numbering **out)
{
  if(!m || !kinds || !essential || !out || nfields < 1) return FE_STATUS_POINTER_ERROR;
  std::vector<unsigned> mask(nfields);
  for(int f = 0; f < nfields; ++f) mask[f] = essential[f] ? ~0u : 0u;
  return fe_numbering_create_groups(m, nfields, kinds, mask.data(), out);
}

int fe_numbering_create_groups(const fe_mesh *m, int nfields, const int *kinds, const unsigned *essentialGroups,
                               fe_numbering **out)
{
  if(!m || !kinds || !essentialGroups || !out || nfields < 1) return FE_STATUS_POINTER_ERROR;
  auto groupOf = [&](int b) { return m->bgroup.empty() ? 0 : m->bgroup[b]; };
  for(int f = 0; f < nfields; ++f) {
    int k = kinds[f];
    if(m->dim == 2 && !(k == FE_TRI_P1 || k == FE_TRI_P2 || k == FE_TRI_VP2 || k == FE_TRI_P3 || k == FE_TRI_P4))
      return FE_STATUS_FORMAT_ERROR;
    if(m->dim == 3 && k != FE_TET_P1) return FE_STATUS_FORMAT_ERROR;
  }
  if(fe_mesh_check_jacobians(m, nullptr, nullptr) != FE_STATUS_OK) return FE_STATUS_ERROR;
  fe_numbering *num = new(std::nothrow) fe_numbering;
  if(!num) return FE_STATUS_OUT_OF_MEMORY;
  num->nfields = nfields;
  num->E = m->E;
  num->kinds.assign(kinds, kinds + nfields);
  const int UNK = -1, ESS = -2;
  std::vector<std::vector<int>> vnum(nfields), enm(nfields), inm(nfields);
  std::vector<int> nvd(nfields), ned(nfields), nin(nfields);
  for(int f = 0; f < nfields; ++f) {
    const int k = kinds[f];
    nvd[f] = k == FE_TRI_VP2 ? 2 : 1;
    // edge DOFs (components inner for the vector kind) and element DOFs
    ned[f] = k == FE_TRI_P2 ? 1 : k == FE_TRI_VP2 ? 2 : k == FE_TRI_P3 ? 2 : k == FE_TRI_P4 ? 3 : 0;
    nin[f] = k == FE_TRI_P3 ? 1 : k == FE_TRI_P4 ? 3 : 0;
    vnum[f].assign((size_t)m->V * nvd[f], UNK);
    enm[f].assign((size_t)m->nEdges * ned[f], UNK);
    inm[f].assign((size_t)m->E * nin[f], UNK);
    if(essentialGroups[f]) {
      for(int b = 0; b < m->nB; ++b) {
        const int grp = groupOf(b);
        if(grp >= 32 || !((essentialGroups[f] >> grp) & 1u)) continue;
        for(int j = 0; j < m->bnv; ++j) {
          int v = m->bconn[(size_t)b * m->bnv + j];
          for(int c = 0; c < nvd[f]; ++c) vnum[f][(size_t)v * nvd[f] + c] = ESS;
        }
        if(ned[f] > 0) {
          int ed = m->bedge[b];
          if(ed < 0) { delete num; return FE_STATUS_FORMAT_ERROR; }
          for(int c = 0; c < ned[f]; ++c) enm[f][(size_t)ed * ned[f] + c] = ESS;
        }
      }
    }
  }
  int g = 0;
  for(int f = 0; f < nfields; ++f) {
    for(int &d : vnum[f]) if(d == UNK) d = g++;
    for(int &d : enm[f]) if(d == UNK) d = g++;
    for(int &d : inm[f]) d = g++;  // element DOFs are never on the boundary
  }
  num->nInc = g;
  for(int f = 0; f < nfields; ++f) {
    for(int &d : vnum[f]) if(d == ESS) d = g++;
    if(ned[f] > 0) {
      for(int b = 0; b < m->nB; ++b)
        for(int c = 0; c < ned[f]; ++c) {
          int &d = enm[f][(size_t)m->bedge[b] * ned[f] + c];
          if(d == ESS) d = g++;
        }
      for(int &d : enm[f]) if(d == ESS) d = g++;
    }
  }
  num->nDofs = g;
  num->vnum = vnum;
  num->enm = enm;
  num->nvd = nvd;
  num->ned = ned;
  num->adr.resize(nfields);
  num->dofXyz.assign((size_t)3 * g, 0.);
  num->dofField.assign(g, -1);
  num->dofComp.assign(g, 0);
  for(int f = 0; f < nfields; ++f) {
    const int nl = nlocOf(kinds[f]);
    std::vector<int> &a = num->adr[f];
    a.resize((size_t)m->E * nl);
    for(int e = 0; e < m->E; ++e) {
      int *ae = a.data() + (size_t)e * nl;
      const int *ce = m->conn.data() + (size_t)e * m->nvpe;
      if(kinds[f] == FE_TRI_P1 || kinds[f] == FE_TET_P1) {
        for(int j = 0; j < nl; ++j) ae[j] = vnum[f][ce[j]];
      } else if(kinds[f] == FE_TRI_P2) {
        for(int j = 0; j < 3; ++j) ae[j] = vnum[f][ce[j]];
        for(int k = 0; k < 3; ++k) ae[3 + k] = enm[f][std::abs(m->edges[3 * e + k]) - 1];
      } else if(kinds[f] == FE_TRI_P3 || kinds[f] == FE_TRI_P4) {
        // an edge's DOFs run from its first vertex (as the first element
        // that meets it walks it); an element walking it the other way takes
        // them in reverse
        const int pe = ned[f];
        for(int j = 0; j < 3; ++j) ae[j] = vnum[f][ce[j]];
        for(int k = 0; k < 3; ++k) {
          const int ed = std::abs(m->edges[3 * e + k]) - 1;
          const bool fwd = m->edgeA[ed] == ce[k];
          for(int i = 0; i < pe; ++i) ae[3 + pe * k + i] = enm[f][(size_t)ed * pe + (fwd ? i : pe - 1 - i)];
        }
        for(int i = 0; i < nin[f]; ++i) ae[3 + 3 * pe + i] = inm[f][(size_t)e * nin[f] + i];
      } else {
        for(int j = 0; j < 3; ++j)
          for(int c = 0; c < 2; ++c) ae[2 * j + c] = vnum[f][(size_t)ce[j] * 2 + c];
        for(int k = 0; k < 3; ++k)
          for(int c = 0; c < 2; ++c)
            ae[6 + 2 * k + c] = enm[f][(size_t)(std::abs(m->edges[3 * e + k]) - 1) * 2 + c];
      }
      // DOF locations: vertices, and edge midpoints (0.5 x_a + 0.5 x_b, the
      // P1 geometric interpolation at the edge DOF's reference point).
      for(int j = 0; j < nl; ++j) {
        int d = ae[j];
        int comp = (kinds[f] == FE_TRI_VP2) ? (j % 2) : 0;
        int node = (kinds[f] == FE_TRI_VP2) ? j / 2 : j;
        double x[3];
        if(kinds[f] == FE_TRI_P3 || kinds[f] == FE_TRI_P4) {
          double r, t;
          feTriNodeRS(kinds[f], j, r, t);
          for(int k = 0; k < 3; ++k) {
            const double x0 = m->xyz[3 * (size_t)ce[0] + k];
            x[k] = x0 + r * (m->xyz[3 * (size_t)ce[1] + k] - x0) + t * (m->xyz[3 * (size_t)ce[2] + k] - x0);
          }
        } else if(node < m->nvpe) {
          for(int k = 0; k < 3; ++k) x[k] = m->xyz[3 * (size_t)ce[node] + k];
        } else {
          int k0 = node - 3, va = ce[k0], vb = ce[(k0 + 1) % 3];
          for(int k = 0; k < 3; ++k) x[k] = 0.5 * m->xyz[3 * (size_t)va + k] + 0.5 * m->xyz[3 * (size_t)vb + k];
        }
        for(int k = 0; k < 3; ++k) num->dofXyz[3 * (size_t)d + k] = x[k];
        num->dofField[d] = f;
        num->dofComp[d] = comp;
      }
    }
  }
  *out = num;
  return FE_STATUS_OK;
}

int fe_numbering_periodic(const fe_mesh *m, const fe_numbering *num, int field, int masterGroup, int slaveGroup,
                          const double *offset, int *master, int *slave, int cap, int *count)
{
  if(!m || !num || !offset || !count || field < 0 || field >= num->nfields) return FE_STATUS_POINTER_ERROR;
  if(m->dim != 2 || num->vnum.empty()) return FE_STATUS_FORMAT_ERROR;
  const int ngr = m->groups.empty() ? 1 : (int)m->groups.size();
  if(masterGroup < 0 || masterGroup >= ngr || slaveGroup < 0 || slaveGroup >= ngr) return FE_STATUS_POINTER_ERROR;
  const double tol = 1e-8;  // the reference's matching tolerance
  auto groupOf = [&](int b) { return m->bgroup.empty() ? 0 : m->bgroup[b]; };
  auto X = [&](int v, int k) { return m->xyz[3 * (size_t)v + k]; };
  auto near = [&](int v0, int v1) {  // x(v1) = x(v0) + offset
    double d = 0.;
    for(int k = 0; k < 3; ++k) d += (X(v1, k) - X(v0, k) - offset[k]) * (X(v1, k) - X(v0, k) - offset[k]);
    return std::sqrt(d) < tol;
  };
  // vertices and edges of each group, in boundary-element order (unique)
  std::vector<int> v0, v1, e0, e1;
  std::vector<char> seenV(m->V, 0), seenE(m->nEdges, 0);
  for(int pass = 0; pass < 2; ++pass) {
    std::fill(seenV.begin(), seenV.end(), 0);
    std::fill(seenE.begin(), seenE.end(), 0);
    const int grp = pass == 0 ? masterGroup : slaveGroup;
    std::vector<int> &vs = pass == 0 ? v0 : v1, &es = pass == 0 ? e0 : e1;
    for(int b = 0; b < m->nB; ++b) {
      if(groupOf(b) != grp) continue;
      for(int j = 0; j < m->bnv; ++j) {
        const int v = m->bconn[(size_t)b * m->bnv + j];
        if(!seenV[v]) { seenV[v] = 1; vs.push_back(v); }
      }
      const int ed = m->bedge[b];
      if(ed >= 0 && !seenE[ed]) { seenE[ed] = 1; es.push_back(ed); }
    }
  }
  const int k = num->kinds[field], nvd = num->nvd[field], ned = num->ned[field];
  const bool vec = k == FE_TRI_VP2;  // edge DOFs are components, not positions
  const std::vector<int> &vn = num->vnum[field], &en = num->enm[field];
  std::map<int, int> pairs;  // master -> slave, the first pair of a master kept
  int nv = 0, ne = 0;
  for(int a : v0)
    for(int c : v1)
      if(near(a, c)) {
        ++nv;
        for(int i = 0; i < nvd; ++i) pairs.insert({vn[(size_t)a * nvd + i], vn[(size_t)c * nvd + i]});
        break;
      }
  for(int a : e0) {
    // the edge's two vertices (its first vertex, as its DOFs run, first)
    const int a0 = m->edgeA[a];
    int a1 = -1;
    for(int b = 0; b < m->nB && a1 < 0; ++b)
      if(m->bedge[b] == a)
        for(int j = 0; j < 2; ++j)
          if(m->bconn[(size_t)b * m->bnv + j] != a0) a1 = m->bconn[(size_t)b * m->bnv + j];
    for(int c : e1) {
      const int c0 = m->edgeA[c];
      int c1 = -1;
      for(int b = 0; b < m->nB && c1 < 0; ++b)
        if(m->bedge[b] == c)
          for(int j = 0; j < 2; ++j)
            if(m->bconn[(size_t)b * m->bnv + j] != c0) c1 = m->bconn[(size_t)b * m->bnv + j];
      const bool same = near(a0, c0) && near(a1, c1), rev = near(a0, c1) && near(a1, c0);
      if(!same && !rev) continue;
      ++ne;
      for(int i = 0; i < ned; ++i) {
        const int j = (vec || same) ? i : ned - 1 - i;  // positions matched along the edge
        pairs.insert({en[(size_t)a * ned + i], en[(size_t)c * ned + j]});
      }
      break;
    }
  }
  if(nv == 0 && ne == 0) return FE_STATUS_ERROR;  // no matching vertex or edge (feNumber.cpp:299-306)
  *count = (int)pairs.size();
  if(master && slave) {
    if((int)pairs.size() > cap) return FE_STATUS_ERROR;
    int q = 0;
    for(const auto &p : pairs) {
      master[q] = p.first;
      slave[q] = p.second;
      ++q;
    }
  }
  return FE_STATUS_OK;
}

int fe_numbering_destroy(fe_numbering *num)
{
  delete num;
  return FE_STATUS_OK;
}

int fe_numbering_info(const fe_numbering *num, int *nInc, int *nDofs)
{
  if(!num) return FE_STATUS_POINTER_ERROR;
  if(nInc) *nInc = num->nInc;
  if(nDofs) *nDofs = num->nDofs;
  return FE_STATUS_OK;
}

int fe_numbering_adr(const fe_numbering *num, int field, int *adr)
{
  if(!num || !adr) return FE_STATUS_POINTER_ERROR;
  if(field < 0 || field >= num->nfields) return FE_STATUS_RANGE_ERROR;
  std::copy(num->adr[field].begin(), num->adr[field].end(), adr);
  return FE_STATUS_OK;
}

int fe_numbering_dof_info(const fe_numbering *num, double *xyz, int *field, int *component)
{
  if(!num) return FE_STATUS_POINTER_ERROR;
  if(xyz) std::copy(num->dofXyz.begin(), num->dofXyz.end(), xyz);
  if(field) std::copy(num->dofField.begin(), num->dofField.end(), field);
  if(component) std::copy(num->dofComp.begin(), num->dofComp.end(), component);
  return FE_STATUS_OK;
}

// ---------------------------------------------------------------------------
// First-fit greedy element coloring over vertex adjacency: an element takes
// the lowest color not yet used by any element sharing one of its vertices.
int fe_color_elements(const fe_mesh *m, int *colorOf, int *ncolors)
{
  if(!m || !colorOf) return FE_STATUS_POINTER_ERROR;
  const int W = 4;  // 256 colors max
  std::vector<uint64_t> used((size_t)m->V * W, 0);
  int nc = 0;
  for(int e = 0; e < m->E; ++e) {
    uint64_t forb[W] = {0, 0, 0, 0};
    const int *ce = m->conn.data() + (size_t)e * m->nvpe;
    for(int j = 0; j < m->nvpe; ++j)
      for(int w = 0; w < W; ++w) forb[w] |= used[(size_t)ce[j] * W + w];
    int c = -1;
    for(int w = 0; w < W && c < 0; ++w)
      if(~forb[w]) c = 64 * w + __builtin_ctzll(~forb[w]);
    if(c < 0) return FE_STATUS_RANGE_ERROR;
    colorOf[e] = c;
    nc = std::max(nc, c + 1);
    for(int j = 0; j < m->nvpe; ++j) used[(size_t)ce[j] * W + c / 64] |= (1ull << (c % 64));
  }
  if(ncolors) *ncolors = nc;
  return FE_STATUS_OK;
}

// ---------------------------------------------------------------------------
int fe_coeff_blocks(const fe_mesh *m, int B, double lo, double hi, uint64_t seed, double *k)
{
  if(!m || !k || B < 1 || lo <= 0. || hi < lo) return FE_STATUS_ERROR;
  SplitMix64 rng(seed);
  std::vector<double> blk((size_t)B * B);
  const double llo = std::log(lo), lhi = std::log(hi);
  for(auto &b : blk) b = std::exp(llo + (lhi - llo) * rng.uniform());
  for(int e = 0; e < m->E; ++e) {
    double cx = 0., cy = 0.;
    for(int j = 0; j < m->nvpe; ++j) {
      int v = m->conn[(size_t)e * m->nvpe + j];
      cx += m->xyz[3 * (size_t)v];
      cy += m->xyz[3 * (size_t)v + 1];
    }
    cx /= m->nvpe;
    cy /= m->nvpe;
    int bi = std::min(B - 1, std::max(0, (int)std::floor(cx * B)));
    int bj = std::min(B - 1, std::max(0, (int)std::floor(cy * B)));
    k[e] = blk[(size_t)bi + (size_t)B * bj];
  }
  return FE_STATUS_OK;
}

// Random Fourier features (Rahimi & Recht 2007) for the squared-exponential
// kernel: g(x) = sigma sqrt(2/M) sum_m cos(w_m . x + b_m), w ~ N(0, I/ell^2),
// b ~ U[0, 2 pi).  Evaluated at the vertices, averaged to the centroid.
int fe_coeff_grf_cube(const fe_mesh *m, double ell, double sigma, int nfeat, uint64_t seed,
                      double *k)
{
  if(!m || !k || m->dim != 3 || ell <= 0. || nfeat < 1) return FE_STATUS_ERROR;
  SplitMix64 rng(seed);
  std::vector<double> w((size_t)3 * nfeat), b(nfeat);
  for(int f = 0; f < nfeat; ++f) {
    for(int d = 0; d < 3; ++d) w[3 * f + d] = rng.normal() / ell;
    b[f] = 2. * M_PI * rng.uniform();
  }
  const double scale = sigma * std::sqrt(2.0 / nfeat);
  std::vector<double> g(m->V);
#pragma omp parallel for schedule(static)
  for(int v = 0; v < m->V; ++v) {
    const double *x = &m->xyz[3 * (size_t)v];
    double s = 0.;
    for(int f = 0; f < nfeat; ++f)
      s += std::cos(w[3 * f] * x[0] + w[3 * f + 1] * x[1] + w[3 * f + 2] * x[2] + b[f]);
    g[v] = scale * s;
  }
  for(int e = 0; e < m->E; ++e) {
    double s = 0.;
    for(int j = 0; j < 4; ++j) s += g[m->conn[(size_t)e * 4 + j]];
    k[e] = std::exp(0.25 * s);
  }
  return FE_STATUS_OK;
}

}  // extern "C"
