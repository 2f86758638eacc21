// bfa_host.cpp -- host-side mesh construction and SWG topology.
//
// Reference behaviour followed (HomerReid/buff-em):
//   ReadGMSHFile            src/libs/libbuff/ReadGMSHFile.cc:66-261
//   NewSWGTet               src/libs/libbuff/SWGVolume.cc:50-85
//   InitSWGFace/InitFaceList src/libs/libbuff/InitFaceList.cc:41-279
//   CompareBFs / CompareTets src/libs/libbuff/SWGVolume.cc:295-400
// The Kuhn-sphere generator builds the synthetic meshes of the benchmark
// workloads (no reference counterpart: the reference reads GMSH files).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <unordered_map>

#include "bfa_internal.h"

namespace bfa {

static thread_local std::string g_err;

int fail(int status, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return status;
}

static uint64_t splitmix64(uint64_t &s)
{
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static double unit01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double signed_volume(const double *V, const int *vi)
{
  const double *P0 = V + 3 * vi[0], *P1 = V + 3 * vi[1], *P2 = V + 3 * vi[2], *P3 = V + 3 * vi[3];
  double A[3], B[3], C[3];
  for (int i = 0; i < 3; i++) { A[i] = P1[i] - P0[i]; B[i] = P2[i] - P0[i]; C[i] = P3[i] - P0[i]; }
  return (A[0] * (B[1] * C[2] - B[2] * C[1]) + A[1] * (B[2] * C[0] - B[0] * C[2])
          + A[2] * (B[0] * C[1] - B[1] * C[0])) / 6.0;
}

}  // namespace bfa

using namespace bfa;

extern "C" const char *bfa_last_error(void) { return g_err.c_str(); }

extern "C" const char *bfa_status_string(int status)
{
  switch (status) {
    case BFA_OK: return "ok";
    case BFA_ERR_ARG: return "invalid argument";
    case BFA_ERR_MESH: return "invalid mesh";
    case BFA_ERR_SINGULAR: return "singular matrix";
    case BFA_ERR_OOM: return "out of memory";
    case BFA_ERR_HIP: return "HIP runtime error";
    case BFA_ERR_STATE: return "invalid call order";
    case BFA_ERR_IO: return "I/O error";
  }
  return "unknown status";
}

extern "C" int bfa_version(void) { return 1; }

extern "C" int bfa_mesh_kuhn_sphere(int n, double radius, double jitter, uint64_t seed,
                                    int *nv, int *nt, double *V, int *tetVI)
{
  if (n < 1 || n > 200 || !(radius > 0.0) || jitter < 0.0 || jitter >= 0.25 || !nv || !nt)
    return fail(BFA_ERR_ARG, "bfa_mesh_kuhn_sphere: bad arguments (n=%d R=%g jitter=%g)", n, radius, jitter);
  const int np = n + 1;
  *nv = np * np * np;
  *nt = 6 * n * n * n;
  if (!V || !tetVI) return BFA_OK;
  auto id = [np](int i, int j, int k) { return i + np * (j + np * k); };
  // lattice, radial map x -> R x |x|_inf / |x|_2
  for (int k = 0; k <= n; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        double x[3] = {-1.0 + 2.0 * i / n, -1.0 + 2.0 * j / n, -1.0 + 2.0 * k / n};
        double ninf = std::max(std::fabs(x[0]), std::max(std::fabs(x[1]), std::fabs(x[2])));
        double n2 = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        double s = n2 > 0.0 ? radius * ninf / n2 : 0.0;
        double *P = V + 3 * id(i, j, k);
        for (int d = 0; d < 3; d++) P[d] = s * x[d];
      }
  // jitter interior nodes, node-index order, x then y then z
  uint64_t st = seed;
  const double h = 2.0 * radius / n;
  for (int v = 0; v < *nv; v++) {
    int i = v % np, j = (v / np) % np, k = v / (np * np);
    if (i == 0 || j == 0 || k == 0 || i == n || j == n || k == n) continue;
    for (int d = 0; d < 3; d++) V[3 * v + d] += jitter * h * (2.0 * unit01(st) - 1.0);
  }
  // 6 Kuhn tetrahedra per subcube: monotone lattice paths v000 -> v111
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int t = 0;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
        for (int p = 0; p < 6; p++) {
          int c[3] = {i, j, k};
          int *vi = tetVI + 4 * t;
          vi[0] = id(c[0], c[1], c[2]);
          for (int s = 0; s < 3; s++) {
            c[perms[p][s]]++;
            vi[s + 1] = id(c[0], c[1], c[2]);
          }
          if (signed_volume(V, vi) < 0.0) std::swap(vi[2], vi[3]);
          t++;
        }
  return BFA_OK;
}

// ---- GMSH reader --------------------------------------------------------
extern "C" int bfa_mesh_read_vmsh(const char *path, int *nv, int *nt, double *V, int *tetVI)
{
  if (!path || !nv || !nt) return fail(BFA_ERR_ARG, "bfa_mesh_read_vmsh: null argument");
  FILE *f = fopen(path, "r");
  if (!f) return fail(BFA_ERR_IO, "could not open mesh file %s", path);
  char buf[256];
  bool legacy = false, found = false;
  while (fgets(buf, sizeof buf, f)) {
    if (!strncmp(buf, "$NOD", 4) && strncmp(buf, "$Nodes", 6)) { legacy = true; found = true; break; }
    if (!strncmp(buf, "$Nodes", 6)) { legacy = false; found = true; break; }
  }
  if (!found) { fclose(f); return fail(BFA_ERR_MESH, "%s: failed to find node start keyword", path); }
  int NV = 0;
  if (!fgets(buf, sizeof buf, f) || (NV = atoi(buf)) <= 0) {
    fclose(f); return fail(BFA_ERR_MESH, "%s: invalid number of nodes", path);
  }
  std::vector<double> X(3 * (size_t)NV);
  std::vector<int> g2h(2 * (size_t)NV + 1, -1);
  for (int v = 0; v < NV; v++) {
    int idx;
    if (!fgets(buf, sizeof buf, f) || sscanf(buf, "%d %lf %lf %lf", &idx, &X[3 * v], &X[3 * v + 1], &X[3 * v + 2]) != 4) {
      fclose(f); return fail(BFA_ERR_MESH, "%s: invalid node specification", path);
    }
    if (idx < 0 || idx > 2 * NV) { fclose(f); return fail(BFA_ERR_MESH, "%s: node index out of range", path); }
    g2h[idx] = v;
  }
  // merge vertices closer than 1e-6 (ReadGMSHFile.cc:130-141)
  for (int i = 0; i < NV; i++)
    for (int j = i + 1; j < NV; j++) {
      double dx = X[3 * i] - X[3 * j], dy = X[3 * i + 1] - X[3 * j + 1], dz = X[3 * i + 2] - X[3 * j + 2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) < 1.0e-6)
        for (auto &m : g2h) if (m == j) m = i;
    }
  if (!fgets(buf, sizeof buf, f) || strncmp(buf, legacy ? "$ENDNOD" : "$EndNodes", legacy ? 7 : 9)) {
    fclose(f); return fail(BFA_ERR_MESH, "%s: nodes section not terminated", path);
  }
  if (!fgets(buf, sizeof buf, f) || strncmp(buf, legacy ? "$ELM" : "$Elements", legacy ? 4 : 9)) {
    fclose(f); return fail(BFA_ERR_MESH, "%s: elements section not initiated", path);
  }
  int NE = 0;
  if (!fgets(buf, sizeof buf, f) || sscanf(buf, "%d", &NE) != 1 || NE < 0) {
    fclose(f); return fail(BFA_ERR_MESH, "%s: invalid number of elements", path);
  }
  std::vector<int> tets;
  for (int e = 0; e < NE; e++) {
    if (!fgets(buf, sizeof buf, f)) { fclose(f); return fail(BFA_ERR_MESH, "%s: too few elements", path); }
    int type = 0, vi[4] = {0, 0, 0, 0};
    if (legacy) {
      int num, rp, re, nn;
      if (sscanf(buf, "%d %d %d %d %d %d %d %d %d", &num, &type, &rp, &re, &nn, vi, vi + 1, vi + 2, vi + 3) < 5) {
        fclose(f); return fail(BFA_ERR_MESH, "%s: invalid element", path);
      }
    } else {
      int num, ntags, pos = 0, rd = 0;
      if (sscanf(buf, "%d %d %d%n", &num, &type, &ntags, &rd) < 3) {
        fclose(f); return fail(BFA_ERR_MESH, "%s: invalid element", path);
      }
      pos = rd;
      for (int t = 0; t < ntags; t++) { int dummy; sscanf(buf + pos, "%d%n", &dummy, &rd); pos += rd; }
      int nc = sscanf(buf + pos, "%d %d %d %d", vi, vi + 1, vi + 2, vi + 3);
      if (type == 4 && nc != 4) { fclose(f); return fail(BFA_ERR_MESH, "%s: invalid tetrahedron", path); }
    }
    if (type != 4) continue;
    for (int q = 0; q < 4; q++) {
      if (vi[q] < 0 || vi[q] > 2 * NV || g2h[vi[q]] < 0) {
        fclose(f); return fail(BFA_ERR_MESH, "%s: element references unknown node", path);
      }
      tets.push_back(g2h[vi[q]]);
    }
  }
  fclose(f);
  if (tets.empty()) return fail(BFA_ERR_MESH, "%s: no tetrahedra found", path);
  *nv = NV;
  *nt = (int)(tets.size() / 4);
  if (V) memcpy(V, X.data(), sizeof(double) * 3 * NV);
  if (tetVI) memcpy(tetVI, tets.data(), sizeof(int) * tets.size());
  return BFA_OK;
}

// ---- topology --------------------------------------------------------------
static int count_common(const int *a, const int *b)
{
  int n = 0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) n += (a[i] == b[j]);
  return n;
}

// CompareTets ordering: common vertices first (in A's order), then the
// non-common vertices of each tetrahedron in its own order.
static void order_common(const int *A, const int *B, int *OA, int *OB)
{
  bool ca[4] = {false, false, false, false}, cb[4] = {false, false, false, false};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      if (A[i] == B[j]) ca[i] = cb[j] = true;
  int na = 0, nb = 0;
  for (int i = 0; i < 4; i++)
    if (ca[i]) { OA[na++] = A[i]; OB[nb++] = A[i]; }
  for (int i = 0; i < 4; i++) {
    if (!ca[i]) OA[na++] = A[i];
    if (!cb[i]) OB[nb++] = B[i];
  }
}

extern "C" int bfa_geometry_create(int nv, const double *V, int nt, const int *tetVI, bfa_geometry **out)
{
  if (!out) return fail(BFA_ERR_ARG, "bfa_geometry_create: null output");
  *out = nullptr;
  if (nv <= 0 || nt <= 0 || !V || !tetVI) return fail(BFA_ERR_ARG, "bfa_geometry_create: empty mesh");
  for (int i = 0; i < 4 * nt; i++)
    if (tetVI[i] < 0 || tetVI[i] >= nv) return fail(BFA_ERR_MESH, "tetrahedron %d references vertex %d", i / 4, tetVI[i]);
  bfa_geometry *g = new bfa_geometry();
  g->nv = nv; g->nt = nt;
  g->V.assign(V, V + 3 * (size_t)nv);
  g->tetVI.assign(tetVI, tetVI + 4 * (size_t)nt);
  g->tetFI.assign(4 * (size_t)nt, -1);
  g->tetVol.resize(nt);
  g->tetCen.resize(3 * (size_t)nt);
  for (int t = 0; t < nt; t++) {
    const int *vi = tetVI + 4 * t;
    g->tetVol[t] = signed_volume(V, vi);
    for (int d = 0; d < 3; d++)
      g->tetCen[3 * t + d] = 0.25 * (V[3 * vi[0] + d] + V[3 * vi[1] + d] + V[3 * vi[2] + d] + V[3 * vi[3] + d]);
  }
  // faces keyed by minimum vertex, created in visit order
  struct F { int v1, v2, v3, qp, qm, pt, mt, pi, mi, index; };
  std::vector<F> faces;
  faces.reserve(2 * (size_t)nt + 16);
  std::vector<std::vector<int>> bymin(nv);
  int nint = 0;
  for (int t = 0; t < nt; t++)
    for (int nf = 0; nf < 4; nf++) {
      const int *vi = tetVI + 4 * t;
      int a[3] = {vi[(nf + 1) % 4], vi[(nf + 2) % 4], vi[(nf + 3) % 4]};
      std::sort(a, a + 3);
      int q = vi[nf];
      int found = -1;
      for (int fi : bymin[a[0]])
        if (faces[fi].v2 == a[1] && faces[fi].v3 == a[2]) { found = fi; break; }
      if (found >= 0) {
        F &f = faces[found];
        if (f.mt != -1) {
          delete g;
          return fail(BFA_ERR_MESH, "invalid mesh topology: face %d of tet %d also belongs to tets %d and %d",
                      nf, t, f.pt, f.mt);
        }
        f.qm = q; f.mt = t; f.mi = nf; f.index = nint++;
      } else {
        faces.push_back(F{a[0], a[1], a[2], q, -1, t, -1, nf, -1, -1});
        bymin[a[0]].push_back((int)faces.size() - 1);
      }
    }
  const int ntot = (int)faces.size();
  g->nbf = nint; g->nftot = ntot;
  g->fV.resize(3 * (size_t)ntot); g->fQP.resize(ntot); g->fQM.resize(ntot);
  g->fPT.resize(ntot); g->fMT.resize(ntot); g->fPI.resize(ntot); g->fMI.resize(ntot);
  g->fArea.resize(ntot); g->fCen.resize(3 * (size_t)ntot); g->fRad.assign(ntot, 0.0);
  int next = 0;
  for (int v = 0; v < nv; v++)
    for (auto it = bymin[v].rbegin(); it != bymin[v].rend(); ++it) {   // list order = LIFO
      F &f = faces[*it];
      if (f.index == -1) f.index = nint + next++;
      int k = f.index;
      g->fV[3 * k] = f.v1; g->fV[3 * k + 1] = f.v2; g->fV[3 * k + 2] = f.v3;
      g->fQP[k] = f.qp; g->fQM[k] = f.qm; g->fPT[k] = f.pt; g->fMT[k] = f.mt;
      g->fPI[k] = f.pi; g->fMI[k] = f.mi;
      g->tetFI[4 * f.pt + f.pi] = k;
      if (f.mt >= 0) g->tetFI[4 * f.mt + f.mi] = k;
      const double *P1 = V + 3 * f.v1, *P2 = V + 3 * f.v2, *P3 = V + 3 * f.v3;
      double A[3], B[3];
      for (int d = 0; d < 3; d++) {
        g->fCen[3 * k + d] = (P1[d] + P2[d] + P3[d]) / 3.0;
        A[d] = P2[d] - P1[d]; B[d] = P3[d] - P1[d];
      }
      double X = A[1] * B[2] - A[2] * B[1], Y = A[2] * B[0] - A[0] * B[2], Z = A[0] * B[1] - A[1] * B[0];
      g->fArea[k] = 0.5 * std::sqrt(X * X + Y * Y + Z * Z);
      if (f.mt >= 0) {
        int vs[5] = {f.qp, f.v1, f.v2, f.v3, f.qm};
        double r = 0.0;
        for (int s = 0; s < 5; s++) {
          const double *P = V + 3 * vs[s];
          double dx = P[0] - g->fCen[3 * k], dy = P[1] - g->fCen[3 * k + 1], dz = P[2] - g->fCen[3 * k + 2];
          r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        g->fRad[k] = r;
      }
    }
  if (nint == 0) {
    delete g;
    return fail(BFA_ERR_MESH, "mesh has no interior faces");
  }

  // touching BF pairs a <= b (CompareBFs > 0), via vertex incidence
  const int N = nint;
  std::vector<std::vector<int>> vbf(nv);
  for (int a = 0; a < N; a++) {
    int vs[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
    for (int s = 0; s < 5; s++) vbf[vs[s]].push_back(a);
  }
  g->rowptr.assign(N + 1, 0);
  std::vector<char> mark(N, 0);
  std::vector<int> tmp;
  for (int a = 0; a < N; a++) {
    tmp.clear();
    int vs[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
    for (int s = 0; s < 5; s++)
      for (int b : vbf[vs[s]])
        if (b >= a && !mark[b]) { mark[b] = 1; tmp.push_back(b); }
    std::sort(tmp.begin(), tmp.end());
    for (int b : tmp) { mark[b] = 0; g->cols.push_back(b); }
    g->rowptr[a + 1] = (int)g->cols.size();
  }

  // Taylor-Duffy tetrahedron pairs (>= 2 shared vertices) of touching pairs
  const int64_t nnz = (int64_t)g->cols.size();
  g->pairTD.assign(4 * (size_t)nnz, -1);
  std::unordered_map<uint64_t, int> tdmap;
  tdmap.reserve((size_t)nt * 16);
  for (int a = 0; a < N; a++)
    for (int j = g->rowptr[a]; j < g->rowptr[a + 1]; j++) {
      int b = g->cols[j];
      for (int sa = 0; sa < 2; sa++)
        for (int sb = 0; sb < 2; sb++) {
          int ta = sa == 0 ? g->fPT[a] : g->fMT[a];
          int tb = sb == 0 ? g->fPT[b] : g->fMT[b];
          int ncv = count_common(&g->tetVI[4 * ta], &g->tetVI[4 * tb]);
          if (ncv < 2) continue;
          int lo = std::min(ta, tb), hi = std::max(ta, tb);
          uint64_t key = ((uint64_t)lo << 32) | (uint32_t)hi;
          auto it = tdmap.find(key);
          int idx;
          if (it == tdmap.end()) {
            idx = (int)g->tdA.size();
            tdmap.emplace(key, idx);
            g->tdA.push_back(lo); g->tdB.push_back(hi); g->tdNcv.push_back(ncv);
            int OA[4], OB[4];
            order_common(&g->tetVI[4 * lo], &g->tetVI[4 * hi], OA, OB);
            for (int q = 0; q < 4; q++) g->tdOV.push_back(OA[q]);
            for (int q = 0; q < 4; q++) g->tdOV.push_back(OB[q]);
          } else {
            idx = it->second;
          }
          g->pairTD[4 * (size_t)j + 2 * sa + sb] = 2 * idx + (ta > tb ? 1 : 0);
        }
    }
  *out = g;
  return BFA_OK;
}

extern "C" int bfa_geometry_info(const bfa_geometry *g, int *nv, int *nt, int *nbf, int *nfaces_total)
{
  if (!g) return fail(BFA_ERR_ARG, "null geometry");
  if (nv) *nv = g->nv;
  if (nt) *nt = g->nt;
  if (nbf) *nbf = g->nbf;
  if (nfaces_total) *nfaces_total = g->nftot;
  return BFA_OK;
}

extern "C" int bfa_geometry_faces(const bfa_geometry *g, int *fV, int *fQP, int *fQM, int *fPT,
                                  int *fMT, int *fPI, int *fMI, double *fArea)
{
  if (!g) return fail(BFA_ERR_ARG, "null geometry");
  size_t n = (size_t)g->nftot;
  if (fV) memcpy(fV, g->fV.data(), 3 * n * sizeof(int));
  if (fQP) memcpy(fQP, g->fQP.data(), n * sizeof(int));
  if (fQM) memcpy(fQM, g->fQM.data(), n * sizeof(int));
  if (fPT) memcpy(fPT, g->fPT.data(), n * sizeof(int));
  if (fMT) memcpy(fMT, g->fMT.data(), n * sizeof(int));
  if (fPI) memcpy(fPI, g->fPI.data(), n * sizeof(int));
  if (fMI) memcpy(fMI, g->fMI.data(), n * sizeof(int));
  if (fArea) memcpy(fArea, g->fArea.data(), n * sizeof(double));
  return BFA_OK;
}

extern "C" int bfa_geometry_touch_count(const bfa_geometry *g, int64_t *nnz)
{
  if (!g || !nnz) return fail(BFA_ERR_ARG, "null argument");
  *nnz = (int64_t)g->cols.size();
  return BFA_OK;
}

extern "C" int bfa_geometry_touch_pairs(const bfa_geometry *g, int *rowptr, int *cols)
{
  if (!g) return fail(BFA_ERR_ARG, "null geometry");
  if (rowptr) memcpy(rowptr, g->rowptr.data(), g->rowptr.size() * sizeof(int));
  if (cols) memcpy(cols, g->cols.data(), g->cols.size() * sizeof(int));
  return BFA_OK;
}
