/*
 * oracle_geom.c -- SWG mesh topology (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see buffem_oracle.h).
 *
 * Restates: NewSWGTet (src/libs/libbuff/SWGVolume.cc:50-85),
 * InitSWGFace / InitFaceList (src/libs/libbuff/InitFaceList.cc:41-279),
 * CompareBFs (SWGVolume.cc:295-344), CompareTets (SWGVolume.cc:357-400).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "buffem_oracle.h"

typedef struct fnode {
  int v1, v2, v3, qp, qm, pt, mt, pi, mi, index;
  struct fnode *next;
} fnode;

int orc_geom_build(int nv, const double *V, int nt, const int *tetVI, orc_geom **out)
{
  *out = NULL;
  if (nv <= 0 || nt <= 0) return -1;
  for (int i = 0; i < 4 * nt; i++)
    if (tetVI[i] < 0 || tetVI[i] >= nv) return -2;
  orc_geom *g = (orc_geom *)calloc(1, sizeof *g);
  g->nv = nv; g->nt = nt;
  g->V = (double *)malloc(3 * sizeof(double) * nv);
  memcpy(g->V, V, 3 * sizeof(double) * nv);
  g->tetVI = (int *)malloc(4 * sizeof(int) * nt);
  memcpy(g->tetVI, tetVI, 4 * sizeof(int) * nt);
  g->tetFI = (int *)malloc(4 * sizeof(int) * nt);
  g->tetVol = (double *)malloc(sizeof(double) * nt);
  g->tetCen = (double *)malloc(3 * sizeof(double) * nt);

  /* NewSWGTet: centroid and signed volume (SWGVolume.cc:63-79) */
  for (int t = 0; t < nt; t++) {
    const double *P[4];
    for (int i = 0; i < 4; i++) P[i] = V + 3 * tetVI[4 * t + i];
    double A[3], B[3], C[3];
    for (int i = 0; i < 3; i++) {
      g->tetCen[3 * t + i] = 0.25 * (P[0][i] + P[1][i] + P[2][i] + P[3][i]);
      A[i] = P[1][i] - P[0][i]; B[i] = P[2][i] - P[0][i]; C[i] = P[3][i] - P[0][i];
    }
    g->tetVol[t] = (A[0] * (B[1] * C[2] - B[2] * C[1]) + A[1] * (B[2] * C[0] - B[0] * C[2])
                  + A[2] * (B[0] * C[1] - B[1] * C[0])) / 6.0;
    for (int i = 0; i < 4; i++) g->tetFI[4 * t + i] = -1;
  }

  /* InitFaceList: faces chained per minimum vertex (InitFaceList.cc:141-230) */
  fnode **lists = (fnode **)calloc(nv, sizeof(fnode *));
  fnode *pool = (fnode *)malloc(sizeof(fnode) * 4 * nt);
  int npool = 0, nint = 0, ntot = 0;
  for (int t = 0; t < nt; t++) {
    for (int nf = 0; nf < 4; nf++) {
      int vi[3] = {tetVI[4 * t + (nf + 1) % 4], tetVI[4 * t + (nf + 2) % 4], tetVI[4 * t + (nf + 3) % 4]};
      int imin = 0, imax = 0;
      if (vi[1] < vi[imin]) imin = 1;
      if (vi[2] < vi[imin]) imin = 2;
      if (vi[1] > vi[imax]) imax = 1;
      if (vi[2] > vi[imax]) imax = 2;
      int vmin = vi[imin], vmid = vi[3 - imin - imax], vmax = vi[imax];
      int q = tetVI[4 * t + nf];
      fnode *F;
      for (F = lists[vmin]; F; F = F->next)
        if (F->v2 == vmid && F->v3 == vmax) break;
      if (F) {
        if (F->mt != -1) { free(lists); free(pool); orc_geom_free(g); return -3; }
        F->qm = q; F->mt = t; F->mi = nf; F->index = nint++;
      } else {
        ntot++;
        F = &pool[npool++];
        F->next = lists[vmin]; lists[vmin] = F;
        F->v1 = vmin; F->v2 = vmid; F->v3 = vmax;
        F->qp = q; F->qm = -1; F->pt = t; F->mt = -1; F->pi = nf; F->mi = -1; F->index = -1;
      }
    }
  }
  g->nf_int = nint; g->nf_tot = ntot;
  int nf = ntot;
  g->fV = (int *)malloc(3 * sizeof(int) * nf);
  g->fQP = (int *)malloc(sizeof(int) * nf); g->fQM = (int *)malloc(sizeof(int) * nf);
  g->fPT = (int *)malloc(sizeof(int) * nf); g->fMT = (int *)malloc(sizeof(int) * nf);
  g->fPI = (int *)malloc(sizeof(int) * nf); g->fMI = (int *)malloc(sizeof(int) * nf);
  g->fArea = (double *)malloc(sizeof(double) * nf);
  g->fCen = (double *)malloc(3 * sizeof(double) * nf);
  g->fRad = (double *)malloc(sizeof(double) * nf);
  /* numbering: interior faces first, exterior in list order (:243-260) */
  int next = 0;
  for (int v = 0; v < nv; v++)
    for (fnode *F = lists[v]; F; F = F->next) {
      if (F->index == -1) F->index = nint + next++;
      int k = F->index;
      g->fV[3 * k] = F->v1; g->fV[3 * k + 1] = F->v2; g->fV[3 * k + 2] = F->v3;
      g->fQP[k] = F->qp; g->fQM[k] = F->qm; g->fPT[k] = F->pt; g->fMT[k] = F->mt;
      g->fPI[k] = F->pi; g->fMI[k] = F->mi;
      g->tetFI[4 * F->pt + F->pi] = k;
      if (F->mt != -1) g->tetFI[4 * F->mt + F->mi] = k;
      /* InitSWGFace (InitFaceList.cc:41-61) */
      const double *V1 = V + 3 * F->v1, *V2 = V + 3 * F->v2, *V3 = V + 3 * F->v3;
      double A[3], B[3], X[3];
      for (int i = 0; i < 3; i++) {
        g->fCen[3 * k + i] = (V1[i] + V2[i] + V3[i]) / 3.0;
        A[i] = V2[i] - V1[i]; B[i] = V3[i] - V1[i];
      }
      X[0] = A[1] * B[2] - A[2] * B[1]; X[1] = A[2] * B[0] - A[0] * B[2]; X[2] = A[0] * B[1] - A[1] * B[0];
      g->fArea[k] = 0.5 * sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
      /* bounding radius: max distance from centroid to the 5 BF vertices (:195-200) */
      double r = 0.0;
      int vs[5] = {F->qp, F->v1, F->v2, F->v3, F->qm};
      for (int j = 0; j < 5; j++) {
        if (vs[j] < 0) continue;
        const double *P = V + 3 * vs[j];
        double d = sqrt((P[0] - g->fCen[3 * k]) * (P[0] - g->fCen[3 * k]) +
                        (P[1] - g->fCen[3 * k + 1]) * (P[1] - g->fCen[3 * k + 1]) +
                        (P[2] - g->fCen[3 * k + 2]) * (P[2] - g->fCen[3 * k + 2]));
        if (d > r) r = d;
      }
      g->fRad[k] = r;
    }
  free(lists); free(pool);
  *out = g;
  return 0;
}

void orc_geom_free(orc_geom *g)
{
  if (!g) return;
  free(g->V); free(g->tetVI); free(g->tetFI); free(g->tetVol); free(g->tetCen);
  free(g->fV); free(g->fQP); free(g->fQM); free(g->fPT); free(g->fMT); free(g->fPI);
  free(g->fMI); free(g->fArea); free(g->fCen); free(g->fRad);
  free(g);
}

int orc_geom_info(const orc_geom *g, int *nv, int *nt, int *nf_int, int *nf_tot)
{
  if (nv) *nv = g->nv;
  if (nt) *nt = g->nt;
  if (nf_int) *nf_int = g->nf_int;
  if (nf_tot) *nf_tot = g->nf_tot;
  return 0;
}

void orc_geom_faces(const orc_geom *g, int *fV, int *fQP, int *fQM, int *fPT,
                    int *fMT, int *fPI, int *fMI, double *fArea, int *tetFI)
{
  int n = g->nf_tot;
  if (fV) memcpy(fV, g->fV, 3 * sizeof(int) * n);
  if (fQP) memcpy(fQP, g->fQP, sizeof(int) * n);
  if (fQM) memcpy(fQM, g->fQM, sizeof(int) * n);
  if (fPT) memcpy(fPT, g->fPT, sizeof(int) * n);
  if (fMT) memcpy(fMT, g->fMT, sizeof(int) * n);
  if (fPI) memcpy(fPI, g->fPI, sizeof(int) * n);
  if (fMI) memcpy(fMI, g->fMI, sizeof(int) * n);
  if (fArea) memcpy(fArea, g->fArea, sizeof(double) * n);
  if (tetFI) memcpy(tetFI, g->tetFI, 4 * sizeof(int) * g->nt);
}

/* CompareBFs (SWGVolume.cc:295-344): 25 vertex-index comparisons */
int orc_compare_bfs(const orc_geom *g, int a, int b)
{
  int va[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
  int vb[5] = {g->fQP[b], g->fV[3 * b], g->fV[3 * b + 1], g->fV[3 * b + 2], g->fQM[b]};
  int n = 0;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++)
      if (va[i] == vb[j]) n++;
  return n;
}

/* CompareTets (SWGVolume.cc:357-400): common vertices first, in TA order */
int orc_compare_tets(const orc_geom *g, int ta, int tb, int *OVIA, int *OVIB)
{
  const int *A = g->tetVI + 4 * ta, *B = g->tetVI + 4 * tb;
  int ca[4] = {0, 0, 0, 0}, cb[4] = {0, 0, 0, 0}, n = 0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      if (A[i] == B[j]) { ca[i] = cb[j] = 1; n++; }
  if (!OVIA || !OVIB) return n;
  int na = 0, nb = 0;
  for (int i = 0; i < 4; i++)
    if (ca[i]) { OVIA[na++] = A[i]; OVIB[nb++] = A[i]; }
  for (int i = 0; i < 4; i++) {
    if (!ca[i]) OVIA[na++] = A[i];
    if (!cb[i]) OVIB[nb++] = B[i];
  }
  return n;
}
