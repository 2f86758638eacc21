/*
 * oracle_vie.c -- VIE matrix assembly, RHS and EMT power (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see buffem_oracle.h).
 *
 * Restates, from the reference HomerReid/buff-em:
 *   TetCR4/16/33 rules            src/libs/libbuff/TetCR.cc:35-109
 *   TetInt/BFInt/TetTetInt/BFBFInt (fixed-rule branches)
 *                                 src/libs/libbuff/Cubature.cc:161-527
 *   GetExpRelTable                src/libs/libbuff/GMatrixElements.cc:59-93
 *   GMEIntegrand                  GMatrixElements.cc:105-167
 *   GetGME_BFBFInt/TetTetInt      GMatrixElements.cc:270-365
 *   GetGMatrixElement             GMatrixElements.cc:380-424
 *   Invert3x3Matrix, GetOverlaps, AssembleOverlapBlocks, AssembleGBlock,
 *   AssembleVIEMatrix             src/libs/libbuff/VIEMatrix.cc:51-72,122-460
 *   AssembleRHSVector             src/libs/libbuff/RHSVector.cc:62-124
 *   GetEMTPFT (Pabs, Pscat)       src/libs/libbuff/EMTPFT.cc:85-546
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "buffem_oracle.h"

#define ZVAC 376.73031346177       /* libbuff.h:49-51 */
typedef double _Complex cplx;

/* ---------------- TetCR rules (TetCR.cc:35-95), {u1,u2,u3,w} ---------- */
static double CR4[16], CR16[64], CR33[132];
static int cr_ready = 0;

/* append the permutations of barycentric orbit (b0,b1,b2,b3); the stored
 * triple is (b1,b2,b3) as the coefficients of L1,L2,L3 in x = Q + sum u L */
static int add_orbit(double *tab, int n, const double b[4], double w)
{
  int perm[24][4], np = 0;
  int idx[4] = {0, 1, 2, 3};
  /* generate all 24 permutations, keep distinct value tuples */
  for (int a = 0; a < 4; a++)
    for (int c = 0; c < 4; c++)
      for (int d = 0; d < 4; d++)
        for (int e = 0; e < 4; e++) {
          if (a == c || a == d || a == e || c == d || c == e || d == e) continue;
          perm[np][0] = idx[a]; perm[np][1] = idx[c]; perm[np][2] = idx[d]; perm[np][3] = idx[e];
          np++;
        }
  for (int p = 0; p < np; p++) {
    double v[4] = {b[perm[p][0]], b[perm[p][1]], b[perm[p][2]], b[perm[p][3]]};
    int dup = 0;
    for (int q = 0; q < n && !dup; q++)
      if (tab[4 * q] == v[1] && tab[4 * q + 1] == v[2] && tab[4 * q + 2] == v[3]) dup = 1;
    if (dup) continue;
    tab[4 * n] = v[1]; tab[4 * n + 1] = v[2]; tab[4 * n + 2] = v[3]; tab[4 * n + 3] = w;
    n++;
  }
  return n;
}

static void init_rules(void) __attribute__((constructor));
static void init_rules(void)
{
  if (cr_ready) return;
  /* TetCR4: one S31 orbit, a = 0.13819660112501051 (degree 2) */
  { double a = 0.13819660112501051, b[4] = {1.0 - 3 * a, a, a, a};
    add_orbit(CR4, 0, b, 0.041666666666666666); }
  /* TetCR16: S31 + S211 (degree 4) */
  { int n = 0;
    double b1[4] = {0.771642902067237, 0.076119032644254315, 0.076119032644254315, 0.076119032644254315};
    n = add_orbit(CR16, n, b1, 8.3956323500204695e-3);
    double b2[4] = {0.07183164526766925, 0.11970052779780202, 0.40423391346726440, 0.40423391346726440};
    n = add_orbit(CR16, n, b2, 0.011090344772215399); }
  /* TetCR33: S4 + 2 x S31 + 2 x S211 (degree 7, negative weights) */
  { int n = 0;
    double c[4] = {0.25, 0.25, 0.25, 0.25};
    n = add_orbit(CR33, n, c, 0.017783308938202055);
    double s1[4] = {0.05777714911165610, 0.31407428362944793, 0.31407428362944793, 0.31407428362944793};
    n = add_orbit(CR33, n, s1, -0.019099406118263309);
    double s2[4] = {0.81235958507558810, 0.06254680497480394, 0.06254680497480394, 0.06254680497480394};
    n = add_orbit(CR33, n, s2, 3.8252276536630236e-3);
    double s3[4] = {0.61738099120939840, 0.04120213879937276, 0.04120213879937276, 0.30021473119185615};
    n = add_orbit(CR33, n, s3, 0.0035475499144103266);
    double s4[4] = {0.44311853122966216, 0.05292032110257111, 0.25198057383388334, 0.25198057383388334};
    n = add_orbit(CR33, n, s4, 0.013950789384495153); }
  cr_ready = 1;
}

const double *orc_tetcr(int npts)
{
  init_rules();
  if (npts == 4) return CR4;
  if (npts == 16) return CR16;
  if (npts == 33) return CR33;
  return NULL;
}

/* ---------------- ExpRel series (GMatrixElements.cc:59-93) -------------- */
static void exprel_table(cplx x, int nmin, int nmax, cplx *tab)
{
  const double tol2 = 1.0e-8 * 1.0e-8;
  cplx term = 1.0, sum;
  int n;
  for (n = 1; n < nmin; n++) term *= x / (double)n;
  for (n = nmin; n <= nmax; n++) { term *= x / (double)n; tab[n - nmin] = term; }
  for (sum = 0.0; n < 100; n++) {
    term *= x / (double)n;
    sum += term;
    double m2t = creal(term) * creal(term) + cimag(term) * cimag(term);
    double m2s = creal(sum) * creal(sum) + cimag(sum) * cimag(sum);
    if (m2t < tol2 * m2s) break;
  }
  tab[nmax - nmin] += sum;
  for (n = nmax - 1; n >= nmin; n--) tab[n - nmin] += tab[n - nmin + 1];
}

void orc_exprel(double zr, double zi, int nmin, int nmax, double *tab)
{
  cplx t[16];
  if (nmax - nmin + 1 > 16) return;
  exprel_table(zr + I * zi, nmin, nmax, t);
  for (int i = 0; i <= nmax - nmin; i++) { tab[2 * i] = creal(t[i]); tab[2 * i + 1] = cimag(t[i]); }
}

/* ---------------- tetrahedron cubature points --------------------------- */
/* TetInt/TetTetInt fixed branch (Cubature.cc:175-226, :422-467):
 * Q = V[VI[iQ]], L_k = V[VI[(iQ+k)%4]] - Q, x = Q + u.L, w = 6 Vol w_cr,
 * PreFac = Sign*Area(FI[iQ])/(3 Vol)                                     */
typedef struct { double x[33][3], w[33], Q[3], pf; int n; } tetpts;

static void tet_points(const orc_geom *g, int t, int iQ, double sign, int npts, tetpts *P)
{
  const double *cr = orc_tetcr(npts);
  const int *VI = g->tetVI + 4 * t;
  const double *Q = g->V + 3 * VI[iQ];
  double L[3][3];
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) L[k][d] = g->V[3 * VI[(iQ + k + 1) % 4] + d] - Q[d];
  double vol = g->tetVol[t];
  P->n = npts;
  for (int d = 0; d < 3; d++) P->Q[d] = Q[d];
  P->pf = sign * g->fArea[g->tetFI[4 * t + iQ]] / (3.0 * vol);
  for (int i = 0; i < npts; i++) {
    for (int d = 0; d < 3; d++)
      P->x[i][d] = Q[d] + cr[4 * i] * L[0][d] + cr[4 * i + 1] * L[1][d] + cr[4 * i + 2] * L[2][d];
    P->w[i] = 6.0 * vol * cr[4 * i + 3];
  }
}

/* ---------------- G matrix elements -------------------------------------- */
enum { K_HELM = 0, K_DESING = 1, K_STATIC = 2 };

/* one TetTetInt with GMEIntegrand (Cubature.cc:422-467 + GMatrixElements.cc:105-167);
 * out: HELM/DESING -> out[0] complex; STATIC -> out[0..5] real             */
static void tettet_gme(const tetpts *A, const tetpts *B, int kernel, cplx k, cplx *zout, double *sout)
{
  cplx acc = 0.0;
  double s[6] = {0, 0, 0, 0, 0, 0};
  double divdiv = (3.0 * A->pf) * (3.0 * B->pf);
  cplx ik = I * k;
  for (int i = 0; i < A->n; i++) {
    double bA[3];
    for (int d = 0; d < 3; d++) bA[d] = A->pf * (A->x[i][d] - A->Q[d]);
    for (int j = 0; j < B->n; j++) {
      double R[3], bB[3];
      for (int d = 0; d < 3; d++) {
        R[d] = A->x[i][d] - B->x[j][d];
        bB[d] = B->pf * (B->x[j][d] - B->Q[d]);
      }
      double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      if (r < 1.0e-12) continue;                     /* GMatrixElements.cc:122-125 */
      double dot = bA[0] * bB[0] + bA[1] * bB[1] + bA[2] * bB[2];
      double ww = A->w[i] * B->w[j];
      if (kernel == K_HELM) {
        cplx phi = cexp(ik * r) / (4.0 * M_PI * r);
        acc += ww * (dot - divdiv / (k * k)) * phi;
      } else if (kernel == K_DESING) {
        cplx ef[1];
        exprel_table(ik * r, 3, 3, ef);
        cplx phi = ef[0] / (4.0 * M_PI * r);
        acc += ww * (dot - divdiv / (k * k)) * phi;
      } else {
        double r0 = 1.0 / (4.0 * M_PI * r), r1 = 1.0 / (4.0 * M_PI), r2 = r / (8.0 * M_PI);
        s[0] += ww * r0 * dot; s[1] += ww * r0 * divdiv;
        s[2] += ww * r1 * dot; s[3] += ww * r1 * divdiv;
        s[4] += ww * r2 * dot; s[5] += ww * r2 * divdiv;
      }
    }
  }
  if (zout) *zout = acc;
  if (sout) memcpy(sout, s, sizeof s);
}

/* BFBFInt with NumPts=4 (Cubature.cc:476-527) */
static cplx bfbf_gme(const orc_geom *g, int a, int b, int kernel, cplx k)
{
  tetpts PA, MA, PB, MB;
  tet_points(g, g->fPT[a], g->fPI[a], +1.0, 4, &PA);
  tet_points(g, g->fMT[a], g->fMI[a], -1.0, 4, &MA);
  tet_points(g, g->fPT[b], g->fPI[b], +1.0, 4, &PB);
  tet_points(g, g->fMT[b], g->fMI[b], -1.0, 4, &MB);
  cplx r = 0.0, t;
  tettet_gme(&PA, &PB, kernel, k, &t, NULL); r += t;
  tettet_gme(&PA, &MB, kernel, k, &t, NULL); r += t;
  tettet_gme(&MA, &PB, kernel, k, &t, NULL); r += t;
  tettet_gme(&MA, &MB, kernel, k, &t, NULL); r += t;
  return r;
}

/* Taylor-Duffy static record of one tet pair in the form of
 * GetGMETTI_TaylorDuffy (GMatrixElements.cc:174-263) times 4 A_a A_b
 * (:352-355): out6 = {G_p: int b.b' r^p, int divdiv r^p} * GPre_p       */
static int g_td_order = 12;

/* Taylor-Duffy moments of tet pair (ta, tb) in CompareTets vertex order,
 * reference points = centroids of ta, tb; 3 x 8 doubles.                 */
static void td_pair_moments(const orc_geom *g, int ta, int tb, int ncv, int td_order, double mom[3][8])
{
  int OA[4], OB[4];
  orc_compare_tets(g, ta, tb, OA, OB);
  double VA[12], VB[12];
  for (int i = 0; i < 4; i++)
    for (int d = 0; d < 3; d++) { VA[3 * i + d] = g->V[3 * OA[i] + d]; VB[3 * i + d] = g->V[3 * OB[i] + d]; }
  int order = td_order + (ncv == 4 ? 4 : ncv == 3 ? 2 : 0);
  if (order > 32) order = 32;
  orc_td_moments(ncv, VA, VB, g->tetCen + 3 * ta, g->tetCen + 3 * tb, order, mom);
}

/* static record of one tet pair from its moments (x-tet = ta, x'-tet = tb):
 * out6 = {G_p: int (x-QA).(x'-QB) r^p, int r^p} * GPre_p, p = -1,0,1
 * (GetGMETTI_TaylorDuffy, GMatrixElements.cc:174-263, before the
 * 4 A_a A_b / 9 prefactors of :352-355)                                  */
static void td_static_from_moments(const orc_geom *g, int ta, int iQA, int tb, int iQB,
                                   const double mom[3][8], int swapped, double out6[6])
{
  const double *cA = g->tetCen + 3 * ta, *cB = g->tetCen + 3 * tb;
  const double *QA = g->V + 3 * iQA, *QB = g->V + 3 * iQB;
  double dA[3], dB[3];
  for (int d = 0; d < 3; d++) { dA[d] = cA[d] - QA[d]; dB[d] = cB[d] - QB[d]; }
  const int ox = swapped ? 4 : 1, oy = swapped ? 1 : 4;
  const double gpre[3] = {1.0 / (4.0 * M_PI), 1.0 / (4.0 * M_PI), 1.0 / (8.0 * M_PI)};
  for (int p = 0; p < 3; p++) {
    const double *M = mom[p];
    double bb = M[7] + dA[0] * M[oy] + dA[1] * M[oy + 1] + dA[2] * M[oy + 2]
              + dB[0] * M[ox] + dB[1] * M[ox + 1] + dB[2] * M[ox + 2]
              + (dA[0] * dB[0] + dA[1] * dB[1] + dA[2] * dB[2]) * M[0];
    out6[2 * p] = gpre[p] * bb;
    out6[2 * p + 1] = gpre[p] * M[0];
  }
}

/* optional cache of tet-pair moments: sorted keys (lo<<32|hi) */
typedef struct { int64_t n; uint64_t *key; double (*mom)[3][8]; } td_cache;

static const double (*cache_find(const td_cache *c, int lo, int hi))[8]
{
  uint64_t k = ((uint64_t)lo << 32) | (uint32_t)hi;
  int64_t a = 0, b = c->n - 1;
  while (a <= b) {
    int64_t m = (a + b) / 2;
    if (c->key[m] == k) return (const double (*)[8])c->mom[m];
    if (c->key[m] < k) a = m + 1; else b = m - 1;
  }
  return NULL;
}

static void fibbi_impl(const orc_geom *g, int a, int b, int td_order, const td_cache *cache, double G[6]);

/* ComputeFIBBIData = GetGME_TetTetInt(KERNEL_STATIC) (GMatrixElements.cc:295-375) */
static void fibbi_impl(const orc_geom *g, int a, int b, int td_order, const td_cache *cache, double G[6])
{
  memset(G, 0, 6 * sizeof(double));
  for (int sa = 0; sa < 2; sa++)
    for (int sb = 0; sb < 2; sb++) {
      int ta = sa == 0 ? g->fPT[a] : g->fMT[a];
      int tb = sb == 0 ? g->fPT[b] : g->fMT[b];
      int ncv = orc_compare_tets(g, ta, tb, NULL, NULL);
      double tti[6];
      if (ncv <= 1) {
        tetpts A, B;
        tet_points(g, ta, sa == 0 ? g->fPI[a] : g->fMI[a], 1.0, 16, &A);
        tet_points(g, tb, sb == 0 ? g->fPI[b] : g->fMI[b], 1.0, 16, &B);
        tettet_gme(&A, &B, K_STATIC, 0.0, NULL, tti);
      } else {
        int iQA = sa == 0 ? g->fQP[a] : g->fQM[a];
        int iQB = sb == 0 ? g->fQP[b] : g->fQM[b];
        int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
        double local[3][8];
        const double (*mom)[8] = cache ? cache_find(cache, lo, hi) : NULL;
        if (!mom) { td_pair_moments(g, lo, hi, ncv, td_order, local); mom = (const double (*)[8])local; }
        double raw[6];
        td_static_from_moments(g, ta, iQA, tb, iQB, mom, ta > tb, raw);
        /* b = (Area/3V)(x-Q), divb = Area/V, volumes unsigned (TD path) */
        double pa = g->fArea[a] / (3.0 * fabs(g->tetVol[ta]));
        double pb = g->fArea[b] / (3.0 * fabs(g->tetVol[tb]));
        for (int p = 0; p < 3; p++) {
          tti[2 * p] = pa * pb * raw[2 * p];
          tti[2 * p + 1] = 9.0 * pa * pb * raw[2 * p + 1];
        }
      }
      double sign = (sa == sb) ? 1.0 : -1.0;
      for (int i = 0; i < 6; i++) G[i] += sign * tti[i];
    }
}

void orc_fibbi(const orc_geom *g, int a, int b, int td_order, double G[6])
{
  if (td_order <= 0) td_order = g_td_order;
  fibbi_impl(g, a, b, td_order, NULL, G);
}

static int cmp_u64(const void *x, const void *y)
{
  uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
  return a < b ? -1 : a > b ? 1 : 0;
}

/* GetGMatrixElement (GMatrixElements.cc:380-424) */
void orc_gme(const orc_geom *g, int a, int b, double omr, double omi,
             const double *fibbi, double out[2])
{
  init_rules();
  cplx k = omr + I * omi;
  int ncv = orc_compare_bfs(g, a, b);
  cplx gme;
  if (ncv == 0) {
    gme = bfbf_gme(g, a, b, K_HELM, k);
  } else {
    gme = bfbf_gme(g, a, b, K_DESING, k);
    double G[6];
    if (fibbi) memcpy(G, fibbi, sizeof G);
    else orc_fibbi(g, a, b, 0, G);
    cplx k2 = k * k, ik = I * k, ik2 = ik * ik;
    gme += (G[0] - G[1] / k2) + ik * (G[2] - G[3] / k2) + ik2 * (G[4] - G[5] / k2);
  }
  out[0] = creal(gme); out[1] = cimag(gme);
}

/* ---------------- touching pairs and the static table ------------------- */
int orc_touch_pairs(const orc_geom *g, int *rowptr, int *cols)
{
  int N = g->nf_int, nnz = 0;
  /* vertex -> BFs incidence */
  int *cnt = (int *)calloc(g->nv + 1, sizeof(int));
  for (int a = 0; a < N; a++) {
    int vs[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
    for (int i = 0; i < 5; i++) cnt[vs[i] + 1]++;
  }
  for (int v = 0; v < g->nv; v++) cnt[v + 1] += cnt[v];
  int *lst = (int *)malloc(sizeof(int) * cnt[g->nv]);
  int *fill = (int *)calloc(g->nv, sizeof(int));
  for (int a = 0; a < N; a++) {
    int vs[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
    for (int i = 0; i < 5; i++) lst[cnt[vs[i]] + fill[vs[i]]++] = a;
  }
  char *mark = (char *)calloc(N, 1);
  int *tmp = (int *)malloc(sizeof(int) * N);
  rowptr[0] = 0;
  for (int a = 0; a < N; a++) {
    int nt = 0;
    int vs[5] = {g->fQP[a], g->fV[3 * a], g->fV[3 * a + 1], g->fV[3 * a + 2], g->fQM[a]};
    for (int i = 0; i < 5; i++)
      for (int j = cnt[vs[i]]; j < cnt[vs[i] + 1]; j++) {
        int b = lst[j];
        if (b >= a && !mark[b]) { mark[b] = 1; tmp[nt++] = b; }
      }
    /* sort ascending */
    for (int i = 1; i < nt; i++) {
      int v = tmp[i], j = i - 1;
      while (j >= 0 && tmp[j] > v) { tmp[j + 1] = tmp[j]; j--; }
      tmp[j + 1] = v;
    }
    for (int i = 0; i < nt; i++) {
      mark[tmp[i]] = 0;
      if (cols) cols[nnz + i] = tmp[i];
    }
    nnz += nt;
    rowptr[a + 1] = nnz;
  }
  free(cnt); free(lst); free(fill); free(mark); free(tmp);
  return nnz;
}

void orc_static_table(const orc_geom *g, const int *rowptr, const int *cols,
                      int td_order, double *table)
{
  int N = g->nf_int;
  if (td_order <= 0) td_order = g_td_order;
  /* unique Taylor-Duffy tet pairs of all touching BF pairs */
  int64_t cap = 16, n = 0;
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * cap);
  for (int a = 0; a < N; a++)
    for (int j = rowptr[a]; j < rowptr[a + 1]; j++) {
      int b = cols[j];
      for (int s = 0; s < 4; s++) {
        int ta = (s >> 1) == 0 ? g->fPT[a] : g->fMT[a];
        int tb = (s & 1) == 0 ? g->fPT[b] : g->fMT[b];
        if (orc_compare_tets(g, ta, tb, NULL, NULL) < 2) continue;
        int lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
        if (n == cap) { cap *= 2; keys = (uint64_t *)realloc(keys, sizeof(uint64_t) * cap); }
        keys[n++] = ((uint64_t)lo << 32) | (uint32_t)hi;
      }
    }
  qsort(keys, n, sizeof(uint64_t), cmp_u64);
  int64_t u = 0;
  for (int64_t i = 0; i < n; i++)
    if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
  td_cache c;
  c.n = u; c.key = keys;
  c.mom = (double (*)[3][8])malloc(sizeof(double) * 24 * (u > 0 ? u : 1));
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < u; i++) {
    int lo = (int)(keys[i] >> 32), hi = (int)(keys[i] & 0xffffffffu);
    int ncv = orc_compare_tets(g, lo, hi, NULL, NULL);
    td_pair_moments(g, lo, hi, ncv, td_order, c.mom[i]);
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < N; a++)
    for (int j = rowptr[a]; j < rowptr[a + 1]; j++)
      fibbi_impl(g, a, cols[j], td_order, &c, table + 6 * (size_t)j);
  free(c.mom);
  free(keys);
}

/* ---------------- materials (SVTensor::Evaluate semantics) -------------- */
void orc_eps(const orc_material *m, const double *x, double eps[3][3][2])
{
  memset(eps, 0, 18 * sizeof(double));
  double e[3][2];
  switch (m->kind) {
    case 1:
      e[0][0] = m->p[0]; e[0][1] = m->p[1]; e[1][0] = m->p[2]; e[1][1] = m->p[3];
      e[2][0] = m->p[4]; e[2][1] = m->p[5];
      break;
    case 2: {
      double re = m->p[0] + m->p[2] * sin(M_PI * x[0]) * cos(M_PI * x[1]);
      for (int i = 0; i < 3; i++) { e[i][0] = re; e[i][1] = m->p[1]; }
    } break;
    case 3: {
      int up = x[2] > 0.0;
      for (int i = 0; i < 3; i++) { e[i][0] = up ? m->p[0] : m->p[2]; e[i][1] = up ? m->p[1] : m->p[3]; }
    } break;
    default:
      for (int i = 0; i < 3; i++) { e[i][0] = m->p[0]; e[i][1] = m->p[1]; }
  }
  for (int i = 0; i < 3; i++) { eps[i][i][0] = e[i][0]; eps[i][i][1] = e[i][1]; }
}

/* Invert3x3Matrix (VIEMatrix.cc:51-72) */
static void inv3(cplx M[3][3], cplx W[3][3])
{
  cplx M11 = M[0][0], M12 = M[0][1], M13 = M[0][2];
  cplx M21 = M[1][0], M22 = M[1][1], M23 = M[1][2];
  cplx M31 = M[2][0], M32 = M[2][1], M33 = M[2][2];
  cplx Det = M12 * M23 * M31 + M13 * M21 * M32 + M11 * M22 * M33
           - M13 * M22 * M31 - M11 * M23 * M32 - M12 * M21 * M33;
  W[0][0] = (M22 * M33 - M23 * M32) / Det;
  W[0][1] = (M13 * M32 - M12 * M33) / Det;
  W[0][2] = (M12 * M23 - M13 * M22) / Det;
  W[1][0] = (M23 * M31 - M21 * M33) / Det;
  W[1][1] = (M11 * M33 - M13 * M31) / Det;
  W[1][2] = (M13 * M21 - M11 * M23) / Det;
  W[2][0] = (M21 * M32 - M22 * M31) / Det;
  W[2][1] = (M12 * M31 - M11 * M32) / Det;
  W[2][2] = (M11 * M22 - M12 * M21) / Det;
}

/* one TetInt(NumPts=33, iQ=0) of GetOverlapIntegrand's VInv part
 * (VIEMatrix.cc:122-180, 252-253) */
static cplx overlap_tet(const orc_geom *g, int t, const double *QA, double pfA,
                        const double *QB, double pfB, double omega, const orc_material *mat)
{
  tetpts P;
  tet_points(g, t, 0, 1.0, 33, &P);
  cplx acc = 0.0;
  for (int i = 0; i < 33; i++) {
    double e[3][3][2];
    orc_eps(mat, P.x[i], e);
    cplx Em1[3][3], Inv[3][3];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Em1[r][c] = e[r][c][0] + I * e[r][c][1];
    for (int r = 0; r < 3; r++) Em1[r][r] -= 1.0;
    inv3(Em1, Inv);
    double FA[3], FB[3];
    for (int d = 0; d < 3; d++) {
      FA[d] = pfA * (P.x[i][d] - QA[d]);
      FB[d] = pfB * (P.x[i][d] - QB[d]);
    }
    cplx v = 0.0;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) v += FA[r] * Inv[r][c] * FB[c];
    acc += P.w[i] * v;
  }
  return acc * (-1.0 / (omega * omega));
}

int orc_overlaps(const orc_geom *g, int a, double omega, const orc_material *mat,
                 int idx[7], double val[7][2])
{
  init_rules();
  cplx v[7];
  idx[0] = a; v[0] = 0.0;
  int nnz = 1;
  for (int side = 0; side < 2; side++) {
    int t = side == 0 ? g->fPT[a] : g->fMT[a];
    const double *QA = g->V + 3 * (side == 0 ? g->fQP[a] : g->fQM[a]);
    double pfA = (side == 0 ? 1.0 : -1.0) * g->fArea[a] / (3.0 * g->tetVol[t]);
    for (int ifb = 0; ifb < 4; ifb++) {
      int b = g->tetFI[4 * t + ifb];
      if (b >= g->nf_int) continue;
      const double *QB;
      double pfB;
      if (g->fPT[b] == t) { QB = g->V + 3 * g->fQP[b]; pfB = g->fArea[b] / (3.0 * g->tetVol[t]); }
      else { QB = g->V + 3 * g->fQM[b]; pfB = -1.0 * g->fArea[b] / (3.0 * g->tetVol[t]); }
      cplx o = overlap_tet(g, t, QA, pfA, QB, pfB, omega, mat);
      if (b == a) v[0] += o;
      else { idx[nnz] = b; v[nnz] = o; nnz++; }
    }
  }
  for (int i = 0; i < nnz; i++) { val[i][0] = creal(v[i]); val[i][1] = cimag(v[i]); }
  return nnz;
}

/* AssembleVIEMatrix (VIEMatrix.cc:431-460) for one object */
void orc_assemble(const orc_geom *g, double omr, double omi, const orc_material *mat,
                  const int *rowptr, const int *cols, const double *table, double *M)
{
  init_rules();
  int N = g->nf_int;
  cplx *Z = (cplx *)M;
  /* AssembleGBlock: upper triangle, both triangles written (:401-416) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < N; a++) {
    int j = rowptr[a];
    for (int b = a; b < N; b++) {
      const double *fib = NULL;
      if (j < rowptr[a + 1] && cols[j] == b) { fib = table + 6 * (size_t)j; j++; }
      double out[2];
      if (fib) orc_gme(g, a, b, omr, omi, fib, out);
      else {
        cplx k = omr + I * omi;
        cplx z = bfbf_gme(g, a, b, K_HELM, k);
        out[0] = creal(z); out[1] = cimag(z);
      }
      Z[(size_t)b * N + a] = out[0] + I * out[1];
      Z[(size_t)a * N + b] = out[0] + I * out[1];
    }
  }
  /* AssembleOverlapBlocks: AddEntry(nr,nc) (:321-360) */
  for (int a = 0; a < N; a++) {
    int idx[7];
    double val[7][2];
    int nnz = orc_overlaps(g, a, omr, mat, idx, val);
    for (int i = 0; i < nnz; i++) Z[(size_t)idx[i] * N + a] += val[i][0] + I * val[i][1];
  }
  /* serial symmetrisation: lower <- upper (:454-456) */
  for (int c = 0; c < N; c++)
    for (int r = c + 1; r < N; r++) Z[(size_t)c * N + r] = Z[(size_t)r * N + c];
}

/* Warm assembly of the upper-triangle part (b >= a) of selected rows, as the
 * row loop of AssembleGBlock (VIEMatrix.cc:401-416) does; used to time a
 * bounded sample.  table may be NULL (static records taken as zero: the
 * warm cost does not depend on their values).  out: [nrows][N] complex,
 * entries b < a left untouched.                                          */
void orc_assemble_rows(const orc_geom *g, double omr, double omi, int nrows, const int *rows,
                       const int *rowptr, const int *cols, const double *table, double *out)
{
  int N = g->nf_int;
  cplx *Z = (cplx *)out;
  const double zero6[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < nrows; i++) {
    int a = rows[i];
    int j = rowptr[a];
    for (int b = a; b < N; b++) {
      const double *fib = NULL;
      if (j < rowptr[a + 1] && cols[j] == b) { fib = table ? table + 6 * (size_t)j : zero6; j++; }
      double v[2];
      if (fib) orc_gme(g, a, b, omr, omi, fib, v);
      else {
        cplx z = bfbf_gme(g, a, b, K_HELM, omr + I * omi);
        v[0] = creal(z); v[1] = cimag(z);
      }
      Z[(size_t)i * N + b] = v[0] + I * v[1];
    }
  }
}

/* ---------------- incident field, RHS, EMT ------------------------------ */
static void pw_field(const orc_planewave *pw, double k, const double *x, cplx E[3])
{
  double ph = k * (pw->dir[0] * x[0] + pw->dir[1] * x[1] + pw->dir[2] * x[2]);
  cplx f = cexp(I * ph);
  for (int d = 0; d < 3; d++) E[d] = (pw->E0[2 * d] + I * pw->E0[2 * d + 1]) * f;
}

/* BFInt(NumPts=33) of b.E (Cubature.cc:234-258, RHSVector.cc:62-80) */
static cplx bf_dot_e(const orc_geom *g, int a, const orc_planewave *pw, double k)
{
  cplx acc = 0.0;
  for (int side = 0; side < 2; side++) {
    tetpts P;
    int t = side == 0 ? g->fPT[a] : g->fMT[a];
    tet_points(g, t, side == 0 ? g->fPI[a] : g->fMI[a], side == 0 ? 1.0 : -1.0, 33, &P);
    for (int i = 0; i < 33; i++) {
      cplx E[3];
      pw_field(pw, k, P.x[i], E);
      double b[3];
      for (int d = 0; d < 3; d++) b[d] = P.pf * (P.x[i][d] - P.Q[d]);
      acc += P.w[i] * (b[0] * E[0] + b[1] * E[1] + b[2] * E[2]);
    }
  }
  return acc;
}

void orc_rhs(const orc_geom *g, double omega, const orc_planewave *pw, int npw, double *B)
{
  init_rules();
  int N = g->nf_int;
  cplx pre = -1.0 / (I * omega * ZVAC);
  cplx *Z = (cplx *)B;
#pragma omp parallel for schedule(dynamic, 16)
  for (int a = 0; a < N; a++)
    for (int j = 0; j < npw; j++) Z[(size_t)j * N + a] = pre * bf_dot_e(g, a, &pw[j], omega);
}

/* ScatteredPFTIntegrand, same object: Im Q_PSCAT (EMTPFT.cc:216-291) */
static double scat_tettet(const tetpts *A, const tetpts *B, double k)
{
  double acc = 0.0;
  double divdiv = (3.0 * A->pf) * (3.0 * B->pf);
  for (int i = 0; i < A->n; i++) {
    double bA[3];
    for (int d = 0; d < 3; d++) bA[d] = A->pf * (A->x[i][d] - A->Q[d]);
    for (int j = 0; j < B->n; j++) {
      double R[3], bB[3];
      for (int d = 0; d < 3; d++) {
        R[d] = A->x[i][d] - B->x[j][d];
        bB[d] = B->pf * (B->x[j][d] - B->Q[d]);
      }
      double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
      double r = sqrt(r2), kr = k * r;
      double dot = bA[0] * bB[0] + bA[1] * bB[1] + bA[2] * bB[2];
      double pefie = dot - divdiv / (k * k);
      double imphi;
      if (fabs(kr) < 2.0e-2) imphi = (1.0 - kr * kr / 6.0) * k / (4.0 * M_PI);
      else imphi = sin(kr) / (4.0 * M_PI * r);
      acc += A->w[i] * B->w[j] * k * pefie * imphi;
    }
  }
  return acc;
}

/* Im Q_PSCAT of GetScatteredPFTIntegrals for BF pair (a, b): the rule by
 * common vertices and k*radius (EMTPFT.cc:294-318), four tetrahedron pairs */
static double emt_pair_q(const orc_geom *g, int a, int b, double omega)
{
  int ncv = orc_compare_bfs(g, a, b);
  double kR = fabs(omega) * fmax(g->fRad[a], g->fRad[b]);
  int hf = kR > 5.0;
  int np = hf ? (ncv > 0 ? 33 : 16) : (ncv > 0 ? 16 : 4);
  tetpts PA, MA, PB, MB;
  tet_points(g, g->fPT[a], g->fPI[a], +1.0, np, &PA);
  tet_points(g, g->fMT[a], g->fMI[a], -1.0, np, &MA);
  tet_points(g, g->fPT[b], g->fPI[b], +1.0, np, &PB);
  tet_points(g, g->fMT[b], g->fMI[b], -1.0, np, &MB);
  return scat_tettet(&PA, &PB, omega) + scat_tettet(&PA, &MB, omega) + scat_tettet(&MA, &PB, omega)
       + scat_tettet(&MA, &MB, omega);
}

void orc_emt_power(const orc_geom *g, double omega, const orc_planewave *pw, int npw,
                   const double *J, double *out)
{
  init_rules();
  int N = g->nf_int;
  const cplx *Z = (const cplx *)J;
  double k = omega;
  double *pscat = (double *)calloc(npw, sizeof(double));
  /* scattered power: sum over a<=b (EMTPFT.cc:389-427) */
#pragma omp parallel
  {
    double *loc = (double *)calloc(npw, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int a = 0; a < N; a++) {
      for (int b = a; b < N; b++) {
        double q = emt_pair_q(g, a, b, omega);
        for (int j = 0; j < npw; j++) {
          cplx jj = ZVAC * conj(Z[(size_t)j * N + a]) * Z[(size_t)j * N + b];
          loc[j] += creal(jj) * q;
        }
      }
    }
#pragma omp critical
    for (int j = 0; j < npw; j++) pscat[j] += loc[j];
    free(loc);
  }
  /* extinction (EMTPFT.cc:151-211): Pext = 1/2 sum Re(J* Q) */
  for (int j = 0; j < npw; j++) {
    double pext = 0.0;
    for (int a = 0; a < N; a++) {
      cplx q = bf_dot_e(g, a, &pw[j], k);
      pext += creal(conj(Z[(size_t)j * N + a]) * q);
    }
    pext *= 0.5;
    out[2 * j] = pext - pscat[j];
    out[2 * j + 1] = pscat[j];
  }
  free(pscat);
}

/* GetEMTPFT with a DRMatrix and no incident field (EMTPFT.cc:389-476,
 * GetJJ GetPFT.cc:108-115): u0JJ = Z0 DR(b, a) for a <= b, P_scat = sum
 * Re(u0JJ) Im Q, P_abs = -P_scat.  DR: N x N complex, column-major.      */
void orc_emt_power_dr(const orc_geom *g, double omega, const double *DR, double *out)
{
  init_rules();
  int N = g->nf_int;
  const cplx *Z = (const cplx *)DR;
  double pscat = 0.0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : pscat)
  for (int a = 0; a < N; a++)
    for (int b = a; b < N; b++) pscat += creal(ZVAC * Z[(size_t)a * N + b]) * emt_pair_q(g, a, b, omega);
  out[0] = -pscat;
  out[1] = pscat;
}
