/*
 * oracle_surr.c -- oracle of the tetrahedron-pair blocks behind the learned
 * near-field surrogate.  TEST INFRASTRUCTURE ONLY (see buffem_oracle.h).
 *
 * Block of a tetrahedron pair (S, T) with vertices v0..v3 / w0..w3: face i
 * of S is opposite v_i, with area A_i; f_i(x) = A_i/(3|V_S|) (x - v_i),
 * div f_i = A_i/|V_S| (the SWG restriction to one tetrahedron without the
 * +/- sign, src/libs/libbuff/Cubature.cc:174,211-218); likewise g_j on T.
 *   G_ij(k) = int_S int_T [f_i(x).g_j(y) - div f_i div g_j / k^2]
 *                          exp(ikr) / (4 pi r)                      (r=|x-y|)
 * i.e. the Helmholtz integrand of GMEIntegrand (GMatrixElements.cc:105-145)
 * for one tetrahedron pair; a far-field matrix element is the signed sum of
 * four such entries (GetGME_BFBFInt, GMatrixElements.cc:270-288).
 *
 * orc_tp_block integrates with a product of TetCR rules on each tetrahedron
 * refined `level` times by regular (8-way) subdivision; level 0 with the
 * 4-point rule is the reference's fixed far-field rule (NumPts=4,
 * GMatrixElements.cc:282).  Deeper levels with the 16/33-point rules stand
 * in for the reference's adaptive cubature (Cubature.cc:419, hcubature) that
 * labels the surrogate's training pairs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "buffem_oracle.h"

#define FOURPI 12.566370614359172954

static void sub3(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static void cross3(const double *a, const double *b, double *o)
{
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static double tet_absvol(const double *v)
{
  double a[3], b[3], c[3], x[3];
  sub3(v + 3, v, a); sub3(v + 6, v, b); sub3(v + 9, v, c);
  cross3(b, c, x);
  return fabs(dot3(a, x)) / 6.0;
}

static double tri_area(const double *p, const double *q, const double *r)
{
  double a[3], b[3], x[3];
  sub3(q, p, a); sub3(r, p, b);
  cross3(a, b, x);
  return 0.5 * sqrt(dot3(x, x));
}

/* face areas A_i (face opposite vertex i) */
static void face_areas(const double *v, double A[4])
{
  A[0] = tri_area(v + 3, v + 6, v + 9);
  A[1] = tri_area(v + 0, v + 6, v + 9);
  A[2] = tri_area(v + 0, v + 3, v + 9);
  A[3] = tri_area(v + 0, v + 3, v + 6);
}

/* Regular subdivision: the 8 children of tetrahedron v (4 corner
 * tetrahedra + the octahedron split along the x02-x13 diagonal). */
static void subdivide(const double *v, double *kids /* 8 x 12 */)
{
  double m[4][4][3];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < 3; c++) m[i][j][c] = 0.5 * (v[3 * i + c] + v[3 * j + c]);
  /* corner tetrahedra */
  int ct[4][4][2] = {{{0, 0}, {0, 1}, {0, 2}, {0, 3}},
                     {{0, 1}, {1, 1}, {1, 2}, {1, 3}},
                     {{0, 2}, {1, 2}, {2, 2}, {2, 3}},
                     {{0, 3}, {1, 3}, {2, 3}, {3, 3}}};
  /* octahedron x01 x02 x03 x12 x13 x23 split around the x02-x13 axis */
  int oc[4][4][2] = {{{0, 2}, {1, 3}, {0, 1}, {0, 3}},
                     {{0, 2}, {1, 3}, {0, 1}, {1, 2}},
                     {{0, 2}, {1, 3}, {2, 3}, {0, 3}},
                     {{0, 2}, {1, 3}, {2, 3}, {1, 2}}};
  for (int t = 0; t < 4; t++)
    for (int k = 0; k < 4; k++)
      for (int c = 0; c < 3; c++) {
        kids[12 * t + 3 * k + c] = m[ct[t][k][0]][ct[t][k][1]][c];
        kids[12 * (4 + t) + 3 * k + c] = m[oc[t][k][0]][oc[t][k][1]][c];
      }
}

/* cubature points {x,y,z,w} of tetrahedron v refined `level` times; w
 * includes the sub-tetrahedron volume (sum w = |V|).  Returns count. */
static int tet_points(const double *v, int npts, int level, double *out)
{
  const double *cr = orc_tetcr(npts);
  int ntet = 1;
  for (int l = 0; l < level; l++) ntet *= 8;
  double *tets = (double *)malloc(sizeof(double) * 12 * ntet);
  double *tmp = (double *)malloc(sizeof(double) * 12 * ntet);
  memcpy(tets, v, sizeof(double) * 12);
  int cur = 1;
  for (int l = 0; l < level; l++) {
    for (int t = 0; t < cur; t++) subdivide(tets + 12 * t, tmp + 96 * t);
    cur *= 8;
    memcpy(tets, tmp, sizeof(double) * 12 * cur);
  }
  int n = 0;
  for (int t = 0; t < cur; t++) {
    const double *q = tets + 12 * t;
    const double vol = tet_absvol(q);
    for (int p = 0; p < npts; p++) {
      const double u1 = cr[4 * p], u2 = cr[4 * p + 1], u3 = cr[4 * p + 2], w = cr[4 * p + 3];
      for (int c = 0; c < 3; c++)
        out[4 * n + c] = q[c] + u1 * (q[3 + c] - q[c]) + u2 * (q[6 + c] - q[c]) + u3 * (q[9 + c] - q[c]);
      out[4 * n + 3] = 6.0 * vol * w;
      n++;
    }
  }
  free(tets);
  free(tmp);
  return n;
}

void orc_tp_block(const double *VS, const double *VT, double kr, double ki, int npts, int level,
                  double out[32])
{
  int ntet = 1;
  for (int l = 0; l < level; l++) ntet *= 8;
  double *ps = (double *)malloc(sizeof(double) * 4 * npts * ntet);
  double *pt = (double *)malloc(sizeof(double) * 4 * npts * ntet);
  const int ns = tet_points(VS, npts, level, ps);
  const int nt = tet_points(VT, npts, level, pt);
  double AS[4], AT[4];
  face_areas(VS, AS);
  face_areas(VT, AT);
  const double vs = tet_absvol(VS), vt = tet_absvol(VT);
  /* 1/k^2 */
  const double k2r = kr * kr - ki * ki, k2i = 2.0 * kr * ki, k2n = k2r * k2r + k2i * k2i;
  const double ik2r = k2r / k2n, ik2i = -k2i / k2n;
  /* moments about the centroids cS, cT (x = cS + xi, y = cT + eta):
   * I0 = int g, Ix = int xi g, Iy = int eta g, Ixy = int xi.eta g       */
  double cS[3], cT[3];
  for (int c = 0; c < 3; c++) {
    cS[c] = 0.25 * (VS[c] + VS[3 + c] + VS[6 + c] + VS[9 + c]);
    cT[c] = 0.25 * (VT[c] + VT[3 + c] + VT[6 + c] + VT[9 + c]);
  }
  double I0[2] = {0, 0}, Ix[3][2] = {{0, 0}}, Iy[3][2] = {{0, 0}}, Ixy[2] = {0, 0};
  for (int a = 0; a < ns; a++) {
    const double *x = ps + 4 * a;
    double xi[3];
    sub3(x, cS, xi);
    for (int b = 0; b < nt; b++) {
      const double *y = pt + 4 * b;
      double R[3], eta[3];
      sub3(x, y, R);
      sub3(y, cT, eta);
      const double r = sqrt(dot3(R, R));
      const double amp = exp(-ki * r) / (FOURPI * r) * x[3] * y[3];
      const double gr = amp * cos(kr * r), gi = amp * sin(kr * r);
      I0[0] += gr; I0[1] += gi;
      for (int c = 0; c < 3; c++) {
        Ix[c][0] += gr * xi[c]; Ix[c][1] += gi * xi[c];
        Iy[c][0] += gr * eta[c]; Iy[c][1] += gi * eta[c];
      }
      const double xy = dot3(xi, eta);
      Ixy[0] += gr * xy; Ixy[1] += gi * xy;
    }
  }
  /* with a_i = v_i - cS, b_j = w_j - cT:
   * G_ij = cS_i cT_j [Ixy - b_j.Ix - a_i.Iy + (a_i.b_j) I0] - dS_i dT_j I0 / k^2 */
  for (int i = 0; i < 4; i++) {
    double vi[3];
    sub3(VS + 3 * i, cS, vi);
    const double cs = AS[i] / (3.0 * vs), ds = AS[i] / vs;
    for (int j = 0; j < 4; j++) {
      double wj[3];
      sub3(VT + 3 * j, cT, wj);
      const double ct = AT[j] / (3.0 * vt), dt = AT[j] / vt;
      double re = Ixy[0], im = Ixy[1];
      for (int c = 0; c < 3; c++) {
        re -= wj[c] * Ix[c][0] + vi[c] * Iy[c][0];
        im -= wj[c] * Ix[c][1] + vi[c] * Iy[c][1];
      }
      const double vw = dot3(vi, wj);
      re += vw * I0[0];
      im += vw * I0[1];
      re *= cs * ct;
      im *= cs * ct;
      /* - ds dt I0 / k^2 */
      const double dd = ds * dt;
      re -= dd * (I0[0] * ik2r - I0[1] * ik2i);
      im -= dd * (I0[0] * ik2i + I0[1] * ik2r);
      out[2 * (4 * i + j)] = re;
      out[2 * (4 * i + j) + 1] = im;
    }
  }
  free(ps);
  free(pt);
}
