/*
 * buffem_oracle.h -- CPU oracle for the dense VIE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's algorithm (HomerReid/buff-em, src/libs/libbuff and
 * src/libs/libTTaylorDuffy), used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, as what results are compared against.  No
 * product code links or calls it.
 *
 * Every function cites the reference file:line it restates.  Parity pins
 * (closed forms, known-answer tests, brute force) live in tests/.
 *
 * Conventions: dense matrices are column-major complex128 stored as
 * interleaved (re,im) doubles, M[(c*N + r)*2 + {0,1}].
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- geometry / topology -------------------------------- */
typedef struct {
  int nv, nt, nf_int, nf_tot;
  double *V;          /* [3*nv] vertices                                  */
  int *tetVI;         /* [4*nt] vertex indices (as given)                 */
  int *tetFI;         /* [4*nt] face opposite vertex i                    */
  double *tetVol;     /* [nt] signed volume  (SWGVolume.cc:70-79)          */
  double *tetCen;     /* [3*nt] centroid                                  */
  /* faces: first nf_int are interior (basis functions)                 */
  int *fV;            /* [3*nf_tot] iV1<iV2<iV3                           */
  int *fQP, *fQM;     /* [nf_tot] opposite vertices                       */
  int *fPT, *fMT;     /* [nf_tot] +/- tets                                */
  int *fPI, *fMI;     /* [nf_tot] index of face within +/- tet            */
  double *fArea;      /* [nf_tot]                                          */
  double *fCen;       /* [3*nf_tot]                                        */
  double *fRad;       /* [nf_tot] bounding radius (interior faces)         */
} orc_geom;

int  orc_geom_build(int nv, const double *V, int nt, const int *tetVI, orc_geom **out);
void orc_geom_free(orc_geom *g);
int  orc_geom_info(const orc_geom *g, int *nv, int *nt, int *nf_int, int *nf_tot);
/* copy topology out (any pointer may be NULL) */
void orc_geom_faces(const orc_geom *g, int *fV, int *fQP, int *fQM, int *fPT,
                    int *fMT, int *fPI, int *fMI, double *fArea, int *tetFI);

int orc_compare_bfs(const orc_geom *g, int a, int b);
int orc_compare_tets(const orc_geom *g, int ta, int tb, int *OVIA, int *OVIB);

/* ---------------- cubature / kernels ---------------------------------- */
const double *orc_tetcr(int npts);                 /* {u1,u2,u3,w} x npts */
void orc_exprel(double zr, double zi, int nmin, int nmax, double *tab); /* 2*(nmax-nmin+1) */

/* Taylor-Duffy static moments for a tetrahedron pair with ncv shared
 * vertices (2,3,4).  VA/VB: 4 vertices each (12 doubles), shared ones first
 * and in the same order.  cA/cB: reference points for the moments.
 * out[p][k], p = rpower+1 (r^-1, r^0, r^1), k = {M0, Mx(3), Mx'(3), Mxx'}
 * with Mx = int (x - cA) r^p etc.  norder = Gauss points per base dim.   */
void orc_td_moments(int ncv, const double *VA, const double *VB,
                    const double *cA, const double *cB, int norder,
                    double out[3][8]);

/* static FIBBI record for BF pair (a,b) (GMatrixElements.cc:295-375)     */
void orc_fibbi(const orc_geom *g, int a, int b, int td_order, double G[6]);

/* one G-matrix element (GMatrixElements.cc:380-424); fibbi may be NULL
 * for far pairs (ncv==0).                                               */
void orc_gme(const orc_geom *g, int a, int b, double omr, double omi,
             const double *fibbi, double out[2]);

/* 4x4 complex block of a tetrahedron pair (oracle_surr.c): out[2*(4i+j)]
 * = G_ij, face i of VS / face j of VT, product TetCR(npts) rule on each
 * tetrahedron refined `level` times by 8-way subdivision.                */
void orc_tp_block(const double *VS, const double *VT, double kr, double ki, int npts, int level,
                  double out[32]);

/* ---------------- materials ------------------------------------------- */
/* kind: 0 const isotropic eps=p[0]+i p[1]
 *       1 const diagonal   eps_xx=p0+ip1, yy=p2+ip3, zz=p4+ip5
 *       2 eps(x) = p0 + p2*sin(pi*x)*cos(pi*y) + i p1   (isotropic)
 *       3 Janus: z>0 -> p0+ip1, else p2+ip3            (isotropic)     */
typedef struct { int kind; double p[6]; } orc_material;
void orc_eps(const orc_material *m, const double *x, double eps[3][3][2]);

/* ---------------- assembly ----------------------------------------------*/
/* touching BF pairs (a<=b, ncv>0), CSR by row a                          */
int  orc_touch_pairs(const orc_geom *g, int *rowptr /*nf+1*/, int *cols /*or NULL*/);
/* static table for all touching pairs, [nnz][6] in CSR order            */
void orc_static_table(const orc_geom *g, const int *rowptr, const int *cols,
                      int td_order, double *table);
/* full VIE matrix, column-major complex128 N x N                         */
void orc_assemble(const orc_geom *g, double omr, double omi, const orc_material *mat,
                  const int *rowptr, const int *cols, const double *table,
                  double *M);
void orc_assemble_rows(const orc_geom *g, double omr, double omi, int nrows, const int *rows,
                       const int *rowptr, const int *cols, const double *table, double *out);
/* only the Gram / chi^-1 entries of row a (VIEMatrix.cc:200-316)         */
int  orc_overlaps(const orc_geom *g, int a, double omega, const orc_material *mat,
                  int idx[7], double val[7][2]);

/* plane waves: pol[3*2] complex E0 (re,im interleaved per component),
 * dir[3]; field E = E0 exp(i k dir.x), k = omega.                        */
typedef struct { double E0[6]; double dir[3]; } orc_planewave;
void orc_rhs(const orc_geom *g, double omega, const orc_planewave *pw, int npw,
             double *B /* N x npw column-major complex */);
/* EMT power (EMTPFT.cc:151-211,216-319,324-546): out[2*j] = Pabs, [2*j+1] = Pscat */
void orc_emt_power(const orc_geom *g, double omega, const orc_planewave *pw, int npw,
                   const double *J /* N x npw */, double *out);
void orc_emt_power_dr(const orc_geom *g, double omega, const double *DR, double *out);

#ifdef __cplusplus
}
#endif
