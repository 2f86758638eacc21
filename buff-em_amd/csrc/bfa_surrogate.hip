// bfa_surrogate.hip -- learned near-field surrogate for the far-field
// (non-touching) tetrahedron pairs of the VIE matrix.
//
// A far matrix element is the signed sum of four tetrahedron-pair entries
// (GetGME_BFBFInt, src/libs/libbuff/GMatrixElements.cc:270-288); the
// reference integrates each with the fixed 4-point product rule
// (NumPts=4, GMatrixElements.cc:282), whose error grows for close pairs and
// large kr.  Here every non-touching pair closer than `factor` tetrahedron
// diameters gets a 4x4 complex block
//     G_ij = int_S int_T [f_i.g_j - div f_i div g_j / k^2] exp(ikr)/(4 pi r)
// (face i of S, face j of T, SWG restrictions without the +/- sign) from:
//   rule       the reference's 4-point product rule,
//   cubature   a product TetCR rule on regularly subdivided tetrahedra (the
//              accurate integral: training labels, fallback, yardstick),
//   surrogate  the 16-point product rule R16 plus the ensemble mean of an
//              MLP that predicts R16's remaining error in units of the
//              per-pair error estimate ||R16 - R4||, from the pair geometry
//              canonicalised to the source frame, kr, and the direction and
//              size of R16 - R4 (a learned extrapolation of the rule
//              sequence); cubature wherever the members disagree by more
//              than a tolerance or the input lies outside the training range.
// The far-field assembly then reads the blocks instead of integrating.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "bfa_internal.h"
#include "bfa_mlp.h"
#include "bfa_rules.h"

using namespace bfa;

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? BFA_ERR_OOM : BFA_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

#define FOURPI 12.566370614359172954
#define NGEO 19           // 12 target coordinates + 6 source shape + log(k h)
#define NFEAT 52          // NGEO + unit (R16 - R4) (32) + log10 ||R16 - R4|| / ||R16||
#define NOUT 32           // 16 complex block entries
#define MAXPTS 33

namespace {

struct cd { double re, im; };

// ---------------------------------------------------------------------------
// geometry helpers (device)
__device__ __forceinline__ void d_sub(const double *a, const double *b, double *o)
{ o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
__device__ __forceinline__ double d_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void d_cross(const double *a, const double *b, double *o)
{
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ double d_absvol(const double *v)
{
  double a[3], b[3], c[3], x[3];
  d_sub(v + 3, v, a); d_sub(v + 6, v, b); d_sub(v + 9, v, c);
  d_cross(b, c, x);
  return fabs(d_dot(a, x)) / 6.0;
}
__device__ double d_tri_area(const double *p, const double *q, const double *r)
{
  double a[3], b[3], x[3];
  d_sub(q, p, a); d_sub(r, p, b);
  d_cross(a, b, x);
  return 0.5 * sqrt(d_dot(x, x));
}

// child `c` (0..7) of a regular subdivision: 4 corner tetrahedra, then the
// octahedron split around its x02-x13 diagonal (same table as the oracle)
__constant__ int c_child[8][4][2] = {
    {{0, 0}, {0, 1}, {0, 2}, {0, 3}}, {{0, 1}, {1, 1}, {1, 2}, {1, 3}},
    {{0, 2}, {1, 2}, {2, 2}, {2, 3}}, {{0, 3}, {1, 3}, {2, 3}, {3, 3}},
    {{0, 2}, {1, 3}, {0, 1}, {0, 3}}, {{0, 2}, {1, 3}, {0, 1}, {1, 2}},
    {{0, 2}, {1, 3}, {2, 3}, {0, 3}}, {{0, 2}, {1, 3}, {2, 3}, {1, 2}}};

__device__ void d_subtet(const double *v, int idx, int level, double *out)
{
  double cur[12], nxt[12];
  for (int i = 0; i < 12; i++) cur[i] = v[i];
  // digit of the first subdivision is the most significant (oracle order)
  for (int l = level - 1; l >= 0; l--) {
    const int c = (idx >> (3 * l)) & 7;
    for (int k = 0; k < 4; k++)
      for (int d = 0; d < 3; d++)
        nxt[3 * k + d] = 0.5 * (cur[3 * c_child[c][k][0] + d] + cur[3 * c_child[c][k][1] + d]);
    for (int i = 0; i < 12; i++) cur[i] = nxt[i];
  }
  for (int i = 0; i < 12; i++) out[i] = cur[i];
}

__constant__ double c_rule4[4 * 4], c_rule16[16 * 4], c_rule33[33 * 4];

// ---------------------------------------------------------------------------
// Block of tetrahedron pairs: one workgroup per (pair, frequency) sample.
// Points of both (subdivided) tetrahedra are generated into LDS; threads
// split the source points, accumulate the 8 complex moments about the
// centroids, reduce, and thread 0 forms the 16 entries.
struct BlockJob {
  const int *ps, *pt;           // pair -> tetrahedra
  const int *sample_pair;       // sample -> pair (NULL: sample = pair)
  const double *sample_omega;   // sample -> omega (NULL: omega0)
  const int *sample_list;       // optional indirection of the launch index
  double omega0;
  const double *V;              // vertices [3 nv]
  const int *tetVI;             // [4 nt]
  double *out;                  // [sample][32] (indexed by the launch index's sample)
};

__global__ __launch_bounds__(256) void k_tp_block(BlockJob J, int npts, int level)
{
  extern __shared__ double lds[];
  const int nsub = 1 << (3 * level);
  const int np = npts * nsub;
  double *PS = lds;                 // [np][4]
  double *PT = lds + 4 * np;        // [np][4]
  __shared__ double red[16][256 / 64];
  __shared__ double VS[12], VT[12];
  const int li = blockIdx.x;
  const int sidx = J.sample_list ? J.sample_list[li] : li;
  const int pair = J.sample_pair ? J.sample_pair[sidx] : sidx;
  const double omega = J.sample_omega ? J.sample_omega[sidx] : J.omega0;
  const int s = J.ps[pair], t = J.pt[pair];
  const int tid = threadIdx.x;
  if (tid < 12) {
    VS[tid] = J.V[3 * J.tetVI[4 * s + tid / 3] + tid % 3];
    VT[tid] = J.V[3 * J.tetVI[4 * t + tid / 3] + tid % 3];
  }
  __syncthreads();
  const double *rule = npts == 4 ? c_rule4 : npts == 16 ? c_rule16 : c_rule33;
  for (int idx = tid; idx < 2 * np; idx += blockDim.x) {
    const int which = idx / np, k = idx % np;
    const int sub = k / npts, q = k % npts;
    double sv[12];
    d_subtet(which ? VT : VS, sub, level, sv);
    const double u1 = rule[4 * q], u2 = rule[4 * q + 1], u3 = rule[4 * q + 2], w = rule[4 * q + 3];
    double *o = (which ? PT : PS) + 4 * k;
    for (int d = 0; d < 3; d++)
      o[d] = sv[d] + u1 * (sv[3 + d] - sv[d]) + u2 * (sv[6 + d] - sv[d]) + u3 * (sv[9 + d] - sv[d]);
    o[3] = 6.0 * d_absvol(sv) * w;
  }
  __syncthreads();
  double cS[3], cT[3];
  for (int d = 0; d < 3; d++) {
    cS[d] = 0.25 * (VS[d] + VS[3 + d] + VS[6 + d] + VS[9 + d]);
    cT[d] = 0.25 * (VT[d] + VT[3 + d] + VT[6 + d] + VT[9 + d]);
  }
  // moments: I0, Ix[3], Iy[3], Ixy (re, im) = 16 doubles
  double m[16];
  for (int i = 0; i < 16; i++) m[i] = 0.0;
  const double kr = omega;
  for (int a = tid; a < np; a += blockDim.x) {
    const double *x = PS + 4 * a;
    double xi[3];
    d_sub(x, cS, xi);
    for (int b = 0; b < np; b++) {
      const double *y = PT + 4 * b;
      double R[3], eta[3];
      d_sub(x, y, R);
      d_sub(y, cT, eta);
      const double r = sqrt(d_dot(R, R));
      const double amp = x[3] * y[3] / (FOURPI * r);
      double sn, cs;
      sincos(kr * r, &sn, &cs);
      const double gr = amp * cs, gi = amp * sn;
      m[0] += gr; m[1] += gi;
      for (int d = 0; d < 3; d++) {
        m[2 + 2 * d] += gr * xi[d]; m[3 + 2 * d] += gi * xi[d];
        m[8 + 2 * d] += gr * eta[d]; m[9 + 2 * d] += gi * eta[d];
      }
      const double xy = d_dot(xi, eta);
      m[14] += gr * xy; m[15] += gi * xy;
    }
  }
  // workgroup reduction (wave shuffles, then across waves in fixed order)
  for (int i = 0; i < 16; i++) {
    double v = m[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((tid & 63) == 0) red[i][tid >> 6] = v;
  }
  __syncthreads();
  if (tid == 0) {
    double M[16];
    for (int i = 0; i < 16; i++) {
      double v = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += red[i][w];
      M[i] = v;
    }
    double AS[4], AT[4];
    AS[0] = d_tri_area(VS + 3, VS + 6, VS + 9); AS[1] = d_tri_area(VS, VS + 6, VS + 9);
    AS[2] = d_tri_area(VS, VS + 3, VS + 9);     AS[3] = d_tri_area(VS, VS + 3, VS + 6);
    AT[0] = d_tri_area(VT + 3, VT + 6, VT + 9); AT[1] = d_tri_area(VT, VT + 6, VT + 9);
    AT[2] = d_tri_area(VT, VT + 3, VT + 9);     AT[3] = d_tri_area(VT, VT + 3, VT + 6);
    const double vs = d_absvol(VS), vt = d_absvol(VT);
    const double ik2 = 1.0 / (kr * kr);
    double *o = J.out + (size_t)sidx * NOUT;
    for (int i = 0; i < 4; i++) {
      double ai[3];
      d_sub(VS + 3 * i, cS, ai);
      const double csi = AS[i] / (3.0 * vs), dsi = AS[i] / vs;
      for (int j = 0; j < 4; j++) {
        double bj[3];
        d_sub(VT + 3 * j, cT, bj);
        const double ctj = AT[j] / (3.0 * vt), dtj = AT[j] / vt;
        double re = M[14], im = M[15];
        for (int d = 0; d < 3; d++) {
          re -= bj[d] * M[2 + 2 * d] + ai[d] * M[8 + 2 * d];
          im -= bj[d] * M[3 + 2 * d] + ai[d] * M[9 + 2 * d];
        }
        const double ab = d_dot(ai, bj);
        re = (re + ab * M[0]) * csi * ctj - dsi * dtj * M[0] * ik2;
        im = (im + ab * M[1]) * csi * ctj - dsi * dtj * M[1] * ik2;
        o[2 * (4 * i + j)] = re;
        o[2 * (4 * i + j) + 1] = im;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Canonical features of (pair, omega) in the source frame (origin = source
// centroid, e1 along v1-v0, e2 in the v0 v1 v2 plane, e3 = e1 x e2), lengths
// over the source's mean edge h: the 4 target vertices (12), the source's
// own shape v1-v0 = (a,0,0), v2-v0 = (b,c,0), v3-v0 = (d,e,f) (6), log(k h).
__device__ void d_features(const double *VS, const double *VT, double omega, float *f)
{
  double c[3], e1[3], e2[3], e3[3], u[3];
  for (int d = 0; d < 3; d++) c[d] = 0.25 * (VS[d] + VS[3 + d] + VS[6 + d] + VS[9 + d]);
  double h = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++) {
      double x[3];
      d_sub(VS + 3 * j, VS + 3 * i, x);
      h += sqrt(d_dot(x, x));
    }
  h /= 6.0;
  d_sub(VS + 3, VS, e1);
  double n1 = sqrt(d_dot(e1, e1));
  for (int d = 0; d < 3; d++) e1[d] /= n1;
  d_sub(VS + 6, VS, u);
  const double p = d_dot(u, e1);
  for (int d = 0; d < 3; d++) e2[d] = u[d] - p * e1[d];
  const double n2 = sqrt(d_dot(e2, e2));
  for (int d = 0; d < 3; d++) e2[d] /= n2;
  d_cross(e1, e2, e3);
  for (int j = 0; j < 4; j++) {
    double x[3];
    d_sub(VT + 3 * j, c, x);
    f[3 * j] = (float)(d_dot(x, e1) / h);
    f[3 * j + 1] = (float)(d_dot(x, e2) / h);
    f[3 * j + 2] = (float)(d_dot(x, e3) / h);
  }
  double v2[3], v3[3];
  d_sub(VS + 6, VS, v2);
  d_sub(VS + 9, VS, v3);
  f[12] = (float)(n1 / h);
  f[13] = (float)(d_dot(v2, e1) / h);
  f[14] = (float)(d_dot(v2, e2) / h);
  f[15] = (float)(d_dot(v3, e1) / h);
  f[16] = (float)(d_dot(v3, e2) / h);
  f[17] = (float)(d_dot(v3, e3) / h);
  f[18] = (float)log(omega * h);
}

struct FeatJob {
  const int *ps, *pt, *sample_pair;
  const double *sample_omega;
  double omega0;
  const double *V;
  const int *tetVI;
  int n;
  const double *r4, *r16;   // rule blocks of the samples (NULL: geometric inputs only)
};

// Network inputs: the NGEO geometric ones, then the 4- and 16-point rules'
// difference as a unit vector and its log10 size relative to R16 (the
// network learns how R16's remaining error relates to that difference).
// feat: raw inputs [n][nf] (nf = NGEO without rule blocks, else NFEAT);
// x16: standardised FP16 inputs [n][kp].
__global__ void k_features(FeatJob J, float *feat, _Float16 *x16, int kp, const float *mu, const float *sd)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J.n) return;
  const int pair = J.sample_pair ? J.sample_pair[i] : i;
  const double omega = J.sample_omega ? J.sample_omega[i] : J.omega0;
  double VS[12], VT[12];
  const int s = J.ps[pair], t = J.pt[pair];
  for (int k = 0; k < 12; k++) {
    VS[k] = J.V[3 * J.tetVI[4 * s + k / 3] + k % 3];
    VT[k] = J.V[3 * J.tetVI[4 * t + k / 3] + k % 3];
  }
  float f[NFEAT];
  d_features(VS, VT, omega, f);
  const int nf = J.r4 ? NFEAT : NGEO;
  if (J.r4) {
    const double *a = J.r4 + (size_t)i * NOUT, *b = J.r16 + (size_t)i * NOUT;
    double d = 0.0, nb = 0.0;
    for (int k = 0; k < NOUT; k++) {
      d += (b[k] - a[k]) * (b[k] - a[k]);
      nb += b[k] * b[k];
    }
    d = fmax(sqrt(d), 1e-300);
    for (int k = 0; k < NOUT; k++) f[NGEO + k] = (float)((b[k] - a[k]) / d);
    f[NGEO + NOUT] = (float)log10(d / sqrt(nb));
  }
  if (feat)
    for (int k = 0; k < nf; k++) feat[(size_t)i * nf + k] = f[k];
  if (x16) {
    for (int k = 0; k < kp; k++) {
      const float v = k < nf ? (f[k] - mu[k]) / sd[k] : 0.f;
      x16[(size_t)i * kp + k] = (_Float16)v;
    }
  }
}

// Per-sample scale of the residual: s = ||R16 - R4|| (the 4-point rule's
// error estimate), floored at 1e-14 ||R16||.
__device__ __forceinline__ double resid_scale(const double *r4, const double *r16)
{
  double d = 0.0, n = 0.0;
  for (int k = 0; k < NOUT; k++) {
    const double x = r16[k] - r4[k];
    d += x * x;
    n += r16[k] * r16[k];
  }
  return fmax(sqrt(d), 1e-14 * sqrt(n));
}

// training target: y = (label - R16) / s, standardised later
__global__ void k_targets(const double *label, const double *r4, const double *r16, int n, float *y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sc = resid_scale(r4 + (size_t)i * NOUT, r16 + (size_t)i * NOUT);
  for (int k = 0; k < NOUT; k++)
    y[(size_t)i * NOUT + k] = (float)((label[(size_t)i * NOUT + k] - r16[(size_t)i * NOUT + k]) / sc);
}

__global__ void k_standardise(float *y, int n, int m, const float *mu, const float *sd)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * m) return;
  const int k = (int)(i % m);
  y[i] = (y[i] - mu[k]) / sd[k];
}

// minibatch gather: X16[e][b] = src16[idx[e][b]], T[e][b] = srcT[idx[e][b]],
// Wt[e][b] = srcW[idx[e][b]]
__global__ void k_gather(const _Float16 *src16, int kp, const float *srcT, const float *srcW, const int *idx, int B,
                         _Float16 *X16, float *T, float *Wt)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x, e = blockIdx.y;
  if (b >= B) return;
  const int s = idx[(size_t)e * B + b];
  Wt[(size_t)e * B + b] = srcW[s];
  for (int k = 0; k < kp; k++) X16[((size_t)e * B + b) * kp + k] = src16[(size_t)s * kp + k];
  for (int k = 0; k < NOUT; k++) T[((size_t)e * B + b) * NOUT + k] = srcT[(size_t)s * NOUT + k];
}

// Ensemble combination: block = R16 + s * mean_e(y_e) (s = ||R16 - R4||);
// disagreement = max_e ||y_e - mean||, the members' spread about the
// correction in units of s (the scale of the residual being learned, so
// that the tolerance means the same for close and distant pairs); flag if
// > tol or an input lies outside the training range [lo, hi].
__global__ void k_combine(const float *Y, int E, int n, const float *ymu, const float *ysd, const double *r4,
                          const double *r16, const float *feat, const float *flo, const float *fhi, double tol,
                          double *blocks, float *disagree, int *flag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *a = r4 + (size_t)i * NOUT, *b = r16 + (size_t)i * NOUT;
  const double sc = resid_scale(a, b);
  double mean[NOUT];
  for (int k = 0; k < NOUT; k++) {
    double s = 0.0;
    for (int e = 0; e < E; e++) s += (double)Y[((size_t)e * n + i) * NOUT + k] * ysd[k] + ymu[k];
    mean[k] = s / E;
  }
  for (int k = 0; k < NOUT; k++) blocks[(size_t)i * NOUT + k] = b[k] + sc * mean[k];
  double worst = 0.0;
  for (int e = 0; e < E; e++) {
    double d2 = 0.0;
    for (int k = 0; k < NOUT; k++) {
      const double d = (double)Y[((size_t)e * n + i) * NOUT + k] * ysd[k] + ymu[k] - mean[k];
      d2 += d * d;
    }
    worst = fmax(worst, sqrt(d2));
  }
  bool ood = false;
  for (int k = 0; k < NFEAT; k++) {
    if (k >= NGEO && k < NGEO + NOUT) continue;   // unit-vector components are bounded anyway
    const float v = feat[(size_t)i * NFEAT + k];
    if (v < flo[k] || v > fhi[k]) ood = true;
  }
  disagree[i] = (float)worst;
  flag[i] = (worst > tol || ood) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Far-field assembly from the block table: BF pair (a, b) with no shared
// vertex = sum over its 4 tetrahedron pairs of sigma_a sigma_b G[i_a][j_b];
// pairs not in the table use the 4-point rule on the fly.
struct TabGeom {
  DevGeom g;
  const int *bfLoc;     // [N][2] local face index within P / M tetrahedron
  const int *trow;      // [nt+1] CSR over partner tetrahedra
  const int *tcol;      // partner tetrahedron (sorted per row)
  const int *tidx;      // table index * 2 + (1 if stored transposed)
  const double *blocks; // [npairs][32]
};

__device__ __forceinline__ int tab_find(const TabGeom &G, int s, int t)
{
  int lo = G.trow[s], hi = G.trow[s + 1] - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = G.tcol[mid];
    if (v == t) return G.tidx[mid];
    if (v < t) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

// 4-point rule entry of one tetrahedron pair for faces (i of s, j of t)
__device__ cd rule4_entry(const DevGeom &g, int s, int i, double pfs, const double *Qs, int t, int j, double pft,
                          const double *Qt, double kre, double kim)
{
  const double *qs = g.q4 + (size_t)s * 16, *qt = g.q4 + (size_t)t * 16;
  const double k2re = kre * kre - kim * kim, k2im = 2.0 * kre * kim;
  const double ik2d = k2re * k2re + k2im * k2im;
  const double ik2re = k2re / ik2d, ik2im = -k2im / ik2d;
  double accr = 0.0, acci = 0.0;
  for (int p = 0; p < 4; p++) {
    const double *x = qs + 4 * p;
    for (int q = 0; q < 4; q++) {
      const double *y = qt + 4 * q;
      const double R0 = x[0] - y[0], R1 = x[1] - y[1], R2 = x[2] - y[2];
      const double r = sqrt(R0 * R0 + R1 * R1 + R2 * R2);
      const double wa = x[3] * pfs, wb = y[3] * pft;
      const double dot = wa * wb * ((x[0] - Qs[0]) * (y[0] - Qt[0]) + (x[1] - Qs[1]) * (y[1] - Qt[1]) +
                                    (x[2] - Qs[2]) * (y[2] - Qt[2]));
      const double dd = 9.0 * wa * wb;
      double sn, cs;
      sincos(kre * r, &sn, &cs);
      const double amp = exp(-kim * r) / r;
      const double per = dot - dd * ik2re, pei = -dd * ik2im;
      const double er = amp * cs, ei = amp * sn;
      accr += per * er - pei * ei;
      acci += per * ei + pei * er;
    }
  }
  (void)i; (void)j;
  return {accr / FOURPI, acci / FOURPI};
}

__global__ __launch_bounds__(256) void k_far_tab(TabGeom G, double kre, double kim, cd *M, int64_t ld)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const DevGeom &g = G.g;
  const int N = g.nbf;
  const int b = (int)(idx / N), a = (int)(idx % N);
  if (b >= N || a > b) return;
  // shared-vertex test over the 5 vertices of each BF
  const int *va = g.bfV + 5 * a, *vb = g.bfV + 5 * b;
  for (int x = 0; x < 5; x++)
    for (int y = 0; y < 5; y++)
      if (va[x] == vb[y]) return;
  double re = 0.0, im = 0.0;
  for (int sa = 0; sa < 2; sa++)
    for (int sb = 0; sb < 2; sb++) {
      const int s = g.bfT[2 * a + sa], t = g.bfT[2 * b + sb];
      const int i = G.bfLoc[2 * a + sa], j = G.bfLoc[2 * b + sb];
      const double sig = (sa == sb) ? 1.0 : -1.0;
      const int ti = tab_find(G, s, t);
      if (ti >= 0) {
        const double *B = G.blocks + (size_t)(ti >> 1) * NOUT;
        const int e = (ti & 1) ? (4 * j + i) : (4 * i + j);
        re += sig * B[2 * e];
        im += sig * B[2 * e + 1];
      } else {
        const cd v = rule4_entry(g, s, i, g.bfPF[2 * a + sa], g.bfQ + 6 * a + 3 * sa, t, j, g.bfPF[2 * b + sb],
                                 g.bfQ + 6 * b + 3 * sb, kre, kim);
        // rule4_entry carries the BF signs through the prefactors already
        re += v.re;
        im += v.im;
      }
    }
  M[(size_t)b * ld + a] = {re, im};
  M[(size_t)a * ld + b] = {re, im};
}

}  // namespace

// ===========================================================================
struct bfa_surrogate {
  bfa_geometry *geo = nullptr;
  int E = 4, hidden = 256;
  double factor = 3.0;
  uint64_t seed = 1;
  MLPShape shp{};
  int npairs = 0;
  std::vector<int> ps, pt;
  // device
  int *d_ps = nullptr, *d_pt = nullptr, *d_bfLoc = nullptr, *d_trow = nullptr, *d_tcol = nullptr, *d_tidx = nullptr;
  double *d_blocks = nullptr;        // [npairs][32] current table
  // model
  float *W = nullptr, *m = nullptr, *v = nullptr, *G = nullptr;   // [E][nparam]
  _Float16 *W16 = nullptr;
  float xmu[NFEAT], xsd[NFEAT], xlo[NFEAT], xhi[NFEAT];
  float ymu[NOUT], ysd[NOUT];
  float *d_stats = nullptr;          // xmu xsd xlo xhi ymu ysd
  int adam_step = 0;
  bool trained = false;
  // dataset
  int nsamp = 0;
  int *d_spair = nullptr;
  double *d_somega = nullptr;
  _Float16 *d_X16 = nullptr;         // [nsamp][kp]
  float *d_Y = nullptr;              // [nsamp][32] standardised targets
  float *d_F = nullptr;              // [nsamp][NFEAT] raw features
  float *d_W = nullptr;              // [nsamp] loss weights
  // last prediction statistics
  int last_flagged = 0;
  double last_disagree_max = 0.0;
  ~bfa_surrogate();
};

bfa_surrogate::~bfa_surrogate()
{
  void *ptrs[] = {d_ps, d_pt, d_bfLoc, d_trow, d_tcol, d_tidx, d_blocks, W, m, v, G, W16, d_stats,
                  d_spair, d_somega, d_X16, d_Y, d_F, d_W};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
}

namespace {

int upload_rules()
{
  static bool done = false;
  if (done) return BFA_OK;
  std::vector<double> r4(16), r16(64), r33(132);
  tet_rule(4, r4.data());
  tet_rule(16, r16.data());
  tet_rule(33, r33.data());
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_rule4), r4.data(), r4.size() * 8));
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_rule16), r16.data(), r16.size() * 8));
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_rule33), r33.data(), r33.size() * 8));
  done = true;
  return BFA_OK;
}

template <typename T> int dev_copy(T **dst, const std::vector<T> &v)
{
  HIPCHECK(hipMalloc((void **)dst, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) HIPCHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return BFA_OK;
}

int launch_blocks(bfa_surrogate *S, const int *d_pair, const double *d_omega, double omega0, const int *d_list, int n,
                  int npts, int level, double *out, hipStream_t st)
{
  if (npts != 4 && npts != 16 && npts != 33) return fail(BFA_ERR_ARG, "cubature rule must have 4, 16 or 33 points");
  if (level < 0 || level > 2) return fail(BFA_ERR_ARG, "subdivision level must be 0..2");
  int rc = upload_rules();
  if (rc) return rc;
  const int np = npts << (3 * level);
  const size_t lds = (size_t)8 * np * 8;
  if (lds > 150 * 1024) return fail(BFA_ERR_ARG, "cubature rule %d x level %d too large", npts, level);
  static bool attr = false;
  if (!attr) {
    HIPCHECK(hipFuncSetAttribute((const void *)k_tp_block, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
  BlockJob J{};
  J.ps = S->d_ps; J.pt = S->d_pt; J.sample_pair = d_pair; J.sample_omega = d_omega; J.sample_list = d_list;
  J.omega0 = omega0; J.V = (const double *)S->geo->d.V; J.tetVI = (const int *)S->geo->d.tetVI; J.out = out;
  if (n <= 0) return BFA_OK;
  const int threads = np >= 256 ? 256 : 64;
  hipLaunchKernelGGL(k_tp_block, dim3(n), dim3(threads), lds, st, J, npts, level);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int bfa_surr_create(bfa_geometry *g, double factor, int ensemble, int hidden, uint64_t seed,
                               bfa_surrogate **out)
{
  if (!g || !out || factor <= 0 || ensemble <= 0 || ensemble > 16) return fail(BFA_ERR_ARG, "bfa_surr_create: bad argument");
  if (hidden <= 0 || hidden > 256 || hidden % 64) return fail(BFA_ERR_ARG, "bfa_surr_create: hidden must be 64..256 in steps of 64");
  if (!g->uploaded) return fail(BFA_ERR_STATE, "bfa_surr_create: geometry not uploaded");
  *out = nullptr;
  bfa_surrogate *S = new bfa_surrogate();
  auto bail = [&](int rc) { delete S; return rc; };
  S->geo = g;
  S->E = ensemble;
  S->hidden = hidden;
  S->factor = factor;
  S->seed = seed;
  S->shp = mlp_shape(NFEAT, hidden, NOUT);
  const int nt = g->nt;
  // centroids and diameters
  std::vector<double> cen(3 * (size_t)nt), diam(nt);
  for (int t = 0; t < nt; t++) {
    const int *vi = &g->tetVI[4 * (size_t)t];
    double dm = 0.0;
    for (int d = 0; d < 3; d++)
      cen[3 * t + d] = 0.25 * (g->V[3 * vi[0] + d] + g->V[3 * vi[1] + d] + g->V[3 * vi[2] + d] + g->V[3 * vi[3] + d]);
    for (int i = 0; i < 4; i++)
      for (int j = i + 1; j < 4; j++) {
        double s2 = 0.0;
        for (int d = 0; d < 3; d++) {
          const double x = g->V[3 * vi[i] + d] - g->V[3 * vi[j] + d];
          s2 += x * x;
        }
        dm = std::max(dm, std::sqrt(s2));
      }
    diam[t] = dm;
  }
  // near, non-touching pairs s < t
  std::vector<std::vector<std::pair<int, int>>> rows(nt);
  for (int s = 0; s < nt; s++) {
    const int *vs = &g->tetVI[4 * (size_t)s];
    for (int t = s + 1; t < nt; t++) {
      const int *vt = &g->tetVI[4 * (size_t)t];
      bool touch = false;
      for (int i = 0; i < 4 && !touch; i++)
        for (int j = 0; j < 4; j++)
          if (vs[i] == vt[j]) { touch = true; break; }
      if (touch) continue;
      double d2 = 0.0;
      for (int d = 0; d < 3; d++) {
        const double x = cen[3 * s + d] - cen[3 * t + d];
        d2 += x * x;
      }
      const double lim = factor * std::max(diam[s], diam[t]);
      if (d2 >= lim * lim) continue;
      const int p = (int)S->ps.size();
      S->ps.push_back(s);
      S->pt.push_back(t);
      rows[s].push_back({t, 2 * p});
      rows[t].push_back({s, 2 * p + 1});
    }
  }
  S->npairs = (int)S->ps.size();
  std::vector<int> trow(nt + 1, 0), tcol, tidx;
  for (int s = 0; s < nt; s++) {
    std::sort(rows[s].begin(), rows[s].end());
    trow[s + 1] = trow[s] + (int)rows[s].size();
    for (auto &x : rows[s]) { tcol.push_back(x.first); tidx.push_back(x.second); }
  }
  std::vector<int> bfLoc(2 * (size_t)g->nbf);
  for (int a = 0; a < g->nbf; a++) { bfLoc[2 * a] = g->fPI[a]; bfLoc[2 * a + 1] = g->fMI[a]; }
  int rc;
  if ((rc = dev_copy(&S->d_ps, S->ps)) || (rc = dev_copy(&S->d_pt, S->pt)) || (rc = dev_copy(&S->d_trow, trow)) ||
      (rc = dev_copy(&S->d_tcol, tcol)) || (rc = dev_copy(&S->d_tidx, tidx)) || (rc = dev_copy(&S->d_bfLoc, bfLoc)))
    return bail(rc);
  if (hipMalloc((void **)&S->d_blocks, std::max(1, S->npairs) * (size_t)NOUT * 8) != hipSuccess)
    return bail(fail(BFA_ERR_OOM, "bfa_surr_create: block table"));
  // model parameters: Kaiming-style init from the seed, zero biases
  const int64_t np = S->shp.nparam;
  std::vector<float> w((size_t)S->E * np, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int e = 0; e < S->E; e++)
    for (int l = 0; l < 4; l++) {
      const int nout = l == 3 ? NOUT : hidden, kin = l == 0 ? S->shp.kp : hidden;
      const int kreal = l == 0 ? NFEAT : hidden;   // padding columns stay zero
      const float sc = std::sqrt(2.0f / kreal) * (l == 3 ? 0.1f : 1.0f);
      for (int r = 0; r < nout; r++)
        for (int k = 0; k < kreal; k++) w[(size_t)e * np + S->shp.oW[l] + (size_t)r * kin + k] = sc * nd(rng);
    }
  if (hipMalloc((void **)&S->W, (size_t)S->E * np * 4) != hipSuccess ||
      hipMalloc((void **)&S->m, (size_t)S->E * np * 4) != hipSuccess ||
      hipMalloc((void **)&S->v, (size_t)S->E * np * 4) != hipSuccess ||
      hipMalloc((void **)&S->G, (size_t)S->E * np * 4) != hipSuccess ||
      hipMalloc((void **)&S->W16, (size_t)S->E * np * 2) != hipSuccess ||
      hipMalloc((void **)&S->d_stats, (4 * NFEAT + 2 * NOUT) * 4) != hipSuccess)
    return bail(fail(BFA_ERR_OOM, "bfa_surr_create: model"));
  if (hipMemcpy(S->W, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(S->m, 0, w.size() * 4) != hipSuccess || hipMemset(S->v, 0, w.size() * 4) != hipSuccess)
    return bail(fail(BFA_ERR_HIP, "bfa_surr_create: model upload"));
  if ((rc = to_half(S->W, S->W16, (int64_t)S->E * np, nullptr))) return bail(rc);
  for (int k = 0; k < NFEAT; k++) { S->xmu[k] = 0.f; S->xsd[k] = 1.f; S->xlo[k] = -1e30f; S->xhi[k] = 1e30f; }
  for (int k = 0; k < NOUT; k++) { S->ymu[k] = 0.f; S->ysd[k] = 1.f; }
  *out = S;
  return BFA_OK;
}

extern "C" int bfa_surr_destroy(bfa_surrogate *S)
{
  delete S;
  return BFA_OK;
}

extern "C" int bfa_surr_pairs(const bfa_surrogate *S, int *npairs, int *ps, int *pt)
{
  if (!S) return fail(BFA_ERR_ARG, "bfa_surr_pairs: null");
  if (npairs) *npairs = S->npairs;
  if (ps) memcpy(ps, S->ps.data(), S->ps.size() * 4);
  if (pt) memcpy(pt, S->pt.data(), S->pt.size() * 4);
  return BFA_OK;
}

// Blocks of selected pairs at selected frequencies by a cubature rule
// (host in/out): pair[n], omega[n] -> blocks [n][32].
extern "C" int bfa_surr_cubature(bfa_surrogate *S, const int *pair, const double *omega, int n, int npts, int level,
                                 double *blocks)
{
  if (!S || !pair || !omega || !blocks || n <= 0) return fail(BFA_ERR_ARG, "bfa_surr_cubature: bad argument");
  for (int i = 0; i < n; i++)
    if (pair[i] < 0 || pair[i] >= S->npairs || !(omega[i] > 0)) return fail(BFA_ERR_ARG, "bfa_surr_cubature: bad sample %d", i);
  int *dp = nullptr;
  double *dw = nullptr, *db = nullptr;
  HIPCHECK(hipMalloc(&dp, (size_t)n * 4));
  HIPCHECK(hipMalloc(&dw, (size_t)n * 8));
  HIPCHECK(hipMalloc(&db, (size_t)n * NOUT * 8));
  int rc = BFA_OK;
  if (hipMemcpy(dp, pair, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dw, omega, (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(BFA_ERR_HIP, "bfa_surr_cubature: upload");
  if (!rc) rc = launch_blocks(S, dp, dw, 0.0, nullptr, n, npts, level, db, nullptr);
  if (!rc && hipMemcpy(blocks, db, (size_t)n * NOUT * 8, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(BFA_ERR_HIP, "bfa_surr_cubature: download");
  (void)hipFree(dp);
  (void)hipFree(dw);
  (void)hipFree(db);
  return rc;
}

// Geometric network inputs (host out [n][NGEO]) of pair[n] at omega[n].
extern "C" int bfa_surr_features(bfa_surrogate *S, const int *pair, const double *omega, int n, float *feat)
{
  if (!S || !pair || !omega || !feat || n <= 0) return fail(BFA_ERR_ARG, "bfa_surr_features: bad argument");
  int *dp = nullptr;
  double *dw = nullptr;
  float *df = nullptr;
  HIPCHECK(hipMalloc(&dp, (size_t)n * 4));
  HIPCHECK(hipMalloc(&dw, (size_t)n * 8));
  HIPCHECK(hipMalloc(&df, (size_t)n * NGEO * 4));
  int rc = BFA_OK;
  if (hipMemcpy(dp, pair, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dw, omega, (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(BFA_ERR_HIP, "bfa_surr_features: upload");
  if (!rc) {
    FeatJob J{S->d_ps, S->d_pt, dp, dw, 0.0, (const double *)S->geo->d.V, (const int *)S->geo->d.tetVI, n,
              nullptr, nullptr};
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, nullptr, J, df, (_Float16 *)nullptr, 0,
                       (const float *)nullptr, (const float *)nullptr);
    if (hipGetLastError() != hipSuccess || hipMemcpy(feat, df, (size_t)n * NGEO * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(BFA_ERR_HIP, "bfa_surr_features: kernel");
  }
  (void)hipFree(dp);
  (void)hipFree(dw);
  (void)hipFree(df);
  return rc;
}

// Training set: every pair at `per_pair` frequencies drawn log-uniformly in
// [wmin, wmax] (seeded); features, rule blocks and cubature labels on the
// GPU; standardisation statistics and the input range kept with the model.
extern "C" int bfa_surr_dataset(bfa_surrogate *S, int per_pair, double wmin, double wmax, int label_npts,
                                int label_level)
{
  if (!S || per_pair <= 0 || !(wmin > 0) || !(wmax >= wmin)) return fail(BFA_ERR_ARG, "bfa_surr_dataset: bad argument");
  const int64_t n64 = (int64_t)S->npairs * per_pair;
  if (n64 <= 0 || n64 > (1 << 26)) return fail(BFA_ERR_ARG, "bfa_surr_dataset: %lld samples", (long long)n64);
  const int n = (int)n64;
  std::vector<int> sp(n);
  std::vector<double> sw(n);
  std::mt19937_64 rng(S->seed ^ 0xDA7A5E7ull);
  std::uniform_real_distribution<double> U(std::log(wmin), std::log(wmax));
  for (int p = 0, i = 0; p < S->npairs; p++)
    for (int r = 0; r < per_pair; r++, i++) { sp[i] = p; sw[i] = std::exp(U(rng)); }
  for (void *p : {(void *)S->d_spair, (void *)S->d_somega, (void *)S->d_X16, (void *)S->d_Y, (void *)S->d_F,
                  (void *)S->d_W})
    if (p) (void)hipFree(p);
  S->d_spair = nullptr; S->d_somega = nullptr; S->d_X16 = nullptr; S->d_Y = nullptr; S->d_F = nullptr;
  S->d_W = nullptr;
  int rc;
  if ((rc = dev_copy(&S->d_spair, sp)) || (rc = dev_copy(&S->d_somega, sw))) return rc;
  const int kp = S->shp.kp;
  double *rule = nullptr, *r16 = nullptr, *label = nullptr;
  HIPCHECK(hipMalloc(&S->d_X16, (size_t)n * kp * 2));
  HIPCHECK(hipMalloc(&S->d_Y, (size_t)n * NOUT * 4));
  HIPCHECK(hipMalloc(&S->d_F, (size_t)n * NFEAT * 4));
  HIPCHECK(hipMalloc(&rule, (size_t)n * NOUT * 8));
  HIPCHECK(hipMalloc(&r16, (size_t)n * NOUT * 8));
  HIPCHECK(hipMalloc(&label, (size_t)n * NOUT * 8));
  S->nsamp = n;
  auto go = [&]() -> int {
    int r = launch_blocks(S, S->d_spair, S->d_somega, 0.0, nullptr, n, 4, 0, rule, nullptr);
    if (r) return r;
    r = launch_blocks(S, S->d_spair, S->d_somega, 0.0, nullptr, n, 16, 0, r16, nullptr);
    if (r) return r;
    r = launch_blocks(S, S->d_spair, S->d_somega, 0.0, nullptr, n, label_npts, label_level, label, nullptr);
    if (r) return r;
    FeatJob J{S->d_ps, S->d_pt, S->d_spair, S->d_somega, 0.0, (const double *)S->geo->d.V,
              (const int *)S->geo->d.tetVI, n, rule, r16};
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, nullptr, J, S->d_F, (_Float16 *)nullptr, 0,
                       (const float *)nullptr, (const float *)nullptr);
    hipLaunchKernelGGL(k_targets, dim3((n + 255) / 256), dim3(256), 0, nullptr, (const double *)label,
                       (const double *)rule, (const double *)r16, n, S->d_Y);
    HIPCHECK(hipGetLastError());
    // statistics on the host (one pass over the raw features and targets)
    std::vector<float> F((size_t)n * NFEAT), Y((size_t)n * NOUT);
    HIPCHECK(hipMemcpy(F.data(), S->d_F, F.size() * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(Y.data(), S->d_Y, Y.size() * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < NFEAT; k++) {
      double s = 0, s2 = 0;
      float lo = 1e30f, hi = -1e30f;
      for (int i = 0; i < n; i++) {
        const double x = F[(size_t)i * NFEAT + k];
        s += x; s2 += x * x;
        lo = std::min(lo, (float)x); hi = std::max(hi, (float)x);
      }
      const double mu = s / n, var = std::max(1e-12, s2 / n - mu * mu);
      S->xmu[k] = (float)mu; S->xsd[k] = (float)std::sqrt(var);
      const float pad = 0.02f * (hi - lo);
      S->xlo[k] = lo - pad; S->xhi[k] = hi + pad;
    }
    for (int k = 0; k < NOUT; k++) {
      double s = 0, s2 = 0;
      for (int i = 0; i < n; i++) { const double y = Y[(size_t)i * NOUT + k]; s += y; s2 += y * y; }
      const double mu = s / n, var = std::max(1e-30, s2 / n - mu * mu);
      S->ymu[k] = (float)mu; S->ysd[k] = (float)std::sqrt(var);
    }
    // relative loss: sample weight 1 / (sum_k (y_ik / sd_k)^2 + 1e-2 of its
    // median), normalised to mean 1, so typical pairs are fitted to the same
    // relative accuracy as the large-error tail
    {
      std::vector<double> m2(n);
      for (int i = 0; i < n; i++) {
        double a = 0.0;
        for (int k = 0; k < NOUT; k++) {
          const double z = (double)Y[(size_t)i * NOUT + k] / S->ysd[k];
          a += z * z;
        }
        m2[i] = a;
      }
      std::vector<double> tmp(m2);
      std::nth_element(tmp.begin(), tmp.begin() + n / 2, tmp.end());
      const double floor_ = 1e-2 * std::max(tmp[n / 2], 1e-30);
      std::vector<float> wts(n);
      double sw = 0.0;
      for (int i = 0; i < n; i++) { wts[i] = (float)(1.0 / (m2[i] + floor_)); sw += wts[i]; }
      for (int i = 0; i < n; i++) wts[i] = (float)(wts[i] * n / sw);
      HIPCHECK(hipMalloc(&S->d_W, (size_t)n * 4));
      HIPCHECK(hipMemcpy(S->d_W, wts.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    std::vector<float> st;
    st.insert(st.end(), S->xmu, S->xmu + NFEAT); st.insert(st.end(), S->xsd, S->xsd + NFEAT);
    st.insert(st.end(), S->xlo, S->xlo + NFEAT); st.insert(st.end(), S->xhi, S->xhi + NFEAT);
    st.insert(st.end(), S->ymu, S->ymu + NOUT); st.insert(st.end(), S->ysd, S->ysd + NOUT);
    HIPCHECK(hipMemcpy(S->d_stats, st.data(), st.size() * 4, hipMemcpyHostToDevice));
    const float *mu = S->d_stats, *sd = S->d_stats + NFEAT, *ym = S->d_stats + 4 * NFEAT, *ys = ym + NOUT;
    hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, nullptr, J, (float *)nullptr, S->d_X16, kp,
                       mu, sd);
    hipLaunchKernelGGL(k_standardise, dim3((unsigned)(((int64_t)n * NOUT + 255) / 256)), dim3(256), 0, nullptr,
                       S->d_Y, n, NOUT, ym, ys);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
    return BFA_OK;
  };
  rc = go();
  (void)hipFree(rule);
  (void)hipFree(r16);
  (void)hipFree(label);
  return rc;
}

// Adam training of all members, each on its own shuffle of the set.
// loss_hist (host, may be NULL): [epochs][E] mean loss per epoch.
extern "C" int bfa_surr_train(bfa_surrogate *S, int epochs, int batch, double lr, double *loss_hist)
{
  if (!S || epochs <= 0 || batch <= 0) return fail(BFA_ERR_ARG, "bfa_surr_train: bad argument");
  if (S->nsamp <= 0) return fail(BFA_ERR_STATE, "bfa_surr_train: call bfa_surr_dataset first");
  const int E = S->E, kp = S->shp.kp;
  const int B = std::min(batch, S->nsamp);
  const int64_t np = S->shp.nparam;
  const int steps = S->nsamp / B;
  const int nparts = 16;
  char *mem = nullptr;
  const size_t bytes = (size_t)E * B * kp * 2 + (size_t)E * B * NOUT * 4 * 3 + (size_t)E * B * S->hidden * 3 * 4 +
                       (size_t)E * mlp_work_halfs(S->shp, B) * 2 + (size_t)E * mlp_work_floats(S->shp, B) * 4 +
                       (size_t)E * B * 4 * 2 + (size_t)E * nparts * 8 + 16 * 256;
  HIPCHECK(hipMalloc(&mem, bytes));
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = (o + 255) / 256 * 256; o = r + b; return mem + r; };
  _Float16 *X16 = (_Float16 *)take((size_t)E * B * kp * 2);
  float *T = (float *)take((size_t)E * B * NOUT * 4);
  float *Yp = (float *)take((size_t)E * B * NOUT * 4);
  float *dY = (float *)take((size_t)E * B * NOUT * 4);
  float *Z = (float *)take((size_t)E * B * S->hidden * 3 * 4);
  _Float16 *A = (_Float16 *)take((size_t)E * mlp_work_halfs(S->shp, B) * 2);
  float *work = (float *)take((size_t)E * mlp_work_floats(S->shp, B) * 4);
  int *idx = (int *)take((size_t)E * B * 4);
  float *Wt = (float *)take((size_t)E * B * 4);
  double *lp = (double *)take((size_t)E * nparts * 8);
  std::mt19937_64 rng(S->seed * 7919 + S->adam_step);
  std::vector<std::vector<int>> perm(E, std::vector<int>(S->nsamp));
  std::vector<int> hidx((size_t)E * B);
  std::vector<double> hl((size_t)E * nparts);
  int rc = BFA_OK;
  for (int ep = 0; ep < epochs && !rc; ep++) {
    for (int e = 0; e < E; e++) {
      for (int i = 0; i < S->nsamp; i++) perm[e][i] = i;
      std::shuffle(perm[e].begin(), perm[e].end(), rng);
    }
    std::vector<double> acc(E, 0.0);
    for (int s = 0; s < steps && !rc; s++) {
      for (int e = 0; e < E; e++)
        for (int b = 0; b < B; b++) hidx[(size_t)e * B + b] = perm[e][(size_t)s * B + b];
      if (hipMemcpy(idx, hidx.data(), hidx.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(BFA_ERR_HIP, "bfa_surr_train: upload");
        break;
      }
      hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256, E), dim3(256), 0, nullptr, (const _Float16 *)S->d_X16, kp,
                         (const float *)S->d_Y, (const float *)S->d_W, (const int *)idx, B, X16, T, Wt);
      if ((rc = mlp_forward(S->shp, E, B, X16, (int64_t)B * kp, S->W16, S->W, Yp, Z, A, nullptr))) break;
      if ((rc = mlp_mse(E, B, NOUT, Yp, T, Wt, dY, lp, nparts, nullptr))) break;
      if ((rc = mlp_backward(S->shp, E, B, X16, (int64_t)B * kp, S->W16, Z, A, dY, S->G, work, nullptr))) break;
      S->adam_step++;
      // cosine decay of the step size over the run
      const double frac = (ep * (double)steps + s) / ((double)epochs * steps);
      const float lr_t = (float)(lr * 0.5 * (1.0 + std::cos(M_PI * frac)));
      if ((rc = mlp_adam((int64_t)E * np, S->W, S->G, S->m, S->v, S->W16, lr_t, S->adam_step, nullptr))) break;
      if (loss_hist && (s % 16 == 0 || s == steps - 1)) {
        if (hipMemcpy(hl.data(), lp, hl.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) {
          rc = fail(BFA_ERR_HIP, "bfa_surr_train: loss download");
          break;
        }
        for (int e = 0; e < E; e++) {
          double l = 0.0;
          for (int p = 0; p < nparts; p++) l += hl[(size_t)e * nparts + p];
          acc[e] = l;   // loss of the latest sampled step
        }
      }
    }
    if (loss_hist)
      for (int e = 0; e < E; e++) loss_hist[(size_t)ep * E + e] = acc[e];
  }
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(BFA_ERR_HIP, "bfa_surr_train: kernels failed");
  (void)hipFree(mem);
  if (!rc) S->trained = true;
  return rc;
}

// Block table at one frequency.  mode 0: the 4-point rule; 1: cubature
// (npts, level) for every pair; 2: surrogate with cubature (npts, level)
// for pairs whose ensemble disagreement exceeds tol or whose features fall
// outside the training range.  Returns the number of cubature fallbacks.
extern "C" int bfa_surr_table(bfa_surrogate *S, double omega, int mode, double tol, int npts, int level,
                              int *nfallback, double *max_disagree)
{
  if (!S || !(omega > 0)) return fail(BFA_ERR_ARG, "bfa_surr_table: bad argument");
  const int n = S->npairs;
  if (n == 0) return BFA_OK;
  if (mode == 0) return launch_blocks(S, nullptr, nullptr, omega, nullptr, n, 4, 0, S->d_blocks, nullptr);
  if (mode == 1) return launch_blocks(S, nullptr, nullptr, omega, nullptr, n, npts, level, S->d_blocks, nullptr);
  if (mode != 2) return fail(BFA_ERR_ARG, "bfa_surr_table: mode %d", mode);
  if (!S->trained) return fail(BFA_ERR_STATE, "bfa_surr_table: surrogate not trained");
  const int E = S->E, kp = S->shp.kp;
  const int chunk = std::min(n, 65536);
  char *mem = nullptr;
  const size_t bytes = (size_t)2 * n * NOUT * 8 + (size_t)chunk * NFEAT * 4 + (size_t)chunk * kp * 2 +
                       (size_t)E * chunk * NOUT * 4 + (size_t)E * mlp_work_halfs(S->shp, chunk) * 2 +
                       (size_t)n * 4 * 3 + 16 * 256;
  HIPCHECK(hipMalloc(&mem, bytes));
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = (o + 255) / 256 * 256; o = r + b; return mem + r; };
  double *rule = (double *)take((size_t)n * NOUT * 8);
  double *r16 = (double *)take((size_t)n * NOUT * 8);
  float *F = (float *)take((size_t)chunk * NFEAT * 4);
  _Float16 *X16 = (_Float16 *)take((size_t)chunk * kp * 2);
  float *Y = (float *)take((size_t)E * chunk * NOUT * 4);
  _Float16 *A = (_Float16 *)take((size_t)E * mlp_work_halfs(S->shp, chunk) * 2);
  float *dis = (float *)take((size_t)n * 4);
  int *flag = (int *)take((size_t)n * 4);
  int *list = (int *)take((size_t)n * 4);
  int rc = launch_blocks(S, nullptr, nullptr, omega, nullptr, n, 4, 0, rule, nullptr);
  if (!rc) rc = launch_blocks(S, nullptr, nullptr, omega, nullptr, n, 16, 0, r16, nullptr);
  const float *mu = S->d_stats, *sd = S->d_stats + NFEAT, *lo = S->d_stats + 2 * NFEAT, *hi = S->d_stats + 3 * NFEAT;
  const float *ym = S->d_stats + 4 * NFEAT, *ys = ym + NOUT;
  for (int c0 = 0; c0 < n && !rc; c0 += chunk) {
    const int m = std::min(chunk, n - c0);
    // features of pairs c0..c0+m at omega (sample index = pair index offset)
    FeatJob J{S->d_ps + c0, S->d_pt + c0, nullptr, nullptr, omega, (const double *)S->geo->d.V,
              (const int *)S->geo->d.tetVI, m, rule + (size_t)c0 * NOUT, r16 + (size_t)c0 * NOUT};
    hipLaunchKernelGGL(k_features, dim3((m + 255) / 256), dim3(256), 0, nullptr, J, F, X16, kp, mu, sd);
    if ((rc = mlp_forward(S->shp, E, m, X16, 0, S->W16, S->W, Y, nullptr, A, nullptr))) break;
    hipLaunchKernelGGL(k_combine, dim3((m + 255) / 256), dim3(256), 0, nullptr, (const float *)Y, E, m, ym, ys,
                       (const double *)(rule + (size_t)c0 * NOUT), (const double *)(r16 + (size_t)c0 * NOUT),
                       (const float *)F, lo, hi, tol, S->d_blocks + (size_t)c0 * NOUT, dis + c0, flag + c0);
  }
  std::vector<int> hf(n);
  std::vector<float> hd(n);
  if (!rc && (hipMemcpy(hf.data(), flag, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(hd.data(), dis, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(BFA_ERR_HIP, "bfa_surr_table: download");
  int nf = 0;
  double dmax = 0.0;
  if (!rc) {
    std::vector<int> hl;
    for (int i = 0; i < n; i++) {
      dmax = std::max(dmax, (double)hd[i]);
      if (hf[i]) hl.push_back(i);
    }
    nf = (int)hl.size();
    if (nf > 0) {
      if (hipMemcpy(list, hl.data(), (size_t)nf * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(BFA_ERR_HIP, "bfa_surr_table: upload");
      // sample index = pair index; the launch index walks the fallback list
      if (!rc) rc = launch_blocks(S, nullptr, nullptr, omega, list, nf, npts, level, S->d_blocks, nullptr);
    }
  }
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(BFA_ERR_HIP, "bfa_surr_table: kernels failed");
  (void)hipFree(mem);
  S->last_flagged = nf;
  S->last_disagree_max = dmax;
  if (nfallback) *nfallback = nf;
  if (max_disagree) *max_disagree = dmax;
  return rc;
}

extern "C" int bfa_surr_table_download(const bfa_surrogate *S, double *blocks)
{
  if (!S || !blocks) return fail(BFA_ERR_ARG, "bfa_surr_table_download: bad argument");
  HIPCHECK(hipMemcpy(blocks, S->d_blocks, (size_t)S->npairs * NOUT * 8, hipMemcpyDeviceToHost));
  return BFA_OK;
}

// a block table from the host (e.g. one computed earlier, or edited), in the
// layout of bfa_surr_table_download
extern "C" int bfa_surr_table_upload(bfa_surrogate *S, const double *blocks)
{
  if (!S || !blocks) return fail(BFA_ERR_ARG, "bfa_surr_table_upload: bad argument");
  HIPCHECK(hipMemcpy(S->d_blocks, blocks, (size_t)S->npairs * NOUT * 8, hipMemcpyHostToDevice));
  return BFA_OK;
}

// VIE matrix with the far field (non-touching BF pairs) read from the block
// table (tetrahedron pairs outside it use the 4-point rule), touching pairs
// as in bfa_assemble_vie.
extern "C" int bfa_assemble_vie_table(bfa_surrogate *S, double omega, const bfa_material *mat, void *d_M,
                                      int64_t ld, void *stream)
{
  if (!S || !mat || !d_M || !(omega > 0)) return fail(BFA_ERR_ARG, "bfa_assemble_vie_table: bad argument");
  bfa_geometry *g = S->geo;
  if (!g->static_built) return fail(BFA_ERR_STATE, "bfa_assemble_vie_table: static table not built");
  if (ld < g->nbf) return fail(BFA_ERR_ARG, "bfa_assemble_vie_table: ld < N");
  hipStream_t st = (hipStream_t)stream;
  TabGeom G{g->dg, S->d_bfLoc, S->d_trow, S->d_tcol, S->d_tidx, S->d_blocks};
  const int64_t tot = (int64_t)g->nbf * g->nbf;
  hipLaunchKernelGGL(k_far_tab, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, G, omega, 0.0, (cd *)d_M, ld);
  HIPCHECK(hipGetLastError());
  return bfa::assemble_near_only(g, omega, 0.0, mat, d_M, ld, stream);
}

extern "C" int bfa_surr_info(const bfa_surrogate *S, int *ensemble, int64_t *nparam, int *nsamples, float *stats)
{
  if (!S) return fail(BFA_ERR_ARG, "bfa_surr_info: null");
  if (ensemble) *ensemble = S->E;
  if (nparam) *nparam = S->shp.nparam;
  if (nsamples) *nsamples = S->nsamp;
  if (stats) {
    memcpy(stats, S->xmu, sizeof S->xmu); memcpy(stats + NFEAT, S->xsd, sizeof S->xsd);
    memcpy(stats + 2 * NFEAT, S->xlo, sizeof S->xlo); memcpy(stats + 3 * NFEAT, S->xhi, sizeof S->xhi);
    memcpy(stats + 4 * NFEAT, S->ymu, sizeof S->ymu); memcpy(stats + 4 * NFEAT + NOUT, S->ysd, sizeof S->ysd);
  }
  return BFA_OK;
}

// Get (out != NULL) and/or set (in != NULL) the FP32 parameters [E][nparam].
extern "C" int bfa_surr_params(bfa_surrogate *S, float *out, const float *in)
{
  if (!S) return fail(BFA_ERR_ARG, "bfa_surr_params: null");
  const size_t n = (size_t)S->E * S->shp.nparam;
  if (out) HIPCHECK(hipMemcpy(out, S->W, n * 4, hipMemcpyDeviceToHost));
  if (in) {
    HIPCHECK(hipMemcpy(S->W, in, n * 4, hipMemcpyHostToDevice));
    int rc = to_half(S->W, S->W16, (int64_t)n, nullptr);
    if (rc) return rc;
    HIPCHECK(hipDeviceSynchronize());
  }
  return BFA_OK;
}
