// bfa_neq.hip -- buff-neq's dressed Rytov matrix and the EMT power it feeds.
//
//   bfa_gram_inverse   S^-1, S the SWG overlap (Gram) matrix of GetOverlapElements
//                      (SWGVolume.cc:405-480), inverted once per geometry as
//                      CreateBNEQData does (CreateBNEQData.cc:119-142)
//   bfa_dressed_rytov  ComputeDressedRytovMatrix (GetFlux.cc:155-262):
//                      W^-1 = 1 + S^-1 V S^-1 G, DR = W (S^-1 R S^-1)^H
//   bfa_emt_power_dr   GetEMTPFT with a DRMatrix and no incident field
//                      (EMTPFT.cc:389-476, GetJJ GetPFT.cc:108-115):
//                      sum over a <= b of Re(Z0 DR[b,a]) Im Q_ab
//
// Every product with S^-1 is a dense GEMM on the MFMA path (C -= A B into a
// zeroed C, signs folded into the sparse steps); V and R are the <= 7-entry
// overlap rows, applied by sparse-row kernels.  The reference's step (k)
// (W applied to S^-1 R, GetFlux.cc:257) produces a matrix nothing reads and is
// not performed.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bfa_internal.h"

using namespace bfa;

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? BFA_ERR_OOM : BFA_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

namespace bfa {
int overlap_rows_device(bfa_geometry *g, double omega, const bfa_material *eps, const bfa_material *temperature,
                        double t_env, int *dcnt, int *dcols, void *dV, void *dVI, double *dR, hipStream_t st);
void emt_q_matrix(const bfa_geometry *g, double omega, double *Q, int64_t ldq, hipStream_t st);
}  // namespace bfa

namespace {

constexpr int kRow = 7;                      // MAXOVERLAP: entries per overlap row
constexpr double kZvac = 376.73031346177;

struct zc { double re, im; };

__global__ void k_zero(zc *M, int64_t ld, int m, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = blockIdx.y; c < n; c += gridDim.y)
    if (r < m) M[(size_t)c * ld + r] = {0.0, 0.0};
}

__global__ void k_add_diag(zc *M, int64_t ld, int n, double alpha)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) M[(size_t)a * ld + a].re += alpha;
}

// Row a of the Gram matrix (GetOverlapElements' closed forms): the diagonal
// from both tetrahedra of BF a, then each other interior face of the
// tetrahedron, with L_i = V_i - Q_A for face a's vertices and X - Q_A its
// centroid.  The Area of a face is recovered from the signed prefactor
// +-Area/(3 V) of its side.
__global__ void k_gram_rows(DevGeom g, zc *S, int64_t ld)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= g.nbf) return;
  const int *fv = g.bfV + 5 * a;
  double diag = 0.0;
  for (int side = 0; side < 2; side++) {
    const double sA = side == 0 ? 1.0 : -1.0;
    const int t = g.bfT[2 * a + side];
    const double vol = g.tetVol[t];
    const double *QA = g.bfQ + 6 * a + 3 * side;
    const double areaA = sA * g.bfPF[2 * a + side] * 3.0 * vol;
    double xmq[3] = {0.0, 0.0, 0.0}, l2 = 0.0;
    for (int i = 1; i <= 3; i++) {
      const double *Vi = g.V + 3 * fv[i];
      for (int d = 0; d < 3; d++) {
        const double L = Vi[d] - QA[d];
        l2 += L * L;
        xmq[d] += Vi[d];
      }
    }
    for (int d = 0; d < 3; d++) xmq[d] = xmq[d] / 3.0 - QA[d];
    const double xmq2 = xmq[0] * xmq[0] + xmq[1] * xmq[1] + xmq[2] * xmq[2];
    diag += areaA * areaA * (xmq2 / 20.0 + l2 / 180.0) / vol;
    for (int f = 0; f < 4; f++) {
      const int b = g.tetFI[4 * t + f];
      if (b == a || b >= g.nbf) continue;
      const int sb = g.bfT[2 * b] == t ? 0 : 1;
      const double sB = sb == 0 ? 1.0 : -1.0;
      const double *QB = g.bfQ + 6 * b + 3 * sb;
      const double areaB = sB * g.bfPF[2 * b + sb] * 3.0 * vol;
      const double pre = 2.0 * sA * sB * areaA * areaB / (3.0 * vol);
      const double xd = xmq[0] * (QA[0] - QB[0]) + xmq[1] * (QA[1] - QB[1]) + xmq[2] * (QA[2] - QB[2]);
      S[(size_t)b * ld + a] = {pre * (3.0 * xmq2 / 40.0 + l2 / 120.0 + xd / 8.0), 0.0};
    }
  }
  S[(size_t)a * ld + a] = {diag, 0.0};
}

// Y(a, c) = scale * sum_k V[a][k] X(cols[a][k], c): a sparse row block times
// a dense matrix (one thread per output, columns strided over grid.y)
__global__ void k_sparse_rows(int n, const int *cnt, const int *cols, const zc *V, const zc *X, int64_t ldx, zc *Y,
                              int64_t ldy, double scale)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int na = cnt[a];
  int ci[kRow];
  zc vi[kRow];
  for (int k = 0; k < kRow; k++) {
    ci[k] = k < na ? cols[(size_t)a * kRow + k] : 0;
    vi[k] = k < na ? V[(size_t)a * kRow + k] : zc{0.0, 0.0};
  }
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    const zc *Xc = X + (size_t)c * ldx;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < na; k++) {
      const zc x = Xc[ci[k]];
      re += vi[k].re * x.re - vi[k].im * x.im;
      im += vi[k].re * x.im + vi[k].im * x.re;
    }
    Y[(size_t)c * ldy + a] = {scale * re, scale * im};
  }
}

// RT[c][k] = R(cols[c][k], c): column c of the real overlap block R read from
// the rows (the overlap pattern is symmetric: b is in row c iff c is in row b)
__global__ void k_overlap_cols(int n, const int *cnt, const int *cols, const double *R, double *RT)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int k = 0; k < kRow; k++) {
    double v = 0.0;
    if (k < cnt[c]) {
      const int b = cols[(size_t)c * kRow + k];
      for (int j = 0; j < cnt[b]; j++)
        if (cols[(size_t)b * kRow + j] == c) v = R[(size_t)b * kRow + j];
    }
    RT[(size_t)c * kRow + k] = v;
  }
}

// Y(r, c) = scale * sum_k X(r, cols[c][k]) RT[c][k]: a dense matrix times a
// sparse real block (one thread per output row, columns over grid.y)
__global__ void k_sparse_cols(int n, const int *cnt, const int *cols, const double *RT, const zc *X, int64_t ldx,
                              zc *Y, int64_t ldy, double scale)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = blockIdx.y; c < n; c += gridDim.y) {
    const int nc = cnt[c];
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nc; k++) {
      const zc x = X[(size_t)cols[(size_t)c * kRow + k] * ldx + r];
      const double w = RT[(size_t)c * kRow + k];
      re += x.re * w;
      im += x.im * w;
    }
    Y[(size_t)c * ldy + r] = {scale * re, scale * im};
  }
}

// M := M^H in place, 32 x 32 tiles: the workgroup of tile pair (bi, bj),
// bi <= bj, reads both tiles into LDS, then writes each one's conjugate
// transpose into the other's place (a diagonal tile into itself)
__global__ __launch_bounds__(256) void k_adjoint(int n, zc *M, int64_t ld)
{
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  __shared__ zc A[32][33], B[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int k = ty; k < 32; k += 8) {
    const int r1 = bi * 32 + tx, c1 = bj * 32 + k;
    A[k][tx] = (r1 < n && c1 < n) ? M[(size_t)c1 * ld + r1] : zc{0.0, 0.0};
    const int r2 = bj * 32 + tx, c2 = bi * 32 + k;
    B[k][tx] = (r2 < n && c2 < n) ? M[(size_t)c2 * ld + r2] : zc{0.0, 0.0};
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r1 = bi * 32 + tx, c1 = bj * 32 + k;
    if (r1 < n && c1 < n) M[(size_t)c1 * ld + r1] = {B[tx][k].re, -B[tx][k].im};
    const int r2 = bj * 32 + tx, c2 = bi * 32 + k;
    if (r2 < n && c2 < n) M[(size_t)c2 * ld + r2] = {A[tx][k].re, -A[tx][k].im};
  }
}

// part[a] = sum over b >= a of Re DR(b, a) Q(b, a): one workgroup per column
__global__ __launch_bounds__(256) void k_emt_dr(int n, const zc *DR, int64_t ld, const double *Q, int64_t ldq,
                                                double *part)
{
  const int a = blockIdx.x;
  const zc *d = DR + (size_t)a * ld;
  const double *q = Q + (size_t)a * ldq;
  double s = 0.0;
  for (int b = a + threadIdx.x; b < n; b += blockDim.x) s += d[b].re * q[b];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[a] = red[0];
}

dim3 col_grid(int m, int n) { return dim3((unsigned)((m + 255) / 256), (unsigned)std::min(n, 65535)); }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// layout of the dressed-Rytov workspace
struct NeqWork {
  zc *W1, *W2;
  int *piv, *cnt, *cols;
  zc *V, *VI;
  double *R, *RT;
};

size_t neq_layout(int n, int64_t ld, char *base, NeqWork *w)
{
  const size_t mat = align256((size_t)ld * n * sizeof(zc));
  const size_t rows = (size_t)n * kRow;
  size_t off = 0;
  auto take = [&](size_t bytes) { char *p = base ? base + off : nullptr; off += align256(bytes); return p; };
  char *w1 = take(mat), *w2 = take(mat), *piv = take((size_t)n * 4), *cnt = take((size_t)n * 4);
  char *cols = take(rows * 4), *V = take(rows * 16), *VI = take(rows * 16), *R = take(rows * 8), *RT = take(rows * 8);
  if (w) *w = {(zc *)w1, (zc *)w2, (int *)piv, (int *)cnt, (int *)cols, (zc *)V, (zc *)VI, (double *)R, (double *)RT};
  return off;
}

}  // namespace

extern "C" size_t bfa_dressed_rytov_work_bytes(int n, int64_t ld)
{
  if (n <= 0 || ld < n) return 0;
  return neq_layout(n, ld, nullptr, nullptr);
}

extern "C" int bfa_gram_inverse(bfa_geometry *g, void *d_Sinv, int64_t ld, void *stream)
{
  if (!g || !d_Sinv) return fail(BFA_ERR_ARG, "bfa_gram_inverse: null argument");
  if (!g->uploaded) return fail(BFA_ERR_STATE, "bfa_gram_inverse: geometry not uploaded");
  const int n = g->nbf;
  if (ld < n) return fail(BFA_ERR_ARG, "bfa_gram_inverse: ld < N");
  hipStream_t st = (hipStream_t)stream;
  zc *S = nullptr;
  int *piv = nullptr;
  HIPCHECK(hipMalloc(&S, (size_t)n * n * sizeof(zc)));
  if (hipMalloc(&piv, (size_t)n * 4) != hipSuccess) {
    (void)hipFree(S);
    return fail(BFA_ERR_OOM, "bfa_gram_inverse: allocation");
  }
  hipLaunchKernelGGL(k_zero, col_grid(n, n), dim3(256), 0, st, S, (int64_t)n, n, n);
  hipLaunchKernelGGL(k_gram_rows, dim3((n + 127) / 128), dim3(128), 0, st, g->dg, S, (int64_t)n);
  hipLaunchKernelGGL(k_zero, col_grid(n, n), dim3(256), 0, st, (zc *)d_Sinv, ld, n, n);
  hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, st, (zc *)d_Sinv, ld, n, 1.0);
  int info = 0;
  int rc = bfa_lu_factor_z(S, n, n, piv, &info, stream);
  if (!rc && info) rc = fail(BFA_ERR_SINGULAR, "bfa_gram_inverse: overlap matrix singular at %d", info);
  if (!rc) rc = bfa_lu_solve_z(S, n, n, piv, d_Sinv, n, ld, stream);
  (void)hipFree(S);
  (void)hipFree(piv);
  return rc;
}

extern "C" int bfa_dressed_rytov(bfa_geometry *g, double omega, const bfa_material *eps,
                                 const bfa_material *temperature, double t_env, const void *d_Sinv, void *d_DR,
                                 int64_t ld, void *d_work, void *stream)
{
  if (!g || !eps || !d_Sinv || !d_DR || !d_work) return fail(BFA_ERR_ARG, "bfa_dressed_rytov: null argument");
  if (!g->uploaded) return fail(BFA_ERR_STATE, "bfa_dressed_rytov: geometry not uploaded");
  if (omega == 0.0) return fail(BFA_ERR_ARG, "bfa_dressed_rytov: omega = 0");
  const int n = g->nbf;
  if (ld < n) return fail(BFA_ERR_ARG, "bfa_dressed_rytov: ld < N");
  hipStream_t st = (hipStream_t)stream;
  NeqWork w;
  neq_layout(n, ld, (char *)d_work, &w);
  const zc *Sinv = (const zc *)d_Sinv;
  zc *DR = (zc *)d_DR;
  const dim3 cg = col_grid(n, n), zg = col_grid(n, n);
  // (a) DR <- G
  int rc = bfa_assemble_g(g, omega, 0.0, d_DR, ld, stream);
  if (rc) return rc;
  // (b) W1 <- -S^-1 G
  hipLaunchKernelGGL(k_zero, zg, dim3(256), 0, st, w.W1, ld, n, n);
  if ((rc = bfa_zgemm_sub(n, n, n, Sinv, ld, DR, ld, w.W1, ld, stream))) return rc;
  // (c, d) DR <- V W1 = -V S^-1 G
  if ((rc = overlap_rows_device(g, omega, eps, temperature, t_env, w.cnt, w.cols, w.V, w.VI, w.R, st))) return rc;
  hipLaunchKernelGGL(k_sparse_rows, cg, dim3(256), 0, st, n, w.cnt, w.cols, w.V, w.W1, ld, DR, ld, 1.0);
  // (e, f) W2 <- S^-1 V S^-1 G + 1, then (g) its LU
  hipLaunchKernelGGL(k_zero, zg, dim3(256), 0, st, w.W2, ld, n, n);
  if ((rc = bfa_zgemm_sub(n, n, n, Sinv, ld, DR, ld, w.W2, ld, stream))) return rc;
  hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, st, w.W2, ld, n, 1.0);
  int info = 0;
  if ((rc = bfa_lu_factor_z(w.W2, n, ld, w.piv, &info, stream))) return rc;
  if (info) return fail(BFA_ERR_SINGULAR, "bfa_dressed_rytov: 1 + S^-1 V S^-1 G singular at %d", info);
  // (h, i) W1 <- -S^-1 R ; (j) DR <- S^-1 R S^-1
  hipLaunchKernelGGL(k_overlap_cols, dim3((n + 255) / 256), dim3(256), 0, st, n, w.cnt, w.cols, w.R, w.RT);
  hipLaunchKernelGGL(k_sparse_cols, cg, dim3(256), 0, st, n, w.cnt, w.cols, w.RT, Sinv, ld, w.W1, ld, -1.0);
  hipLaunchKernelGGL(k_zero, zg, dim3(256), 0, st, DR, ld, n, n);
  if ((rc = bfa_zgemm_sub(n, n, n, w.W1, ld, Sinv, ld, DR, ld, stream))) return rc;
  // (l) DR <- DR^H ; (m) DR <- W DR
  const int nt = (n + 31) / 32;
  hipLaunchKernelGGL(k_adjoint, dim3(nt, nt), dim3(32, 8), 0, st, n, DR, ld);
  HIPCHECK(hipGetLastError());
  return bfa_lu_solve_z(w.W2, n, ld, w.piv, d_DR, n, ld, stream);
}

extern "C" int bfa_emt_power_dr(bfa_geometry *g, double omega, const void *d_DR, int64_t ld, double *out,
                                void *stream)
{
  if (!g || !d_DR || !out) return fail(BFA_ERR_ARG, "bfa_emt_power_dr: null argument");
  if (!g->uploaded) return fail(BFA_ERR_STATE, "bfa_emt_power_dr: geometry not uploaded");
  if (omega == 0.0) return fail(BFA_ERR_ARG, "bfa_emt_power_dr: omega = 0");
  const int n = g->nbf;
  if (ld < n) return fail(BFA_ERR_ARG, "bfa_emt_power_dr: ld < N");
  hipStream_t st = (hipStream_t)stream;
  double *Q = nullptr, *part = nullptr;
  HIPCHECK(hipMalloc(&Q, (size_t)n * n * 8));
  if (hipMalloc(&part, (size_t)n * 8) != hipSuccess) {
    (void)hipFree(Q);
    return fail(BFA_ERR_OOM, "bfa_emt_power_dr: allocation");
  }
  emt_q_matrix(g, omega, Q, n, st);
  hipLaunchKernelGGL(k_emt_dr, dim3(n), dim3(256), 0, st, n, (const zc *)d_DR, ld, (const double *)Q, (int64_t)n,
                     part);
  std::vector<double> h(n);
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = hipMemcpyAsync(h.data(), part, (size_t)n * 8, hipMemcpyDeviceToHost, st);
  hipError_t e3 = hipStreamSynchronize(st);
  (void)hipFree(Q);
  (void)hipFree(part);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(BFA_ERR_HIP, "bfa_emt_power_dr kernels failed");
  double s = 0.0;
  for (int a = 0; a < n; a++) s += h[a];
  // no incident field: P_abs = -P_scat (EMTPFT.cc:444-462)
  out[0] = -kZvac * s;
  out[1] = kZvac * s;
  return BFA_OK;
}
