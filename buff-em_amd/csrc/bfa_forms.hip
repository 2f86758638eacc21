// bfa_forms.hip -- the precision form of the dense solve.
//
// Complex LU in FP32 on MFMA (the complex64 instantiation of the blocked LU,
// 4-real-product trailing updates with FP32 accumulation), then FP64
// iterative refinement against the device-resident FP64 VIE matrix:
//     x_0 = 0, r_0 = b;  repeat  d = LU32^-1 r ;  x += d ;  r = b - M64 x
// until max_j ||r_j|| / ||b_j|| <= tol.  Pivot search, the residual and the
// solution stay FP64.  No reference counterpart: it replaces
// HMatrix::LUFactorize / LUSolve (buff-scatter.cc:285,314) at lower cost.
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>

#include "bfa_internal.h"
#include "bfa_lu.h"

using namespace bfa;

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? BFA_ERR_OOM : BFA_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

namespace {

struct z2 { double re, im; };
struct c2 { float re, im; };

// A32(:, c) = (complex64) A64(:, c), columns [0, n)
__global__ void k_z2c(const z2 *A, int64_t lda, c2 *B, int64_t ldb, int m, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m || c >= n) return;
  const z2 v = A[(size_t)c * lda + r];
  B[(size_t)c * ldb + r] = {(float)v.re, (float)v.im};
}

__global__ void k_c2z_add(const c2 *D, int64_t ldd, z2 *X, int64_t ldx, int m, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m || c >= n) return;
  const c2 d = D[(size_t)c * ldd + r];
  z2 &x = X[(size_t)c * ldx + r];
  x.re += (double)d.re;
  x.im += (double)d.im;
}

// Residual partials: P[chunk][j][r] = sum_{c in chunk} M(r, c) X(c, j).
// Grid: (row blocks of 256, column chunks); one thread per row, M streamed
// once, coalesced along rows; X for the chunk staged in LDS.
#define RES_CHUNK 256
#define RES_MAXRHS 8
__global__ __launch_bounds__(256) void k_residual_part(const z2 *__restrict__ M, int64_t ldm, int n,
                                                       const z2 *__restrict__ X, int64_t ldx, int nrhs, z2 *P)
{
  __shared__ z2 xs[RES_MAXRHS][RES_CHUNK];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = blockIdx.y * RES_CHUNK;
  const int nc = min(RES_CHUNK, n - c0);
  for (int idx = threadIdx.x; idx < nrhs * RES_CHUNK; idx += blockDim.x) {
    const int j = idx / RES_CHUNK, c = idx % RES_CHUNK;
    xs[j][c] = c < nc ? X[(size_t)j * ldx + c0 + c] : z2{0, 0};
  }
  __syncthreads();
  if (r >= n) return;
  double ar[RES_MAXRHS], ai[RES_MAXRHS];
#pragma unroll
  for (int j = 0; j < RES_MAXRHS; j++) { ar[j] = 0.0; ai[j] = 0.0; }
  const z2 *col = M + (size_t)c0 * ldm + r;
#pragma unroll 16
  for (int c = 0; c < nc; c++) {
    const z2 m = col[(size_t)c * ldm];
#pragma unroll
    for (int j = 0; j < RES_MAXRHS; j++) {
      if (j < nrhs) {
        const z2 x = xs[j][c];
        ar[j] += m.re * x.re - m.im * x.im;
        ai[j] += m.re * x.im + m.im * x.re;
      }
    }
  }
  for (int j = 0; j < nrhs; j++) P[((size_t)blockIdx.y * nrhs + j) * n + r] = {ar[j], ai[j]};
}

// R(:, j) = B(:, j) - sum_chunks P ; partial norms of R and B per block
__global__ void k_residual_final(const z2 *P, int nchunk, int n, int nrhs, const z2 *B, int64_t ldb, z2 *R,
                                 int64_t ldr, double *norms)
{
  __shared__ double sr[256], sb[256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  double nr = 0.0, nb = 0.0;
  if (r < n) {
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nchunk; k++) {
      const z2 p = P[((size_t)k * nrhs + j) * n + r];
      re += p.re; im += p.im;
    }
    const z2 b = B[(size_t)j * ldb + r];
    const z2 v = {b.re - re, b.im - im};
    R[(size_t)j * ldr + r] = v;
    nr = v.re * v.re + v.im * v.im;
    nb = b.re * b.re + b.im * b.im;
  }
  sr[threadIdx.x] = nr; sb[threadIdx.x] = nb;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) { sr[threadIdx.x] += sr[threadIdx.x + off]; sb[threadIdx.x] += sb[threadIdx.x + off]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    norms[2 * ((size_t)j * gridDim.x + blockIdx.x)] = sr[0];
    norms[2 * ((size_t)j * gridDim.x + blockIdx.x) + 1] = sb[0];
  }
}

}  // namespace

extern "C" int bfa_convert_z2c(const void *d_A64, int m, int n, int64_t lda, void *d_A32, int64_t lda32, void *stream)
{
  if (!d_A64 || !d_A32 || m < 0 || n < 0 || lda < m || lda32 < m) return fail(BFA_ERR_ARG, "bfa_convert_z2c: bad argument");
  if (m == 0 || n == 0) return BFA_OK;
  hipLaunchKernelGGL(k_z2c, dim3((m + 255) / 256, n), dim3(256), 0, (hipStream_t)stream, (const z2 *)d_A64, lda,
                     (c2 *)d_A32, lda32, m, n);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

extern "C" int bfa_lu_factor_c(void *d_A, int n, int64_t lda, int *d_ipiv, int *info, void *stream)
{
  if (!d_A || !d_ipiv || n <= 0 || lda < n) return fail(BFA_ERR_ARG, "bfa_lu_factor_c: bad argument");
  return bfa::lu_factor_c(d_A, n, lda, d_ipiv, info, (hipStream_t)stream, 256);
}

extern "C" int bfa_lu_solve_c(const void *d_A, int n, int64_t lda, const int *d_ipiv, void *d_B, int nrhs, int64_t ldb,
                              void *stream)
{
  if (!d_A || !d_ipiv || !d_B || n <= 0 || nrhs <= 0 || lda < n || ldb < n)
    return fail(BFA_ERR_ARG, "bfa_lu_solve_c: bad argument");
  void *scratch = nullptr;
  HIPCHECK(hipMalloc(&scratch, bfa::lu_solve_scratch_bytes(n, nrhs, 8)));
  int rc = bfa::lu_solve_c(d_A, n, lda, d_ipiv, d_B, nrhs, ldb, (hipStream_t)stream, scratch);
  if (rc == BFA_OK) rc = bfa::lu_solve_check(scratch, n, (hipStream_t)stream);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  (void)hipFree(scratch);
  if (rc) return rc;
  if (e != hipSuccess) return fail(BFA_ERR_HIP, "bfa_lu_solve_c: %s", hipGetErrorString(e));
  return BFA_OK;
}

// Mixed-precision solve with FP64 iterative refinement.  d_B (n x nrhs,
// complex128) holds the right-hand sides on entry and the solution on exit.
extern "C" int bfa_solve_refine(const void *d_M64, int64_t ldm, const void *d_LU32, int64_t ldlu, const int *d_ipiv,
                                int n, void *d_B, int nrhs, int64_t ldb, double tol, int maxit, int *iters,
                                double *resid, void *stream)
{
  if (!d_M64 || !d_LU32 || !d_ipiv || !d_B || n <= 0 || nrhs <= 0 || ldm < n || ldlu < n || ldb < n)
    return fail(BFA_ERR_ARG, "bfa_solve_refine: bad argument");
  hipStream_t st = (hipStream_t)stream;
  if (maxit <= 0) maxit = 10;
  const int nchunk = (n + RES_CHUNK - 1) / RES_CHUNK;
  const int nblk = (n + 255) / 256;
  const size_t zsz = (size_t)n * nrhs * sizeof(z2);
  // workspace: B copy (z), X (z), R (z), D (c), partials, norms, solve scratch
  const size_t part = (size_t)nchunk * std::min(nrhs, RES_MAXRHS) * n * sizeof(z2);
  const size_t norms_b = (size_t)2 * nblk * nrhs * sizeof(double);
  const size_t scr = bfa::lu_solve_scratch_bytes(n, nrhs, 8);
  char *w = nullptr;
  const size_t total = 3 * zsz + (size_t)n * nrhs * sizeof(c2) + part + norms_b + scr + 5 * 256;
  HIPCHECK(hipMalloc((void **)&w, total));
  auto align = [](size_t x) { return (x + 255) / 256 * 256; };
  z2 *Bc = (z2 *)w;
  z2 *X = (z2 *)(w + align(zsz));
  z2 *R = (z2 *)(w + 2 * align(zsz));
  c2 *Dd = (c2 *)(w + 3 * align(zsz));
  z2 *P = (z2 *)(w + 3 * align(zsz) + align((size_t)n * nrhs * sizeof(c2)));
  double *norms = (double *)((char *)P + align(part));
  void *scratch = (char *)norms + align(norms_b);
  int rc = BFA_OK;
  std::vector<double> hn((size_t)2 * nblk * nrhs);
  double worst = 0.0;
  int it = 0;
  const z2 *M = (const z2 *)d_M64;
  hipError_t e;
  if ((e = hipMemcpy2DAsync(Bc, n * sizeof(z2), d_B, ldb * sizeof(z2), n * sizeof(z2), nrhs, hipMemcpyDeviceToDevice,
                            st)) != hipSuccess ||
      (e = hipMemsetAsync(X, 0, zsz, st)) != hipSuccess ||
      (e = hipMemcpyAsync(R, Bc, zsz, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
    (void)hipFree(w);
    return fail(BFA_ERR_HIP, "bfa_solve_refine setup: %s", hipGetErrorString(e));
  }
  rc = bfa::lu_solve_prep_c(d_LU32, n, ldlu, d_ipiv, st, scratch);
  for (it = 0; rc == BFA_OK && it <= maxit; it++) {
    if (it > 0) {
      // r = b - M64 x, in chunks of at most RES_MAXRHS right-hand sides
      for (int j0 = 0; j0 < nrhs; j0 += RES_MAXRHS) {
        const int nj = std::min(RES_MAXRHS, nrhs - j0);
        hipLaunchKernelGGL(k_residual_part, dim3(nblk, nchunk), dim3(256), 0, st, M, ldm, n,
                           (const z2 *)(X + (size_t)j0 * n), (int64_t)n, nj, P);
        hipLaunchKernelGGL(k_residual_final, dim3(nblk, nj), dim3(256), 0, st, (const z2 *)P, nchunk, n, nj,
                           (const z2 *)(Bc + (size_t)j0 * n), (int64_t)n, R + (size_t)j0 * n, (int64_t)n,
                           norms + (size_t)2 * nblk * j0);
      }
      if ((e = hipMemcpyAsync(hn.data(), norms, hn.size() * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess ||
          (e = hipStreamSynchronize(st)) != hipSuccess) {
        rc = fail(BFA_ERR_HIP, "bfa_solve_refine residual: %s", hipGetErrorString(e));
        break;
      }
      worst = 0.0;
      for (int j = 0; j < nrhs; j++) {
        double sr = 0.0, sb = 0.0;
        for (int b = 0; b < nblk; b++) { sr += hn[2 * ((size_t)j * nblk + b)]; sb += hn[2 * ((size_t)j * nblk + b) + 1]; }
        const double rel = sb > 0.0 ? std::sqrt(sr / sb) : std::sqrt(sr);
        worst = std::max(worst, rel);
      }
      if (!(worst > tol) || it == maxit) break;
    }
    // d = LU32^-1 r (complex64), x += d
    hipLaunchKernelGGL(k_z2c, dim3(nblk, nrhs), dim3(256), 0, st, (const z2 *)R, (int64_t)n, Dd, (int64_t)n, n, nrhs);
    rc = bfa::lu_solve_apply_c(d_LU32, n, ldlu, Dd, nrhs, n, st, scratch);
    if (rc) break;
    hipLaunchKernelGGL(k_c2z_add, dim3(nblk, nrhs), dim3(256), 0, st, (const c2 *)Dd, (int64_t)n, X, (int64_t)n, n,
                       nrhs);
  }
  if (rc == BFA_OK) rc = bfa::lu_solve_check(scratch, n, st);
  if (rc == BFA_OK) {
    e = hipMemcpy2DAsync(d_B, ldb * sizeof(z2), X, n * sizeof(z2), n * sizeof(z2), nrhs, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(BFA_ERR_HIP, "bfa_solve_refine copy-out: %s", hipGetErrorString(e));
  }
  (void)hipFree(w);
  if (iters) *iters = it;
  if (resid) *resid = worst;
  return rc;
}
