// bfa_lu.hip -- blocked complex LU with partial pivoting and triangular
// solves on CDNA4, in complex128 and complex64.
//
// Replaces HMatrix::LUFactorize / LUSolve (LAPACK zgetrf/zgetrs) as called
// by the reference driver (src/applications/buff-scatter/buff-scatter.cc:285,
// :314) and by multi-RHS callers (src/applications/buff-neq/GetFlux.cc:265).
//
// Layout: column-major, interleaved complex, leading dimension ld.
// Algorithm: right-looking blocking (outer block NB); the panel is factored
// recursively down to width PW, where a persistent kernel (all workgroups
// co-resident) keeps the panel's rows in LDS and performs one grid barrier
// per column (pivot = first index of max |re|+|im|, the izamax rule).  The
// trailing update C -= A*B runs on MFMA (4 real products per complex
// product, v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <rocblas/rocblas.h>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>

#include "bfa_internal.h"
#include "bfa_lu.h"
#include "bfa_group.h"

using namespace bfa;

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? BFA_ERR_OOM : BFA_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

namespace {

template <typename T> struct cplx { T re, im; };

template <typename T> __device__ __forceinline__ cplx<T> c_mul(cplx<T> a, cplx<T> b)
{ return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <typename T> __device__ __forceinline__ cplx<T> c_sub(cplx<T> a, cplx<T> b)
{ return {a.re - b.re, a.im - b.im}; }
template <typename T> __device__ __forceinline__ cplx<T> c_add(cplx<T> a, cplx<T> b)
{ return {a.re + b.re, a.im + b.im}; }
template <typename T> __device__ __forceinline__ cplx<T> c_inv(cplx<T> a)
{ T d = a.re * a.re + a.im * a.im; return {a.re / d, -a.im / d}; }
template <typename T> __device__ __forceinline__ T c_abs1(cplx<T> a) { return fabs(a.re) + fabs(a.im); }

// ---------------------------------------------------------------------------
// MFMA wrappers: D(16x16) += A(16x4) * B(4x16).  Lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; register j of lane l holds D(row(j, l/16), l%16), where
// the f64 instruction interleaves rows (4j + l/16) -- measured on gfx950 --
// and the f32 one keeps them blocked (4(l/16) + j).
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
template <typename T> struct mf;
template <> struct mf<double> {
  typedef d4 acc;
  __device__ static __forceinline__ acc op(double a, double b, acc c)
  { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static __forceinline__ int row(int j, int lq) { return 4 * j + lq; }
};
template <> struct mf<float> {
  typedef f4 acc;
  __device__ static __forceinline__ acc op(float a, float b, acc c)
  { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static __forceinline__ int row(int j, int lq) { return 4 * lq + j; }
};

// ---------------------------------------------------------------------------
// C(m x n) -= A(m x k) * B(k x n).  Workgroup tile 64 x 64 complex, 4 waves
// in a 2 x 2 layout of 32 x 32 each, K staged through LDS 16 at a time with
// real and imaginary planes stored separately ([k][row] and [k][col]).
#define GBM 64
#define GBN 64
#define GBK 16
// LDS row stride of the GEMM tiles: 16 elements of padding per k row.  An
// MFMA operand read takes k row lane / 16 and row lane % 16, so with
// unpadded rows of 64 or 128 elements (a multiple of 64 banks x 4 bytes)
// the four lane groups address the same banks; the padding moves each
// group to its own banks
#define GPAD 16

template <typename T>
__device__ __forceinline__ void gemm64_body(int m, int n, int k, const cplx<T> *__restrict__ A, int64_t lda,
                                            const cplx<T> *__restrict__ B, int64_t ldb, cplx<T> *C, int64_t ldc,
                                            int row0, int col0, T (&As)[2][GBK][GBM + GPAD], T (&Bs)[2][GBK][GBN + GPAD])
{
  typedef typename mf<T>::acc acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int l16 = lane & 15, lq = lane >> 4;
  // the accumulators collect -A*B from zero and C is added once at the end:
  // one rounding into C per update, not one per k step at C's magnitude
  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int j = 0; j < 4; j++) cr[rb][cb][j] = ci[rb][cb][j] = 0;
  for (int k0 = 0; k0 < k; k0 += GBK) {
    // A tile: 64 rows x 16 k; thread -> row tid%64, k = tid/64 + 4i
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int kk = (tid >> 6) + 4 * i, r = tid & 63;
      cplx<T> v = {0, 0};
      if (row0 + r < m && k0 + kk < k) v = A[(size_t)(k0 + kk) * lda + row0 + r];
      As[0][kk][r] = v.re; As[1][kk][r] = v.im;
    }
    // B tile: 16 k x 64 cols; thread -> k = tid%16, col = tid/16 + 16i
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int kk = tid & 15, c = (tid >> 4) + 16 * i;
      cplx<T> v = {0, 0};
      if (col0 + c < n && k0 + kk < k) v = B[(size_t)(col0 + c) * ldb + k0 + kk];
      Bs[0][kk][c] = v.re; Bs[1][kk][c] = v.im;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int rb = 0; rb < 2; rb++) {
        ar[rb] = -As[0][kk + lq][wr * 32 + rb * 16 + l16];
        ai[rb] = As[1][kk + lq][wr * 32 + rb * 16 + l16];
      }
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        br[cb] = Bs[0][kk + lq][wc * 32 + cb * 16 + l16];
        bi[cb] = Bs[1][kk + lq][wc * 32 + cb * 16 + l16];
      }
      // Cr -= ar*br - ai*bi ; Ci -= ar*bi + ai*br  (ar holds -Re A); two
      // passes so that consecutive MFMAs write different accumulators
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
          cr[rb][cb] = mf<T>::op(ar[rb], br[cb], cr[rb][cb]);
          ci[rb][cb] = mf<T>::op(ar[rb], bi[cb], ci[rb][cb]);
        }
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
          cr[rb][cb] = mf<T>::op(ai[rb], bi[cb], cr[rb][cb]);
          ci[rb][cb] = mf<T>::op(-ai[rb], br[cb], ci[rb][cb]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int col = col0 + wc * 32 + cb * 16 + l16;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int row = row0 + wr * 32 + rb * 16 + mf<T>::row(j, lq);
        if (row < m && col < n) {
          cplx<T> *c = &C[(size_t)col * ldc + row];
          const cplx<T> v = *c;
          *c = {v.re + cr[rb][cb][j], v.im + ci[rb][cb][j]};
        }
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(int m, int n, int k, const cplx<T> *__restrict__ A, int64_t lda,
                                                  const cplx<T> *__restrict__ B, int64_t ldb, cplx<T> *C, int64_t ldc)
{
  __shared__ T As[2][GBK][GBM + GPAD];
  __shared__ T Bs[2][GBK][GBN + GPAD];
  gemm64_body<T>(m, n, k, A, lda, B, ldb, C, ldc, blockIdx.x * GBM, blockIdx.y * GBN, As, Bs);
}

// Large-tile variant: workgroup tile 128 x 128 complex, 8 waves (2 x 4) of
// 64 x 32 each, K staged 16 at a time through double-buffered LDS with the
// next stage prefetched into registers; workgroups are renumbered so that
// each XCD (dispatch round-robin over 8) sweeps a contiguous band of tiles.
#define G2M 128
#define G2N 128
#define G2K 16
// XCD-aware renumbering of a linear workgroup id: the dispatcher deals
// workgroups round-robin over the 8 XCDs, so XCD x gets a contiguous band of
// tiles (bijective when the tile count is a multiple of 8, identity otherwise)
__device__ __forceinline__ int xcd_remap(int bid, int ntile)
{
  return (ntile % 8 == 0) ? (bid % 8) * (ntile / 8) + bid / 8 : bid;
}

// Epilogue of the 128 x 128 tiles: C += (cr, ci), the accumulators holding
// the product transposed (B's value as the MFMA's first operand), so that
// register j of block (rb, cb) is row 16 rb + l16 and column 16 cb +
// row(j, lq) of the wave's sub-tile: 16 lanes write 16 consecutive rows of
// one column.  Each row block's loads are issued before its stores.
template <typename T, int CB>
__device__ __forceinline__ void gemm_epilogue(int m, int n, cplx<T> *C, int64_t ldc, int row, int col,
                                              int lq, const typename mf<T>::acc (&cr)[4][CB],
                                              const typename mf<T>::acc (&ci)[4][CB])
{
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int r = row + rb * 16;
    cplx<T> v[CB][4];
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int c = col + cb * 16 + mf<T>::row(j, lq);
        v[cb][j] = (r < m && c < n) ? C[(size_t)c * ldc + r] : cplx<T>{0, 0};
      }
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int c = col + cb * 16 + mf<T>::row(j, lq);
        if (r < m && c < n) C[(size_t)c * ldc + r] = {v[cb][j].re + cr[rb][cb][j], v[cb][j].im + ci[rb][cb][j]};
      }
  }
}

template <typename T>
__device__ __forceinline__ void gemm128_body(int m, int n, int k, const cplx<T> *__restrict__ A, int64_t lda,
                                             const cplx<T> *__restrict__ B, int64_t ldb, cplx<T> *C, int64_t ldc,
                                             int row0, int col0, T (&As)[2][2][G2K][G2M + GPAD], T (&Bs)[2][2][G2K][G2N + GPAD])
{
  typedef typename mf<T>::acc acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const int l16 = lane & 15, lq = lane >> 4;
  // -A*B accumulated from zero, C added once in the epilogue
  acc_t cr[4][2], ci[4][2];
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
      for (int j = 0; j < 4; j++) cr[rb][cb][j] = ci[rb][cb][j] = 0;
  cplx<T> pa[4], pb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = tid & 127, kk = (tid >> 7) + 4 * i;
      pa[i] = (row0 + r < m && k0 + kk < k) ? A[(size_t)(k0 + kk) * lda + row0 + r] : cplx<T>{0, 0};
      const int kb = tid & 15, c = (tid >> 4) + 32 * i;
      pb[i] = (col0 + c < n && k0 + kb < k) ? B[(size_t)(col0 + c) * ldb + k0 + kb] : cplx<T>{0, 0};
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = tid & 127, kk = (tid >> 7) + 4 * i;
      As[buf][0][kk][r] = pa[i].re; As[buf][1][kk][r] = pa[i].im;
      const int kb = tid & 15, c = (tid >> 4) + 32 * i;
      Bs[buf][0][kb][c] = pb[i].re; Bs[buf][1][kb][c] = pb[i].im;
    }
  };
  const int nst = (k + G2K - 1) / G2K;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int st = 0; st < nst; st++) {
    const int buf = st & 1;
    if (st + 1 < nst) gload((st + 1) * G2K);
#pragma unroll
    for (int kk = 0; kk < G2K; kk += 4) {
      T ar[4], ai[4], br[2], bi[2];
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        ar[rb] = -As[buf][0][kk + lq][wr * 64 + rb * 16 + l16];
        ai[rb] = As[buf][1][kk + lq][wr * 64 + rb * 16 + l16];
      }
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        br[cb] = Bs[buf][0][kk + lq][wc * 32 + cb * 16 + l16];
        bi[cb] = Bs[buf][1][kk + lq][wc * 32 + cb * 16 + l16];
      }
      // two passes so that consecutive MFMAs write different accumulators
      // (same per-accumulator order: Re A Re B first, then Im A Im B)
#pragma unroll
      for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
          cr[rb][cb] = mf<T>::op(br[cb], ar[rb], cr[rb][cb]);
          ci[rb][cb] = mf<T>::op(bi[cb], ar[rb], ci[rb][cb]);
        }
#pragma unroll
      for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
          cr[rb][cb] = mf<T>::op(bi[cb], ai[rb], cr[rb][cb]);
          ci[rb][cb] = mf<T>::op(br[cb], -ai[rb], ci[rb][cb]);
        }
    }
    if (st + 1 < nst) sstore(buf ^ 1);
    __syncthreads();
  }
  gemm_epilogue<T, 2>(m, n, C, ldc, row0 + wr * 64 + l16, col0 + wc * 32, lq, cr, ci);
}

template <typename T>
__global__ __launch_bounds__(512) void k_gemm_sub2(int m, int n, int k, const cplx<T> *__restrict__ A, int64_t lda,
                                                   const cplx<T> *__restrict__ B, int64_t ldb, cplx<T> *C, int64_t ldc,
                                                   int tiles_m, int tiles_n)
{
  __shared__ T As[2][2][G2K][G2M + GPAD];
  __shared__ T Bs[2][2][G2K][G2N + GPAD];
  const int t = xcd_remap(blockIdx.x, tiles_m * tiles_n);
  gemm128_body<T>(m, n, k, A, lda, B, ldb, C, ldc, (t % tiles_m) * G2M, (t / tiles_m) * G2N, As, Bs);
}

// Optional instrumentation: HIP events around every GEMM launch, so a caller
// can read the kernel's total time and algorithmic flop count.
struct GemmProfile {
  bool on = false;
  std::vector<hipEvent_t> pool;
  // [type * 3 + class]: type 0 complex128 / 1 complex64; class 0 the 64x64
  // kernel (k_gemm_sub), 1 the large updates (k_gemm_sub2 or rocBLAS),
  // 2 large updates run as column slices on the LU's panel CUs
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rec[6];
  double flops[6] = {0, 0, 0, 0, 0, 0};
  size_t used = 0;
};
static GemmProfile g_prof;

static void prof_begin(hipStream_t st, hipEvent_t &a)
{
  if (g_prof.used + 2 > g_prof.pool.size()) {
    for (int i = 0; i < 256; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) == hipSuccess) g_prof.pool.push_back(e);
    }
  }
  a = g_prof.pool[g_prof.used++];
  (void)hipEventRecord(a, st);
}

// The large complex128 trailing updates (m, n >= 256) are plain C -= A B
// products and go to rocBLAS's zgemm (alpha = -1, beta = 1), 70 TF/s on these
// shapes against 48 TF/s for k_gemm_sub2<double>: factorisation 517 -> 484 ms
// at N = 19,872.  The complex64 ones stay on k_gemm_sub2<float>: its 64 KB
// workgroups leave room for the look-ahead panel's workgroups on every CU,
// and with rocBLAS's cgemm beside it the panel slows enough that the
// factorisation takes longer (317 against 297 ms).  BFA_GEMM=own or =blas
// selects one kind for both precisions.  Streams being captured into a graph
// always use the hand-written kernels.
// One rocBLAS handle per (device, stream): the CU-partitioned LU runs
// rocBLAS GEMMs on two streams at the same time, and a handle's workspace
// and state belong to one stream.  The table keeps the first 32 pairs; a
// further stream shares the last entry's handle, re-pointed at it.
struct BlasEntry {
  int dev = -1;
  hipStream_t st = nullptr;
  rocblas_handle h = nullptr;
};
static BlasEntry g_blas[32];
static rocblas_handle blas_handle(hipStream_t st, bool want)
{
  static int mode = -1;     // 0: own kernels, 1: where the caller wants rocBLAS, 2: always
  if (mode < 0) {
    const char *e = getenv("BFA_GEMM");
    mode = !e ? 1 : strcmp(e, "own") == 0 ? 0 : strcmp(e, "blas") == 0 ? 2 : 1;
  }
  if (mode == 0 || (mode == 1 && !want)) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const int ne = (int)(sizeof g_blas / sizeof g_blas[0]);
  int i = 0;
  for (; i < ne && g_blas[i].h; i++)
    if (g_blas[i].dev == dev && g_blas[i].st == st) return g_blas[i].h;
  if (i == ne) {            // table full: share the last entry's handle
    BlasEntry &e = g_blas[ne - 1];
    if (e.dev != dev || rocblas_set_stream(e.h, st) != rocblas_status_success) return nullptr;
    e.st = st;
    return e.h;
  }
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
  if (rocblas_set_stream(h, st) != rocblas_status_success) {
    (void)rocblas_destroy_handle(h);
    return nullptr;
  }
  g_blas[i].dev = dev;
  g_blas[i].st = st;
  g_blas[i].h = h;
  return h;
}

static bool blas_gemm_sub(rocblas_handle h, int m, int n, int k, const cplx<double> *A, int64_t lda,
                          const cplx<double> *B, int64_t ldb, cplx<double> *C, int64_t ldc)
{
  const rocblas_double_complex alpha(-1.0, 0.0), beta(1.0, 0.0);
  return rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha,
                       (const rocblas_double_complex *)A, lda, (const rocblas_double_complex *)B, ldb, &beta,
                       (rocblas_double_complex *)C, ldc) == rocblas_status_success;
}
static bool blas_gemm_sub(rocblas_handle h, int m, int n, int k, const cplx<float> *A, int64_t lda,
                          const cplx<float> *B, int64_t ldb, cplx<float> *C, int64_t ldc)
{
  const rocblas_float_complex alpha(-1.0f, 0.0f), beta(1.0f, 0.0f);
  return rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha,
                       (const rocblas_float_complex *)A, lda, (const rocblas_float_complex *)B, ldb, &beta,
                       (rocblas_float_complex *)C, ldc) == rocblas_status_success;
}

template <typename T>
static int gemm_sub(int m, int n, int k, const cplx<T> *A, int64_t lda, const cplx<T> *B, int64_t ldb, cplx<T> *C,
                    int64_t ldc, hipStream_t st, bool lib_ok = false, bool slice = false)
{
  if (m <= 0 || n <= 0 || k <= 0) return BFA_OK;
  dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int p = sizeof(T) == 8 ? 0 : 1;
  if (g_prof.on) prof_begin(st, e0);
  const bool big = m >= 256 && n >= 256;
  rocblas_handle bh = big ? blas_handle(st, sizeof(T) == 8 || lib_ok) : nullptr;
  if (bh && blas_gemm_sub(bh, m, n, k, A, lda, B, ldb, C, ldc)) {
    // done by rocBLAS
  } else if (big) {
    const int tm = (m + G2M - 1) / G2M, tn = (n + G2N - 1) / G2N;
    hipLaunchKernelGGL(k_gemm_sub2<T>, dim3(tm * tn), dim3(512), 0, st, m, n, k, A, lda, B, ldb, C, ldc, tm, tn);
  } else {
    hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C, ldc);
  }
  if (g_prof.on) {
    prof_begin(st, e1);
    const int cls = big ? (slice ? 2 : 1) : 0;
    g_prof.rec[3 * p + cls].push_back({e0, e1});
    g_prof.flops[3 * p + cls] += 8.0 * m * (double)n * k;
  }
  return BFA_OK;
}

// ---------------------------------------------------------------------------
// Diagonal-block solve, kb <= TRSM_KB: B(kb x n) := L^-1 B (L unit lower) or
// U^-1 B (U upper, non-unit).  One workgroup per TRSM_COLS columns of B held
// in LDS.  Larger triangles are blocked (trsm() below): the off-diagonal
// updates run on the MFMA GEMM.
#define TRSM_KB 32
#define TRSM_COLS 64
template <typename T, bool LOWER>
__device__ __forceinline__ void trsm_body(int kb, int n, const cplx<T> *L, int64_t ldl, cplx<T> *B, int64_t ldb,
                                          int c0, cplx<T> (&X)[TRSM_KB][TRSM_COLS], cplx<T> (&Ls)[TRSM_KB][TRSM_KB])
{
  const int tid = threadIdx.x;
  const int nc = min(TRSM_COLS, n - c0);
  for (int idx = tid; idx < kb * TRSM_COLS; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    X[r][c] = c < nc ? B[(size_t)(c0 + c) * ldb + r] : cplx<T>{0, 0};
  }
  for (int idx = tid; idx < kb * kb; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    Ls[r][c] = L[(size_t)c * ldl + r];
  }
  __syncthreads();
  const int c = tid % TRSM_COLS, rg = tid / TRSM_COLS, nrg = blockDim.x / TRSM_COLS;
  if (LOWER) {
    for (int i = 0; i < kb - 1; i++) {
      const cplx<T> xi = X[i][c];
      for (int r = i + 1 + rg; r < kb; r += nrg) X[r][c] = c_sub(X[r][c], c_mul(Ls[r][i], xi));
      __syncthreads();
    }
  } else {
    for (int i = kb - 1; i >= 0; i--) {
      if (rg == 0) X[i][c] = c_mul(X[i][c], c_inv(Ls[i][i]));
      __syncthreads();
      const cplx<T> xi = X[i][c];
      for (int r = rg; r < i; r += nrg) X[r][c] = c_sub(X[r][c], c_mul(Ls[r][i], xi));
      __syncthreads();
    }
  }
  for (int idx = tid; idx < kb * TRSM_COLS; idx += blockDim.x) {
    const int r = idx % kb, cc = idx / kb;
    if (cc < nc) B[(size_t)(c0 + cc) * ldb + r] = X[r][cc];
  }
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsm(int kb, int n, const cplx<T> *L, int64_t ldl, cplx<T> *B, int64_t ldb)
{
  __shared__ cplx<T> X[TRSM_KB][TRSM_COLS];
  __shared__ cplx<T> Ls[TRSM_KB][TRSM_KB];
  trsm_body<T, LOWER>(kb, n, L, ldl, B, ldb, blockIdx.x * TRSM_COLS, X, Ls);
}

template <typename T, bool LOWER>
static void trsm_diag(int kb, int n, const cplx<T> *L, int64_t ldl, cplx<T> *B, int64_t ldb, hipStream_t st)
{
  if (kb <= 0 || n <= 0) return;
  hipLaunchKernelGGL((k_trsm<T, LOWER>), dim3((n + TRSM_COLS - 1) / TRSM_COLS), dim3(256), 0, st, kb, n, L, ldl, B,
                     ldb);
}

template <typename T> static int gemm_sub(int m, int n, int k, const cplx<T> *A, int64_t lda, const cplx<T> *B,
                                          int64_t ldb, cplx<T> *C, int64_t ldc, hipStream_t st, bool lib_ok,
                                          bool slice);

// B(kb x n) := L^-1 B or U^-1 B for any kb, blocked by TRSM_KB
template <typename T, bool LOWER>
static void trsm(int kb, int n, const cplx<T> *L, int64_t ldl, cplx<T> *B, int64_t ldb, hipStream_t st)
{
  if (kb <= 0 || n <= 0) return;
  if (LOWER) {
    for (int i0 = 0; i0 < kb; i0 += TRSM_KB) {
      const int ib = std::min(TRSM_KB, kb - i0);
      trsm_diag<T, true>(ib, n, L + (size_t)i0 * ldl + i0, ldl, B + i0, ldb, st);
      if (i0 + ib < kb)
        gemm_sub<T>(kb - i0 - ib, n, ib, L + (size_t)i0 * ldl + i0 + ib, ldl, B + i0, ldb, B + i0 + ib, ldb, st);
    }
  } else {
    const int nb = (kb + TRSM_KB - 1) / TRSM_KB;
    for (int bi = nb - 1; bi >= 0; bi--) {
      const int i0 = bi * TRSM_KB, ib = std::min(TRSM_KB, kb - i0);
      trsm_diag<T, false>(ib, n, L + (size_t)i0 * ldl + i0, ldl, B + i0, ldb, st);
      if (i0 > 0) gemm_sub<T>(i0, n, ib, L + (size_t)i0 * ldl, ldl, B + i0, ldb, B, ldb, st);
    }
  }
}

// B(kb x n) := L^-1 B, L unit lower, kb <= TF_MAXK, in one launch.  A
// workgroup owns a strip of TF_SC columns of B (column-major in LDS, 32 KB
// so that it fits beside the trailing-update GEMM's tiles on a CU); per
// 32-row block, the diagonal solve runs in registers (a wave takes two
// columns, lane = row, the solved value passed along by shuffles, no
// workgroup barrier), then every row below subtracts the block's
// contribution from register accumulators.
#define TF_MAXK 256
template <typename T> struct TfShape { static constexpr int SC = 32768 / (TF_MAXK * (int)sizeof(cplx<T>)); };

template <typename T, int SC = TfShape<T>::SC>
__device__ __forceinline__ void trsm_lf_body(int kb, int n, const cplx<T> *__restrict__ L, int64_t ldl, cplx<T> *B,
                                             int64_t ldb, int c0, cplx<T> (&X)[SC][TF_MAXK])
{
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nc = min(SC, n - c0);
  for (int idx = tid; idx < kb * SC; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    X[c][r] = c < nc ? B[(size_t)(c0 + c) * ldb + r] : cplx<T>{0, 0};
  }
  __syncthreads();
  const int row = lane & 31, half = lane >> 5;
  for (int b0 = 0; b0 < kb; b0 += 32) {
    const int bk = min(32, kb - b0);
    cplx<T> lrow[32];
#pragma unroll
    for (int k = 0; k < 32; k++)
      lrow[k] = (k < row && row < bk) ? L[(size_t)(b0 + k) * ldl + b0 + row] : cplx<T>{0, 0};
    for (int pr = wave; pr < SC / 2; pr += 4) {
      const int c = 2 * pr + half;
      cplx<T> x = row < bk ? X[c][b0 + row] : cplx<T>{0, 0};
#pragma unroll
      for (int k = 0; k < 31; k++) {
        const cplx<T> xk = {__shfl(x.re, k + 32 * half, 64), __shfl(x.im, k + 32 * half, 64)};
        x = c_sub(x, c_mul(lrow[k], xk));   // lrow[k] = 0 unless k < row
      }
      if (row < bk) X[c][b0 + row] = x;
    }
    __syncthreads();
    for (int r = b0 + bk + tid; r < kb; r += blockDim.x) {
      cplx<T> acc[SC];
#pragma unroll
      for (int c = 0; c < SC; c++) acc[c] = cplx<T>{0, 0};
      if (bk == 32) {
        // a full block: the row's 32 factor entries are loaded together
        // before use, so their latencies overlap instead of adding up
        cplx<T> fr[32];
#pragma unroll
        for (int k = 0; k < 32; k++) fr[k] = L[(size_t)(b0 + k) * ldl + r];
#pragma unroll
        for (int k = 0; k < 32; k++)
#pragma unroll
          for (int c = 0; c < SC; c++) acc[c] = c_add(acc[c], c_mul(fr[k], X[c][b0 + k]));
      } else {
        for (int k = 0; k < bk; k++) {
          const cplx<T> l = L[(size_t)(b0 + k) * ldl + r];
#pragma unroll
          for (int c = 0; c < SC; c++) acc[c] = c_add(acc[c], c_mul(l, X[c][b0 + k]));
        }
      }
#pragma unroll
      for (int c = 0; c < SC; c++) X[c][r] = c_sub(X[c][r], acc[c]);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < kb * SC; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    if (c < nc) B[(size_t)(c0 + c) * ldb + r] = X[c][r];
  }
}

template <typename T, int SC = TfShape<T>::SC>
__global__ __launch_bounds__(256) void k_trsm_lf(int kb, int n, const cplx<T> *__restrict__ L, int64_t ldl, cplx<T> *B,
                                                 int64_t ldb)
{
  __shared__ cplx<T> X[SC][TF_MAXK];
  trsm_lf_body<T, SC>(kb, n, L, ldl, B, ldb, blockIdx.x * SC, X);
}

// B(kb x n) := U^-1 B, U upper with its diagonal, kb <= TF_MAXK: the mirror
// of trsm_lf_body (32-row blocks from the bottom; in the diagonal block lane
// k divides by U(k,k) when its value is final, then passes it on)
template <typename T, int SC = TfShape<T>::SC>
__global__ __launch_bounds__(256) void k_trsm_uf(int kb, int n, const cplx<T> *__restrict__ U, int64_t ldu, cplx<T> *B,
                                                 int64_t ldb)
{
  __shared__ cplx<T> X[SC][TF_MAXK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c0 = blockIdx.x * SC, nc = min(SC, n - c0);
  for (int idx = tid; idx < kb * SC; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    X[c][r] = c < nc ? B[(size_t)(c0 + c) * ldb + r] : cplx<T>{0, 0};
  }
  __syncthreads();
  const int row = lane & 31, half = lane >> 5;
  for (int b0 = ((kb - 1) / 32) * 32; b0 >= 0; b0 -= 32) {
    const int bk = min(32, kb - b0);
    cplx<T> lrow[32];
#pragma unroll
    for (int k = 0; k < 32; k++)
      lrow[k] = (k > row && k < bk) ? U[(size_t)(b0 + k) * ldu + b0 + row] : cplx<T>{0, 0};
    const cplx<T> dinv = row < bk ? c_inv(U[(size_t)(b0 + row) * ldu + b0 + row]) : cplx<T>{0, 0};
    for (int pr = wave; pr < SC / 2; pr += 4) {
      const int c = 2 * pr + half;
      cplx<T> x = row < bk ? X[c][b0 + row] : cplx<T>{0, 0};
#pragma unroll
      for (int k = 31; k >= 0; k--) {
        if (row == k) x = c_mul(x, dinv);
        const cplx<T> xk = {__shfl(x.re, k + 32 * half, 64), __shfl(x.im, k + 32 * half, 64)};
        if (row < k) x = c_sub(x, c_mul(lrow[k], xk));
      }
      if (row < bk) X[c][b0 + row] = x;
    }
    __syncthreads();
    for (int r = tid; r < b0; r += blockDim.x) {
      cplx<T> acc[SC];
#pragma unroll
      for (int c = 0; c < SC; c++) acc[c] = cplx<T>{0, 0};
      if (bk == 32) {
        // a full block: the row's 32 factor entries are loaded together
        // before use, so their latencies overlap instead of adding up
        cplx<T> fr[32];
#pragma unroll
        for (int k = 0; k < 32; k++) fr[k] = U[(size_t)(b0 + k) * ldu + r];
#pragma unroll
        for (int k = 0; k < 32; k++)
#pragma unroll
          for (int c = 0; c < SC; c++) acc[c] = c_add(acc[c], c_mul(fr[k], X[c][b0 + k]));
      } else {
        for (int k = 0; k < bk; k++) {
          const cplx<T> u = U[(size_t)(b0 + k) * ldu + r];
#pragma unroll
          for (int c = 0; c < SC; c++) acc[c] = c_add(acc[c], c_mul(u, X[c][b0 + k]));
        }
      }
#pragma unroll
      for (int c = 0; c < SC; c++) X[c][r] = c_sub(X[c][r], acc[c]);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < kb * SC; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    if (c < nc) B[(size_t)(c0 + c) * ldb + r] = X[c][r];
  }
}

// B(kb x n) := L^-1 B, L unit lower: one launch up to TF_MAXK rows
template <typename T>
static void trsm_lower(int kb, int n, const cplx<T> *L, int64_t ldl, cplx<T> *B, int64_t ldb, hipStream_t st)
{
  if (kb <= 0 || n <= 0) return;
  if (kb > TF_MAXK) { trsm<T, true>(kb, n, L, ldl, B, ldb, st); return; }
  constexpr int SC = TfShape<T>::SC;
  hipLaunchKernelGGL(k_trsm_lf<T>, dim3((n + SC - 1) / SC), dim3(256), 0, st, kb, n, L, ldl, B, ldb);
}

// ---------------------------------------------------------------------------
// Row interchanges rows k0..k0+kb-1 <-> ipiv[.] (absolute, 0-based) on the
// columns [c0, c1) of A; one thread per column, swaps applied in order.
template <typename T>
__device__ __forceinline__ void laswp_col(cplx<T> *A, int64_t lda, int c, int k0, int kb, const int *ipiv)
{
  cplx<T> *col = A + (size_t)c * lda;
  for (int i = 0; i < kb; i++) {
    const int p = ipiv[k0 + i];
    if (p != k0 + i) {
      cplx<T> t = col[k0 + i];
      col[k0 + i] = col[p];
      col[p] = t;
    }
  }
}

template <typename T>
__global__ void k_laswp(cplx<T> *A, int64_t lda, int c0, int c1, int k0, int kb, const int *ipiv)
{
  const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (c < c1) laswp_col<T>(A, lda, c, k0, kb, ipiv);
}

template <typename T>
static void laswp(cplx<T> *A, int64_t lda, int c0, int c1, int k0, int kb, const int *ipiv, hipStream_t st)
{
  if (c1 <= c0 || kb <= 0) return;
  hipLaunchKernelGGL(k_laswp<T>, dim3((c1 - c0 + 127) / 128), dim3(128), 0, st, A, lda, c0, c1, k0, kb, ipiv);
}

// ---------------------------------------------------------------------------
// Persistent panel factorisation of an m x w panel (w <= PW).  Workgroup g
// owns rows [g*rpw, (g+1)*rpw) in LDS.  Per column: local argmax, publish
// (value, row, row data) and, by the owner, the current row j; grid barrier;
// every workgroup resolves the same pivot, swaps in LDS, scales and updates.
#define PW 32
#define SPIN_LIMIT (1u << 24)

typedef bfa::PanelSlots PanelWork;

// Cross-workgroup data of the panel (slots, candidate rows, row j) is
// written and read only with agent-scope atomic stores/loads, which are
// coherent across the XCDs without cache maintenance; all of it is written
// by wave 0 of each workgroup.  Grid barrier: thread 0 waits until its
// wave's stores have completed, arrives on the counter and polls it, so no
// L2 write-back or invalidate is needed per column.
template <typename V> __device__ __forceinline__ void st_agent(V *p, V v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V> __device__ __forceinline__ V ld_agent(const V *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The panel's barrier counters and pivot keys are spread over PB_LANES
// cache lines (workgroup g uses lane g % PB_LANES, which follows the XCD
// round-robin), so arrivals on one address are ~nwg / 8 instead of nwg.
#define PB_LANES 8
#define PB_BAR_STRIDE 32    // u32 between counters (128 bytes)
#define PB_WORDS (PB_LANES * PB_BAR_STRIDE + PB_LANES * PW * 2)   // u32: counters, then keys [lane][PW]

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)v, off, 64), hi = __shfl_xor((unsigned)(v >> 32), off, 64);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    v = o > v ? o : v;
  }
  return v;
}

// Wave 0 arrives (thread 0, after its wave's stores completed) and polls the
// sum of the counters (lanes < PB_LANES); every wave waits at the workgroup
// barrier that follows.
__device__ __forceinline__ void grid_barrier(unsigned *bar, int g, unsigned target, int *flags)
{
  __syncthreads();
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    if (lane == 0) {
      __builtin_amdgcn_s_waitcnt(0);   // this wave's slot stores and key update have completed
      __hip_atomic_fetch_add(&bar[(g % PB_LANES) * PB_BAR_STRIDE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    for (;;) {
      unsigned c = lane < PB_LANES ? ld_agent(&bar[lane * PB_BAR_STRIDE]) : 0u;
#pragma unroll
      for (int off = 4; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
      if (__shfl(c, 0, 64) >= target) break;
      // a barrier that already timed out elsewhere ends this one at once
      if ((++spins & 1023) == 0 && __shfl(lane == 0 ? ld_agent(&flags[0]) : 0, 0, 64)) break;
      if (spins > SPIN_LIMIT) {
        if (lane == 0) atomicOr(&flags[0], 1);
        break;
      }
    }
  }
  __syncthreads();
}

template <typename T> struct PanelShared {
  cplx<T> prow[PW], rj[PW];
  unsigned long long wkey[4];
  int wrow[4];
};

// first index of the maximum: (v, r) wins over (ov, orr) unless ov is larger
// or equal with a smaller row
template <typename T> __device__ __forceinline__ void argmax_merge(T &v, int &r, int &w, T ov, int orr, int ow)
{
  if (ov > v || (ov == v && orr < r)) { v = ov; r = orr; w = ow; }
}

// Pivot search key: |x|_1 rounded to float (monotone, non-negative: its
// bits order as integers) above the inverted row, so the 64-bit maximum is
// the largest magnitude and, among equal keys, the first row (LAPACK's
// izamax up to the float rounding of near-ties).
__device__ __forceinline__ unsigned long long panel_key(float v, int row)
{
  return ((unsigned long long)__float_as_uint(v) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)row);
}

// Body shared by the single-matrix and the grouped panel kernels: workgroup
// g of the nwg that factor this panel; rows = the workgroup's LDS block.
// Per column: the workgroup's candidate (found while updating the previous
// column) is reduced across waves, its row data published and its key
// merged into the column's atomic maximum, one grid barrier, every
// workgroup reads the same pivot from the key (the winner's rows belong to
// workgroup p / rpw), then each thread swaps/scales/updates its own rows and
// tracks the next column's candidate.
template <typename T>
__device__ __forceinline__ void panel_body(cplx<T> *A, int m, int w, int64_t lda, int *ipiv, int rpw,
                                           const PanelWork &pw, int piv_off, int g, int nwg, cplx<T> *rows,
                                           PanelShared<T> &sh)
{
#define RW(r, c) rows[(size_t)(c) * rpw + (r)]
  // The panel is latency-bound and runs beside the trailing GEMM: its waves
  // take instruction-issue priority over the GEMM's waves on shared SIMDs.
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int r0 = g * rpw, r1 = min(m, r0 + rpw), nr = max(0, r1 - r0);
  cplx<T> *rdata = (cplx<T> *)pw.rowdata;
  cplx<T> *rowjg = (cplx<T> *)pw.rowj;
  for (int idx = tid; idx < nr * w; idx += blockDim.x) {
    const int r = idx % nr, c = idx / nr;
    RW(r, c) = A[(size_t)c * lda + r0 + r];
  }
  __syncthreads();
  // this thread's candidate for column 0
  T best = -1;
  int brow = INT_MAX;
  for (int r = tid; r < nr; r += blockDim.x) {
    int dummy = 0;
    argmax_merge(best, brow, dummy, c_abs1(RW(r, 0)), r0 + r, 0);
  }
  for (int j = 0; j < w; j++) {
    const int buf = j & 1;
    {
      const unsigned long long k = wave_max_u64(brow == INT_MAX ? 0ull : panel_key((float)best, brow));
      if (lane == 0) sh.wkey[wave] = k;
    }
    __syncthreads();
    unsigned long long lkey = sh.wkey[0];
    for (int q = 1; q < nwave; q++) lkey = sh.wkey[q] > lkey ? sh.wkey[q] : lkey;
    const int lrow = lkey ? (int)(0xFFFFFFFFu - (unsigned)lkey) : INT_MAX;
    // (w <= PW = 32: every store below is made by wave 0)
    if (tid == 0 && lkey)
      __hip_atomic_fetch_max(&pw.key[(g % PB_LANES) * PW + j], lkey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lrow != INT_MAX && tid < w) {
      const cplx<T> v = RW(lrow - r0, tid);
      cplx<T> *d = &rdata[((size_t)buf * nwg + g) * PW + tid];
      st_agent(&d->re, v.re);
      st_agent(&d->im, v.im);
    }
    if (j >= r0 && j < r1 && tid < w) {
      const cplx<T> v = RW(j - r0, tid);
      st_agent(&rowjg[buf * PW + tid].re, v.re);
      st_agent(&rowjg[buf * PW + tid].im, v.im);
    }
    grid_barrier(pw.bar, g, (unsigned)((j + 1) * nwg), pw.flags);
    // global pivot (same in every workgroup): the largest key
    if (tid < 64) {
      const unsigned long long k = wave_max_u64(tid < PB_LANES ? ld_agent(&pw.key[tid * PW + j]) : 0ull);
      if (tid == 0) sh.wrow[0] = (int)(0xFFFFFFFFu - (unsigned)k);
    }
    __syncthreads();
    const int p = sh.wrow[0], win = p / rpw;
    if (tid < PW) {
      const cplx<T> *d = &rdata[((size_t)buf * nwg + win) * PW + tid];
      sh.prow[tid] = tid < w ? cplx<T>{ld_agent(&d->re), ld_agent(&d->im)} : cplx<T>{0, 0};
      sh.rj[tid] = tid < w ? cplx<T>{ld_agent(&rowjg[buf * PW + tid].re), ld_agent(&rowjg[buf * PW + tid].im)}
                           : cplx<T>{0, 0};
    }
    __syncthreads();
    const bool scale = c_abs1(sh.prow[j]) != (T)0;
    if (g == 0 && tid == 0) {
      ipiv[j] = p + piv_off;
      if (!scale && pw.flags[1] == 0) pw.flags[1] = piv_off + j + 1;
    }
    // swap rows j and p, scale column j below the diagonal, update the rest;
    // each row belongs to one thread, which also tracks column j+1's candidate
    best = -1;
    brow = INT_MAX;
    const cplx<T> inv = scale ? c_inv(sh.prow[j]) : cplx<T>{0, 0};
    for (int r = tid; r < nr; r += blockDim.x) {
      const int gr = r0 + r;
      if (gr < j) continue;
      if (gr == j) {
        if (p != j)
          for (int c = 0; c < w; c++) RW(r, c) = sh.prow[c];
        continue;
      }
      if (gr == p)
        for (int c = 0; c < w; c++) RW(r, c) = sh.rj[c];
      if (scale) {
        const cplx<T> l = c_mul(RW(r, j), inv);
        RW(r, j) = l;
        for (int c = j + 1; c < w; c++) RW(r, c) = c_sub(RW(r, c), c_mul(l, sh.prow[c]));
      }
      if (j + 1 < w) {
        int dummy = 0;
        argmax_merge(best, brow, dummy, c_abs1(RW(r, j + 1)), gr, 0);
      }
    }
  }
  __syncthreads();
  for (int idx = tid; idx < nr * w; idx += blockDim.x) {
    const int r = idx % nr, c = idx / nr;
    A[(size_t)c * lda + r0 + r] = RW(r, c);
  }
#undef RW
}

// The single-matrix panels alternate between two sets of barrier counters
// and keys: launch i uses one (cleared by launch i - 1, which completed
// before it on the same stream) and its workgroup 0 clears the other for
// launch i + 1, in place of a separate clearing launch per panel.
template <typename T>
__global__ __launch_bounds__(256) void k_panel(cplx<T> *A, int m, int w, int64_t lda, int *ipiv, int rpw, PanelWork pw,
                                               int piv_off, unsigned *clear)
{
  extern __shared__ unsigned char smem_raw[];
  __shared__ PanelShared<T> sh;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < PB_WORDS; i += blockDim.x) st_agent(&clear[i], 0u);
  panel_body<T>(A, m, w, lda, ipiv, rpw, pw, piv_off, blockIdx.x, gridDim.x, (cplx<T> *)smem_raw, sh);
}

__global__ void k_zero_u32(unsigned *p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_add_offset(int *ipiv, int n, int off)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ipiv[i] += off;
}

// ---------------------------------------------------------------------------
// Grouped kernels: one launch, many independent problems (descriptors in
// device memory, sorted by their first workgroup blk0).
template <typename P>
__device__ __forceinline__ int find_prob(const P *__restrict__ d, int np, int bid)
{
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].blk0 <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
__global__ __launch_bounds__(256) void kg_gemm64(const GemmProb *__restrict__ d, int np)
{
  __shared__ T As[2][GBK][GBM + GPAD];
  __shared__ T Bs[2][GBK][GBN + GPAD];
  const GemmProb q = d[find_prob(d, np, blockIdx.x)];
  const int t = blockIdx.x - q.blk0;
  gemm64_body<T>(q.m, q.n, q.k, (const cplx<T> *)q.A, q.lda, (const cplx<T> *)q.B, q.ldb, (cplx<T> *)q.C, q.ldc,
                 (t % q.tiles_m) * GBM, (t / q.tiles_m) * GBN, As, Bs);
}

template <typename T>
__global__ __launch_bounds__(512) void kg_gemm128(const GemmProb *__restrict__ d, int np)
{
  __shared__ T As[2][2][G2K][G2M + GPAD];
  __shared__ T Bs[2][2][G2K][G2N + GPAD];
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const GemmProb q = d[find_prob(d, np, bid)];
  const int t = bid - q.blk0;
  gemm128_body<T>(q.m, q.n, q.k, (const cplx<T> *)q.A, q.lda, (const cplx<T> *)q.B, q.ldb, (cplx<T> *)q.C, q.ldc,
                  (t % q.tiles_m) * G2M, (t / q.tiles_m) * G2N, As, Bs);
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void kg_trsm(const TrsmProb *__restrict__ d, int np)
{
  __shared__ cplx<T> X[TRSM_KB][TRSM_COLS];
  __shared__ cplx<T> Ls[TRSM_KB][TRSM_KB];
  const TrsmProb q = d[find_prob(d, np, blockIdx.x)];
  trsm_body<T, LOWER>(q.kb, q.n, (const cplx<T> *)q.L, q.ldl, (cplx<T> *)q.B, q.ldb,
                      (blockIdx.x - q.blk0) * TRSM_COLS, X, Ls);
}

template <typename T>
__global__ __launch_bounds__(256) void kg_trsm_lf(const TrsmProb *__restrict__ d, int np)
{
  __shared__ cplx<T> X[TfShape<T>::SC][TF_MAXK];
  const TrsmProb q = d[find_prob(d, np, blockIdx.x)];
  trsm_lf_body<T>(q.kb, q.n, (const cplx<T> *)q.L, q.ldl, (cplx<T> *)q.B, q.ldb,
                  (blockIdx.x - q.blk0) * TfShape<T>::SC, X);
}

template <typename T>
__global__ __launch_bounds__(128) void kg_swap(const SwapProb *__restrict__ d, int np)
{
  const SwapProb q = d[find_prob(d, np, blockIdx.x)];
  const int c = q.c0 + (blockIdx.x - q.blk0) * 128 + threadIdx.x;
  if (c < q.c1) laswp_col<T>((cplx<T> *)q.A, q.lda, c, q.k0, q.kb, q.ipiv);
}

template <typename T>
__global__ __launch_bounds__(256) void kg_panel(const PanelProb *__restrict__ d, int np)
{
  extern __shared__ unsigned char smem_raw[];
  __shared__ PanelShared<T> sh;
  const PanelProb q = d[find_prob(d, np, blockIdx.x)];
  panel_body<T>((cplx<T> *)q.A, q.m, q.w, q.lda, q.ipiv, q.rpw, q.slots, q.piv_off, blockIdx.x - q.blk0, q.nwg,
                (cplx<T> *)smem_raw, sh);
}


}  // namespace

// ---------------------------------------------------------------------------
namespace bfa {

struct LUWork {
  int device = -1, maxwg = 0;
  PanelWork pw{};
  void *mem = nullptr;
  int num_cu = 0;
  unsigned *bars = nullptr;      // two sets of panel barrier counters + keys
  unsigned long long npanel = 0; // panel launches so far (selects the set)
  hipStream_t side = nullptr;    // bulk trailing updates of the look-ahead LU
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // CU-partitioned look-ahead: the panel on mask_cus CUs, the updates on the rest
  int mask_cus = -1;             // -1 undecided, 0 off
  bool mask_off = false;         // bfa_lu_partition(0): shared-CU path
  hipStream_t s_pan = nullptr, s_bulk = nullptr;
  hipEvent_t evm[6] = {};
};
static LUWork g_lu[16];

static int lu_work(LUWork *&w)
{
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 16) return fail(BFA_ERR_ARG, "device index %d unsupported", dev);
  w = &g_lu[dev];
  if (w->mem) return BFA_OK;
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, dev));
  w->num_cu = prop.multiProcessorCount;
  w->maxwg = 1024;
  const size_t bytes = 2 * (size_t)w->maxwg * PW * 16 + 2 * PW * 16 + 2 * PB_WORDS * 4 + 64;
  HIPCHECK(hipMalloc(&w->mem, bytes));
  HIPCHECK(hipMemset(w->mem, 0, bytes));
  char *p = (char *)w->mem;
  w->pw.rowdata = p; p += 2 * (size_t)w->maxwg * PW * 16;
  w->pw.rowj = p; p += 2 * PW * 16;
  w->bars = (unsigned *)p;                       // two sets of PB_WORDS, both zero
  w->pw.bar = w->bars;
  w->pw.key = (unsigned long long *)(w->pw.bar + PB_LANES * PB_BAR_STRIDE);
  p += 2 * PB_WORDS * 4;
  w->pw.flags = (int *)p;
  w->device = dev;
  return BFA_OK;
}

template <typename T>
static int panel_base(cplx<T> *A, int m, int w, int64_t lda, int *ipiv, int piv_off, LUWork *lw, hipStream_t st,
                      int cus)
{
  // cus < num_cu: the panel's stream owns cus CUs (CU-partitioned
  // look-ahead) and nothing else runs there, so its workgroups take the
  // largest LDS blocks (fewest workgroups, at most one per CU).
  // complex128: as few workgroups as the LDS allows (fewer arrivals per
  // grid barrier), up to ~150 KB of rows each, at most one per CU.
  // complex64: the look-ahead runs this panel beside the trailing update,
  // whose two 64 KB tiles per CU leave ~32 KB of LDS; workgroups of at most
  // 31 KB of rows start beside them at once instead of waiting for CUs to
  // drain (measured at N = 19,872: 376 ms factorisation against 470 ms).
  static int attr_done[2] = {0, 0};
  const int ti = sizeof(T) == 8 ? 0 : 1;
  size_t lds_cap = 150 * 1024;
  if (!attr_done[ti]) {
    attr_done[ti] = hipFuncSetAttribute((const void *)k_panel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds_cap) == hipSuccess ? 1 : -1;
  }
  if (attr_done[ti] < 0) lds_cap = 64 * 1024;
  const bool masked = cus < lw->num_cu;
  if (masked) {
    static int attr_m[2] = {0, 0};
    if (!attr_m[ti])
      attr_m[ti] = hipFuncSetAttribute((const void *)k_panel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       158 * 1024) == hipSuccess ? 1 : -1;
    if (attr_m[ti] < 0) return fail(BFA_ERR_HIP, "panel LDS attribute for the CU-partitioned LU");
    lds_cap = 158 * 1024;
  }
  const int row_bytes = w * (int)sizeof(cplx<T>);
  int rpw = std::max(16, (int)(lds_cap / row_bytes));
  int nwg = (m + rpw - 1) / rpw;
  if (nwg > cus) { nwg = cus; rpw = (m + nwg - 1) / nwg; }
  // masked: one workgroup per panel CU (at least 16 rows each), so the
  // per-column row updates spread over every panel CU's LDS
  if (masked) { nwg = std::min(cus, std::max(1, (m + 15) / 16)); rpw = (m + nwg - 1) / nwg; }
  if (sizeof(T) == 4 && !masked) {
    const int rco = std::max(16, (31 * 1024) / row_bytes);
    if ((m + rco - 1) / rco <= lw->num_cu) { rpw = rco; nwg = (m + rpw - 1) / rpw; }
  }
  rpw = (m + nwg - 1) / nwg;
  nwg = (m + rpw - 1) / rpw;
  if ((size_t)rpw * row_bytes > lds_cap || nwg > lw->maxwg)
    return fail(BFA_ERR_ARG, "panel of %d rows too tall for LDS-resident factorisation", m);
  size_t lds = (size_t)rpw * w * sizeof(cplx<T>);
  // The grid barrier needs every workgroup resident at once: the grid never
  // exceeds what the occupancy calculator says fits on the device, and the
  // barrier's bounded spin reports (flags[0]) rather than hangs otherwise.
  static int per_cu[2][2] = {{-1, -1}, {-1, -1}};
  int &pc = per_cu[masked ? 1 : 0][ti];
  if (pc < 0) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_panel<T>, 256, lds_cap) != hipSuccess) per = 0;
    pc = per;
  }
  if (nwg > pc * cus)
    return fail(BFA_ERR_HIP, "panel needs %d co-resident workgroups, %d CUs fit %d", nwg, cus, pc * cus);
  const int set = (int)(lw->npanel++ & 1);
  PanelWork pw = lw->pw;
  pw.bar = lw->bars + set * PB_WORDS;
  pw.key = (unsigned long long *)(pw.bar + PB_LANES * PB_BAR_STRIDE);
  hipLaunchKernelGGL(k_panel<T>, dim3(nwg), dim3(256), lds, st, A, m, w, lda, ipiv, rpw, pw, piv_off,
                     lw->bars + (set ^ 1) * PB_WORDS);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    // the launch that would have cleared the other set did not run
    (void)hipMemset(lw->bars, 0, 2 * PB_WORDS * sizeof(unsigned));
    return fail(BFA_ERR_HIP, "panel launch: %s", hipGetErrorString(e));
  }
  return BFA_OK;
}

// recursive panel: m x w, pivots written absolute (offset piv_off + local)
template <typename T>
static int panel_rec(cplx<T> *A, int m, int w, int64_t lda, int *ipiv, int piv_off, LUWork *lw, hipStream_t st,
                     int cus)
{
  if (w <= PW) return panel_base<T>(A, m, w, lda, ipiv, piv_off, lw, st, cus);
  const int w1 = w / 2, w2 = w - w1;
  int rc = panel_rec<T>(A, m, w1, lda, ipiv, piv_off, lw, st, cus);
  if (rc) return rc;
  // apply left pivots to the right half (rows relative to panel top)
  cplx<T> *Ar = A + (size_t)w1 * lda;
  hipLaunchKernelGGL(k_laswp<T>, dim3((w2 + 127) / 128), dim3(128), 0, st, Ar - (ptrdiff_t)piv_off, lda, 0, w2, piv_off,
                     w1, (const int *)(ipiv - piv_off));
  trsm_lower<T>(w1, w2, A, lda, Ar, lda, st);
  gemm_sub<T>(m - w1, w2, w1, A + w1, lda, Ar, lda, Ar + w1, lda, st);
  rc = panel_rec<T>(A + (size_t)w1 * lda + w1, m - w1, w2, lda, ipiv + w1, piv_off + w1, lw, st, cus);
  if (rc) return rc;
  // apply right pivots to the left half
  hipLaunchKernelGGL(k_laswp<T>, dim3((w1 + 127) / 128), dim3(128), 0, st, A - (ptrdiff_t)piv_off, lda, 0, w1,
                     piv_off + w1, w2, (const int *)(ipiv - piv_off));
  return BFA_OK;
}

// CU-partitioned look-ahead.  The panel's stream owns P CUs (BFA_LU_MASK,
// default 64; 0: off) and the update stream the other 256 - P, through
// per-queue CU masks.  Beside an unmasked GEMM the latency-bound panel shares
// SIMDs with it and runs ~3x slower than alone (5.2 against 1.8 ms per 256
// columns at N = 19,872); on its own CUs it keeps its speed while the bulk
// update runs on rocBLAS's GEMM over the rest: factorisation at N = 19,872
// 297 -> 276 ms (complex64), 496 -> 465 ms (complex128).  A matrix whose
// first panel does not fit the panel CUs' LDS takes the shared-CU path.
static int lu_mask_setup(LUWork *lw)
{
  if (lw->mask_cus >= 0) return lw->mask_cus;
  lw->mask_cus = 0;
  const char *e = getenv("BFA_LU_MASK");
  const int P = e ? atoi(e) : 64;
  // 8 XCDs of 32 CUs.  The driver deals mask bits to the XCDs round-robin
  // (bit i -> XCD i % 8; measured on MI355X by the XCC_ID of workgroups run
  // under a mask with one extra bit), so bits [0, P) give each XCD P/8 CUs;
  // a queue with no CU on some XCD could never run that XCD's workgroups.
  // Workgroups are also dealt round-robin over each XCD's shader engines, so
  // P is a multiple of 32 (P/32 panel CUs per engine): with 6 CUs per XCD
  // some panel workgroups found no free CU and the grid barrier timed out.
  if (P <= 0 || P > 128 || P % 32 != 0 || lw->num_cu != 256) return 0;
  uint32_t mp[8] = {}, mb[8];
  for (int b = 0; b < P; b++) mp[b / 32] |= 1u << (b % 32);
  for (int i = 0; i < 8; i++) mb[i] = ~mp[i];
  const uint32_t words = 8;
  if (hipExtStreamCreateWithCUMask(&lw->s_pan, words, mp) != hipSuccess ||
      hipExtStreamCreateWithCUMask(&lw->s_bulk, words, mb) != hipSuccess)
    return 0;
  for (auto &ev : lw->evm)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return 0;
  lw->mask_cus = P;
  return P;
}

// Columns of the bulk update that the panel stream takes after its panel,
// so that both streams finish a step together: panel time t_p(m) plus X
// columns at the panel CUs' GEMM rate equals the other columns at the update
// CUs' rate.  Rates: rocBLAS GEMM on 192 CUs at ~100 (complex64) and ~53
// (complex128) TF/s, a third of that on 64; panel recursion of 256 columns
// ~1.5 ms + 65 ns per row (kernel trace, N = 19,872).  0 when the panel is
// the longer stream; a multiple of 128, and the next block column stays with
// the update stream.
template <typename T> static int bulk_split(int m, int rest, int jb, int nb_next, int P, int ncu)
{
  const double rate_b = (sizeof(T) == 4 ? 100e12 : 53e12) * (ncu - P) / 192.0, rate_p = rate_b * P / (ncu - P);
  const double c = 8.0 * m * jb, t_pan = 1.5e-3 + 65e-9 * m;
  const double X = (rest * c / rate_b - t_pan) / (c / rate_p + c / rate_b);
  int x = X > 0 ? (int)(X / 128) * 128 : 0;
  x = std::min(x, rest - nb_next);
  return x >= 256 ? x : 0;
}

template <typename T>
static int lu_factor_masked(cplx<T> *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb, LUWork *lw)
{
  const int P = lw->mask_cus;
  hipStream_t sp = lw->s_pan, sb = lw->s_bulk;
  hipEvent_t ev_s = lw->evm[0], ev_p = lw->evm[1], ev_n = lw->evm[2], ev_e = lw->evm[3], ev_t = lw->evm[4],
             ev_g = lw->evm[5];
  HIPCHECK(hipMemsetAsync(lw->pw.flags, 0, 8, st));
  HIPCHECK(hipEventRecord(ev_s, st));
  HIPCHECK(hipStreamWaitEvent(sp, ev_s, 0));
  HIPCHECK(hipStreamWaitEvent(sb, ev_s, 0));
  int rc = panel_rec<T>(A, n, std::min(nb, n), lda, ipiv, 0, lw, sp, P);
  if (rc) return rc;
  bool sliced = false;                           // the panel stream updated a column slice last step
  for (int j = 0; j < n; j += nb) {
    const int jb = std::min(nb, n - j);
    cplx<T> *Ajj = A + (size_t)j * lda + j;
    const int next = j + jb;
    HIPCHECK(hipEventRecord(ev_p, sp));          // panel j done
    HIPCHECK(hipStreamWaitEvent(sb, ev_p, 0));
    laswp<T>(A, lda, 0, j, j, jb, ipiv, sb);    // rows of panel j, columns left of it
    if (next >= n) break;
    const int nb2 = std::min(nb, n - next);
    cplx<T> *Anext = A + (size_t)next * lda;
    // block column next on the update CUs, then its panel on the panel CUs
    laswp<T>(A, lda, next, next + nb2, j, jb, ipiv, sb);
    trsm_lower<T>(jb, nb2, Ajj, lda, Anext + j, lda, sb);
    gemm_sub<T>(n - next, nb2, jb, Ajj + jb, lda, Anext + j, lda, Anext + next, lda, sb, true);
    HIPCHECK(hipEventRecord(ev_n, sb));
    HIPCHECK(hipStreamWaitEvent(sp, ev_n, 0));
    const int rest = n - next - nb2;
    int x = 0;
    if (rest > 0) {
      cplx<T> *Arest = A + (size_t)(next + nb2) * lda;
      // last step's slice, updated on the panel stream, is swapped again here
      if (sliced) HIPCHECK(hipStreamWaitEvent(sb, ev_g, 0));
      laswp<T>(A, lda, next + nb2, n, j, jb, ipiv, sb);
      trsm_lower<T>(jb, rest, Ajj, lda, Arest + j, lda, sb);
      x = bulk_split<T>(n - next, rest, jb, std::min(nb, rest), P, lw->num_cu);
      if (x > 0) HIPCHECK(hipEventRecord(ev_t, sb));
      gemm_sub<T>(n - next, rest - x, jb, Ajj + jb, lda, Arest + j, lda, Arest + next, lda, sb, true);
    }
    rc = panel_rec<T>(Anext + next, n - next, nb2, lda, ipiv + next, next, lw, sp, P);
    if (rc) return rc;
    if (x > 0) {
      // the right-most x columns of the bulk update, on the panel CUs
      cplx<T> *As = A + (size_t)(n - x) * lda;
      HIPCHECK(hipStreamWaitEvent(sp, ev_t, 0));
      gemm_sub<T>(n - next, x, jb, Ajj + jb, lda, As + j, lda, As + next, lda, sp, true, true);
      HIPCHECK(hipEventRecord(ev_g, sp));
    }
    sliced = x > 0;
  }
  HIPCHECK(hipEventRecord(ev_e, sb));
  HIPCHECK(hipStreamWaitEvent(st, ev_e, 0));
  HIPCHECK(hipEventRecord(ev_e, sp));
  HIPCHECK(hipStreamWaitEvent(st, ev_e, 0));
  HIPCHECK(hipGetLastError());
  int flags[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(flags, lw->pw.flags, 8, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (flags[0]) return fail(BFA_ERR_HIP, "LU panel barrier timed out");
  if (info) *info = flags[1];
  return BFA_OK;
}

// whether every panel of an n x n factorisation fits the panel CUs at once
// (the first, tallest panel decides): one workgroup of <= 158 KB per CU,
// dealt round-robin to the XCDs, each of which holds P/8 panel CUs
template <typename T> static bool masked_fits(int n, int nb, const LUWork *lw)
{
  const size_t row_bytes = (size_t)std::min(nb, PW) * sizeof(cplx<T>);
  const size_t rpw = ((size_t)n + lw->mask_cus - 1) / lw->mask_cus;
  return rpw * row_bytes <= 158 * 1024;
}

// Blocked right-looking LU with one block of look-ahead.  Stream st carries
// the critical path: the update of the next block column, then its panel.
// A second stream carries the bulk trailing update of the remaining columns
// (row swaps, TRSM, GEMM) and the swaps left of each panel, so each panel
// factorisation overlaps the previous step's big GEMM.  Column ranges
// touched concurrently are disjoint: the panel owns its nb columns, the bulk
// update the columns to its right; the left swaps of panel j+1 are queued
// behind bulk update j on the same stream because both touch panel j's L.
template <typename T>
int lu_factor(cplx<T> *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb)
{
  LUWork *lw = nullptr;
  int rc = lu_work(lw);
  if (rc) return rc;
  if (!lw->mask_off && lu_mask_setup(lw) > 0 && masked_fits<T>(n, nb, lw))
    return lu_factor_masked<T>(A, n, lda, ipiv, info, st, nb, lw);
  if (!lw->side) {
    HIPCHECK(hipStreamCreateWithFlags(&lw->side, hipStreamNonBlocking));
    for (auto &e : lw->ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  hipStream_t sb = lw->side;
  hipEvent_t ev_p = lw->ev[0], ev_b = lw->ev[1], ev_s = lw->ev[2];
  HIPCHECK(hipMemsetAsync(lw->pw.flags, 0, 8, st));
  HIPCHECK(hipEventRecord(ev_s, st));
  HIPCHECK(hipStreamWaitEvent(sb, ev_s, 0));
  rc = panel_rec<T>(A, n, std::min(nb, n), lda, ipiv, 0, lw, st, lw->num_cu);
  if (rc) return rc;
  for (int j = 0; j < n; j += nb) {
    const int jb = std::min(nb, n - j);
    cplx<T> *Ajj = A + (size_t)j * lda + j;
    const int next = j + jb;
    HIPCHECK(hipEventRecord(ev_p, st));          // panel j done
    HIPCHECK(hipStreamWaitEvent(sb, ev_p, 0));
    laswp<T>(A, lda, 0, j, j, jb, ipiv, sb);    // rows of panel j, columns left of it
    if (next >= n) break;
    const int nb2 = std::min(nb, n - next);
    cplx<T> *Anext = A + (size_t)next * lda;
    // critical path: block column next, then its panel
    laswp<T>(A, lda, next, next + nb2, j, jb, ipiv, st);
    trsm_lower<T>(jb, nb2, Ajj, lda, Anext + j, lda, st);
    gemm_sub<T>(n - next, nb2, jb, Ajj + jb, lda, Anext + j, lda, Anext + next, lda, st);
    // bulk: the columns right of block column next
    const int rest = n - next - nb2;
    if (rest > 0) {
      cplx<T> *Arest = A + (size_t)(next + nb2) * lda;
      laswp<T>(A, lda, next + nb2, n, j, jb, ipiv, sb);
      trsm_lower<T>(jb, rest, Ajj, lda, Arest + j, lda, sb);
      gemm_sub<T>(n - next, rest, jb, Ajj + jb, lda, Arest + j, lda, Arest + next, lda, sb);
    }
    HIPCHECK(hipEventRecord(ev_b, sb));
    rc = panel_rec<T>(Anext + next, n - next, nb2, lda, ipiv + next, next, lw, st, lw->num_cu);
    if (rc) return rc;
    // the next block column was updated by this bulk pass
    HIPCHECK(hipStreamWaitEvent(st, ev_b, 0));
  }
  HIPCHECK(hipEventRecord(ev_b, sb));
  HIPCHECK(hipStreamWaitEvent(st, ev_b, 0));
  HIPCHECK(hipGetLastError());
  int flags[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(flags, lw->pw.flags, 8, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (flags[0]) return fail(BFA_ERR_HIP, "LU panel barrier timed out");
  if (info) *info = flags[1];
  return BFA_OK;
}

// ---------------------------------------------------------------------------
// Solve: permutation from ipiv (composed on the host), row gather, blocked
// forward (unit L) and backward (U) substitution.
template <typename T>
__global__ void k_gather_rows(const cplx<T> *B, int64_t ldb, cplx<T> *X, int64_t ldx, int n, int nrhs, const int *perm)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i < n && j < nrhs) X[(size_t)j * ldx + i] = B[(size_t)j * ldb + perm[i]];
}

// triangular solves: SV_KB-row blocks, up to SV_RHS right-hand sides per
// sweep; from SV_BLOCKED_RHS right-hand sides on, the blocked GEMM form
#define SV_KB 64
#define SV_RHS 8
#define SV_BLOCKED_RHS 64

// ---------------------------------------------------------------------------
// Streaming block substitution.  The diagonal SV_KB x SV_KB blocks of L
// (unit) and U are inverted once per factorisation (k_trsv_inv); one launch
// per triangle then solves every row block: a workgroup takes a ticket t
// (row block t of L, or nblk-1-t of U), accumulates the updates of the
// blocks solved before it as they appear, applies its inverted diagonal
// block and stores its solution.  The output starts filled with a
// signalling-NaN tag that arithmetic never produces; solved values are
// stored and polled with agent-scope atomics (coherent across the XCDs, no
// cache maintenance, no separate flag), so a block is visible as soon as
// its values land.  A ticket only ever waits on lower tickets, which belong
// to workgroups already running.
template <typename T> struct InvShared {
  cplx<T> Ls[SV_KB][SV_KB + 1];
  cplx<T> X[SV_KB][SV_KB + 1];
};
template <typename T>
__device__ __forceinline__ void trsv_inv_body(int n, const cplx<T> *__restrict__ A, int64_t lda, cplx<T> *inv, int bi,
                                              bool lower, InvShared<T> &sh)
{
  cplx<T>(&Ls)[SV_KB][SV_KB + 1] = sh.Ls;
  cplx<T>(&X)[SV_KB][SV_KB + 1] = sh.X;
  const int nblk = (n + SV_KB - 1) / SV_KB;
  const int k0 = bi * SV_KB, kb = min(SV_KB, n - k0);
  const int tid = threadIdx.x, c = tid % SV_KB, rg = tid / SV_KB;
  for (int idx = tid; idx < SV_KB * SV_KB; idx += blockDim.x) {
    const int r = idx % SV_KB, cc = idx / SV_KB;
    Ls[r][cc] = (r < kb && cc < kb) ? A[(size_t)(k0 + cc) * lda + k0 + r] : cplx<T>{0, 0};
    X[r][cc] = {(T)(r == cc && r < kb ? 1 : 0), 0};
  }
  __syncthreads();
  if (lower) {
    for (int i = 0; i < kb - 1; i++) {
      const cplx<T> xi = X[i][c];
      for (int r = i + 1 + rg; r < kb; r += 4) X[r][c] = c_sub(X[r][c], c_mul(Ls[r][i], xi));
      __syncthreads();
    }
  } else {
    for (int i = kb - 1; i >= 0; i--) {
      if (rg == 0) X[i][c] = c_mul(X[i][c], c_inv(Ls[i][i]));
      __syncthreads();
      const cplx<T> xi = X[i][c];
      for (int r = rg; r < i; r += 4) X[r][c] = c_sub(X[r][c], c_mul(Ls[r][i], xi));
      __syncthreads();
    }
  }
  cplx<T> *o = inv + ((size_t)(lower ? 0 : nblk) + bi) * SV_KB * SV_KB;
  for (int idx = tid; idx < SV_KB * SV_KB; idx += blockDim.x) o[idx] = X[idx % SV_KB][idx / SV_KB];
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsv_inv(int n, const cplx<T> *__restrict__ A, int64_t lda, cplx<T> *inv)
{
  __shared__ InvShared<T> sh;
  trsv_inv_body<T>(n, A, lda, inv, blockIdx.x, blockIdx.y == 0, sh);
}

template <typename T> struct SvTag;
template <> struct SvTag<float> {
  typedef unsigned long long W;   // re and im stored together
  static constexpr unsigned bits = 0x7FA5A5A5u;
};
template <> struct SvTag<double> {
  typedef unsigned long long W;
  static constexpr unsigned long long bits = 0x7FF4A5A5A5A5A5A5ull;
};

// agent-scope store / load of one complex value with tag detection
__device__ __forceinline__ void sv_store(cplx<float> *p, cplx<float> v)
{
  const unsigned long long w =
      (unsigned long long)__float_as_uint(v.re) | ((unsigned long long)__float_as_uint(v.im) << 32);
  st_agent((unsigned long long *)p, w);
}
__device__ __forceinline__ void sv_store(cplx<double> *p, cplx<double> v)
{
  st_agent((unsigned long long *)p, (unsigned long long)__double_as_longlong(v.re));
  st_agent((unsigned long long *)p + 1, (unsigned long long)__double_as_longlong(v.im));
}
__device__ __forceinline__ bool sv_load(const cplx<float> *p, cplx<float> &v)
{
  const unsigned long long w = ld_agent((const unsigned long long *)p);
  const unsigned lo = (unsigned)w, hi = (unsigned)(w >> 32);
  v = {__uint_as_float(lo), __uint_as_float(hi)};
  return lo != SvTag<float>::bits && hi != SvTag<float>::bits;
}
__device__ __forceinline__ bool sv_load(const cplx<double> *p, cplx<double> &v)
{
  const unsigned long long a = ld_agent((const unsigned long long *)p);
  const unsigned long long b = ld_agent((const unsigned long long *)p + 1);
  v = {__longlong_as_double((long long)a), __longlong_as_double((long long)b)};
  return a != SvTag<double>::bits && b != SvTag<double>::bits;
}

// fill two n x nrhs outputs with the tag, clear the tickets
template <typename T>
__global__ void k_sv_reset(cplx<T> *Y, cplx<T> *X, size_t count, unsigned *tickets, int nt)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    if (sizeof(T) == 4) {
      const unsigned long long w = (unsigned long long)SvTag<float>::bits * 0x100000001ull;
      ((unsigned long long *)Y)[i] = w;
      ((unsigned long long *)X)[i] = w;
    } else {
      ((unsigned long long *)Y)[2 * i] = ((unsigned long long *)Y)[2 * i + 1] = (unsigned long long)SvTag<double>::bits;
      ((unsigned long long *)X)[2 * i] = ((unsigned long long *)X)[2 * i + 1] = (unsigned long long)SvTag<double>::bits;
    }
  }
  if (i < (size_t)nt) tickets[i] = 0u;
}

template <typename T> struct StreamShared {
  cplx<T> xs[SV_RHS][SV_KB];
  cplx<T> part[4][SV_RHS][SV_KB];
  int ticket;
};

// Bin: right-hand sides (n x nrhs, read-only here); Xo: tag-filled output.
// ticket: this sweep's counter; flag: set when a wait gives up.
// x_i = D_i^-1 (b_i - sum over k != p of L_ik x_k) - (D_i^-1 L_ip) x_p, p the
// block solved just before i: M = D_i^-1 L_ip is formed while the chain is
// still upstream, everything but the M x_p term is finished before x_p
// lands, and the step that waits on x_p is one 64 x 64 product between two
// workgroup barriers.
template <typename T, bool LOWER>
__device__ __forceinline__ void trsv_stream_body(int n, int nrhs, const cplx<T> *__restrict__ A, int64_t lda,
                                                 const cplx<T> *__restrict__ inv, const cplx<T> *Bin, cplx<T> *Xo,
                                                 int64_t ldx, unsigned *ticket, unsigned *flag, StreamShared<T> &sh)
{
  const int nblk = (n + SV_KB - 1) / SV_KB;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) sh.ticket = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int t = sh.ticket;
  const int bi = LOWER ? t : nblk - 1 - t;
  const int r0 = bi * SV_KB, kb = min(SV_KB, n - r0);
  const int r = r0 + lane;
  // this wave's columns of the inverted diagonal block, loaded up front
  const cplx<T> *Di = inv + ((size_t)(LOWER ? 0 : nblk) + bi) * SV_KB * SV_KB;
  cplx<T> dcol[SV_KB / 4];
#pragma unroll
  for (int q = 0; q < SV_KB / 4; q++) dcol[q] = Di[(size_t)(wv + 4 * q) * SV_KB + lane];
  // M = D^-1 L_ip, this thread's row (lane) and the wave's columns: row tt of
  // block (i, p) is held by lane tt of the wave, passed along by shuffles
  const int kp = t > 0 ? (LOWER ? t - 1 : nblk - t) : 0;
  const int cp0 = kp * SV_KB, cbp = t > 0 ? min(SV_KB, n - cp0) : 0;
  cplx<T> mcol[SV_KB / 4];
#pragma unroll
  for (int q = 0; q < SV_KB / 4; q++) mcol[q] = {0, 0};
  if (t > 0) {
    cplx<T> lp[SV_KB / 4];
#pragma unroll
    for (int q = 0; q < SV_KB / 4; q++) {
      const int c = wv + 4 * q;
      lp[q] = (lane < kb && c < cbp) ? A[(size_t)(cp0 + c) * lda + r] : cplx<T>{0, 0};
    }
    for (int tt = 0; tt < SV_KB; tt++) {
      const cplx<T> d = Di[(size_t)tt * SV_KB + lane];      // D^-1 (lane, tt)
#pragma unroll
      for (int q = 0; q < SV_KB / 4; q++) {
        const T pr = __shfl(lp[q].re, tt, 64), pi = __shfl(lp[q].im, tt, 64);
        mcol[q].re += d.re * pr - d.im * pi;
        mcol[q].im += d.re * pi + d.im * pr;
      }
    }
  }
  T ar[SV_RHS], ai[SV_RHS];
#pragma unroll
  for (int j = 0; j < SV_RHS; j++) { ar[j] = 0; ai[j] = 0; }
  for (int s = 0; s < t - 1; s++) {
    const int k = LOWER ? s : nblk - 1 - s;
    const int c0 = k * SV_KB, cb = min(SV_KB, n - c0);
    // this wave's columns of block (bi, k), loaded before waiting on x_k
    cplx<T> l[SV_KB / 4];
#pragma unroll
    for (int q = 0; q < SV_KB / 4; q++) {
      const int c = wv + 4 * q;
      l[q] = (lane < kb && c < cb) ? A[(size_t)(c0 + c) * lda + r] : cplx<T>{0, 0};
    }
    // wave 0 polls x_k until every value has landed
    if (wv == 0) {
      unsigned spins = 0;
      for (int idx = lane; idx < cb * nrhs; idx += 64) {
        const int c = idx % cb, j = idx / cb;
        cplx<T> v;
        while (!sv_load(&Xo[(size_t)j * ldx + c0 + c], v)) {
          if (++spins > SPIN_LIMIT) { atomicOr(flag, 1u); break; }
        }
        sh.xs[j][c] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SV_KB / 4; q++) {
      const int c = wv + 4 * q;
#pragma unroll
      for (int j = 0; j < SV_RHS; j++) {
        if (j < nrhs) {
          const cplx<T> x = sh.xs[j][c < cb ? c : 0];
          ar[j] += l[q].re * x.re - l[q].im * x.im;
          ai[j] += l[q].re * x.im + l[q].im * x.re;
        }
      }
    }
    __syncthreads();
  }
  for (int j = 0; j < nrhs; j++) sh.part[wv][j][lane] = {ar[j], ai[j]};
  __syncthreads();
  // raw right-hand side of this block without x_p's term: b - sum of updates
  if (wv == 0) {
    for (int j = 0; j < nrhs; j++) {
      cplx<T> v = {0, 0};
      if (lane < kb) {
        const cplx<T> b = Bin[(size_t)j * ldx + r];
        v.re = b.re - (sh.part[0][j][lane].re + sh.part[1][j][lane].re + sh.part[2][j][lane].re + sh.part[3][j][lane].re);
        v.im = b.im - (sh.part[0][j][lane].im + sh.part[1][j][lane].im + sh.part[2][j][lane].im + sh.part[3][j][lane].im);
      }
      sh.xs[j][lane] = v;
    }
  }
  __syncthreads();
  // z = inv(D) * raw, D the diagonal block
#pragma unroll
  for (int j = 0; j < SV_RHS; j++) { ar[j] = 0; ai[j] = 0; }
#pragma unroll
  for (int q = 0; q < SV_KB / 4; q++) {
    const int c = wv + 4 * q;
    const cplx<T> d = dcol[q];
#pragma unroll
    for (int j = 0; j < SV_RHS; j++) {
      if (j < nrhs) {
        const cplx<T> x = sh.xs[j][c];
        ar[j] += d.re * x.re - d.im * x.im;
        ai[j] += d.re * x.im + d.im * x.re;
      }
    }
  }
  for (int j = 0; j < nrhs; j++) sh.part[wv][j][lane] = {ar[j], ai[j]};
  __syncthreads();
  T zr[SV_RHS], zi[SV_RHS];
#pragma unroll
  for (int j = 0; j < SV_RHS; j++) {
    zr[j] = zi[j] = 0;
    if (wv == 0 && j < nrhs) {
      zr[j] = sh.part[0][j][lane].re + sh.part[1][j][lane].re + sh.part[2][j][lane].re + sh.part[3][j][lane].re;
      zi[j] = sh.part[0][j][lane].im + sh.part[1][j][lane].im + sh.part[2][j][lane].im + sh.part[3][j][lane].im;
    }
  }
  if (t > 0) {
    // wave 0 polls x_p (xs is free: every wave read it before the last barrier)
    if (wv == 0) {
      unsigned spins = 0;
      for (int idx = lane; idx < cbp * nrhs; idx += 64) {
        const int c = idx % cbp, j = idx / cbp;
        cplx<T> v;
        while (!sv_load(&Xo[(size_t)j * ldx + cp0 + c], v)) {
          if (++spins > SPIN_LIMIT) { atomicOr(flag, 1u); break; }
        }
        sh.xs[j][c] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SV_RHS; j++) { ar[j] = 0; ai[j] = 0; }
#pragma unroll
    for (int q = 0; q < SV_KB / 4; q++) {
      const int c = wv + 4 * q;
#pragma unroll
      for (int j = 0; j < SV_RHS; j++) {
        if (j < nrhs) {
          const cplx<T> x = sh.xs[j][c < cbp ? c : 0];
          ar[j] += mcol[q].re * x.re - mcol[q].im * x.im;
          ai[j] += mcol[q].re * x.im + mcol[q].im * x.re;
        }
      }
    }
    for (int j = 0; j < nrhs; j++) sh.part[wv][j][lane] = {ar[j], ai[j]};
    __syncthreads();
    if (wv == 0 && lane < kb) {
      for (int j = 0; j < nrhs; j++) {
        const T mr = sh.part[0][j][lane].re + sh.part[1][j][lane].re + sh.part[2][j][lane].re + sh.part[3][j][lane].re;
        const T mi = sh.part[0][j][lane].im + sh.part[1][j][lane].im + sh.part[2][j][lane].im + sh.part[3][j][lane].im;
        sv_store(&Xo[(size_t)j * ldx + r], cplx<T>{zr[j] - mr, zi[j] - mi});
      }
    }
  } else if (wv == 0 && lane < kb) {
    for (int j = 0; j < nrhs; j++) sv_store(&Xo[(size_t)j * ldx + r], cplx<T>{zr[j], zi[j]});
  }
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_stream(int n, int nrhs, const cplx<T> *__restrict__ A, int64_t lda,
                                                     const cplx<T> *__restrict__ inv, const cplx<T> *Bin,
                                                     cplx<T> *Xo, int64_t ldx, unsigned *ticket, unsigned *flag)
{
  __shared__ StreamShared<T> sh;
  trsv_stream_body<T, LOWER>(n, nrhs, A, lda, inv, Bin, Xo, ldx, ticket, flag, sh);
}

// grouped forms over the problems of a batch (complex64)
__global__ __launch_bounds__(256) void kg_sv_inv(const SvProb *__restrict__ d, int np)
{
  __shared__ InvShared<float> sh;
  const SvProb q = d[find_prob(d, np, blockIdx.x)];
  trsv_inv_body<float>(q.n, (const cplx<float> *)q.LU, q.ld, (cplx<float> *)q.inv, blockIdx.x - q.blk0,
                       blockIdx.y == 0, sh);
}

template <bool LOWER>
__global__ __launch_bounds__(256) void kg_sv_stream(const SvProb *__restrict__ d, int np)
{
  __shared__ StreamShared<float> sh;
  const SvProb q = d[find_prob(d, np, blockIdx.x)];
  if (q.skip && *q.skip) return;
  trsv_stream_body<float, LOWER>(q.n, q.nrhs, (const cplx<float> *)q.LU, q.ld, (const cplx<float> *)q.inv,
                                 (const cplx<float> *)(LOWER ? q.Bin : q.Y), (cplx<float> *)(LOWER ? q.Y : q.X), q.ldx,
                                 q.tick + (LOWER ? 0 : 1), (unsigned *)q.flag, sh);
}

// scratch: [perm n ints][flag 256 B][inverted diagonal blocks]
//          [P b, Y, X: n x nrhs each][tickets]
static size_t sv_align(size_t x) { return (x + 255) / 256 * 256; }
static size_t sv_inv_off(int n) { return sv_align((size_t)n * 4) + 256; }
static size_t sv_x_off(int n, size_t elem)
{
  const size_t nblk = (size_t)(n + SV_KB - 1) / SV_KB;
  return sv_inv_off(n) + sv_align(2 * nblk * SV_KB * SV_KB * elem);
}
static int sv_nticket(int nrhs) { return 2 * ((nrhs + SV_RHS - 1) / SV_RHS); }

// Row permutation from the pivots and the inverted diagonal blocks: once
// per factorisation.
template <typename T>
int lu_solve_prep(const cplx<T> *A, int n, int64_t lda, const int *ipiv, hipStream_t st, void *scratch)
{
  int *perm = (int *)scratch;
  {
    std::vector<int> hp(n), pm(n);
    HIPCHECK(hipMemcpyAsync(hp.data(), ipiv, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++) pm[i] = i;
    for (int i = 0; i < n; i++) {
      const int p = hp[i];
      if (p < 0 || p >= n) return fail(BFA_ERR_ARG, "invalid pivot %d at %d", p, i);
      if (p != i) std::swap(pm[i], pm[p]);
    }
    HIPCHECK(hipMemcpyAsync(perm, pm.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  }
  const int nblk = (n + SV_KB - 1) / SV_KB;
  cplx<T> *inv = (cplx<T> *)((char *)scratch + sv_inv_off(n));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, (unsigned *)((char *)scratch + sv_align((size_t)n * 4)), 1);
  hipLaunchKernelGGL(k_trsv_inv<T>, dim3(nblk, 2), dim3(256), 0, st, n, A, lda, inv);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

// Many right-hand sides (LUSolve(HMatrix*), GetFlux.cc:258-265): blocked by
// TF_MAXK rows, one-launch TRSMs on the diagonal blocks and the MFMA GEMM for
// the updates, so the work runs at GEMM speed (8 n^2 nrhs flop per sweep)
template <typename T>
static int lu_solve_blocked(const cplx<T> *A, int n, int64_t lda, cplx<T> *X, int nrhs, hipStream_t st)
{
  constexpr int SC = TfShape<T>::SC;
  for (int k = 0; k < n; k += TF_MAXK) {
    const int kb = std::min(TF_MAXK, n - k);
    trsm_lower<T>(kb, nrhs, A + (size_t)k * lda + k, lda, X + k, n, st);
    if (k + kb < n) {
      int rc = gemm_sub<T>(n - k - kb, nrhs, kb, A + (size_t)k * lda + k + kb, lda, X + k, n, X + k + kb, n, st);
      if (rc) return rc;
    }
  }
  for (int k = ((n - 1) / TF_MAXK) * TF_MAXK; k >= 0; k -= TF_MAXK) {
    const int kb = std::min(TF_MAXK, n - k);
    hipLaunchKernelGGL(k_trsm_uf<T>, dim3((nrhs + SC - 1) / SC), dim3(256), 0, st, kb, nrhs,
                       A + (size_t)k * lda + k, lda, X + k, (int64_t)n);
    if (k > 0) {
      int rc = gemm_sub<T>(k, nrhs, kb, A + (size_t)k * lda, lda, X + k, n, X, n, st);
      if (rc) return rc;
    }
  }
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

// B <- (LU)^-1 P B with the data lu_solve_prep left in scratch
template <typename T>
int lu_solve_apply(const cplx<T> *A, int n, int64_t lda, cplx<T> *B, int nrhs, int64_t ldb, hipStream_t st,
                   void *scratch)
{
  const int *perm = (const int *)scratch;
  unsigned *flag = (unsigned *)((char *)scratch + sv_align((size_t)n * 4));
  const cplx<T> *inv = (const cplx<T> *)((char *)scratch + sv_inv_off(n));
  const size_t vec = sv_align((size_t)n * nrhs * sizeof(cplx<T>));
  cplx<T> *Pb = (cplx<T> *)((char *)scratch + sv_x_off(n, sizeof(cplx<T>)));
  hipLaunchKernelGGL(k_gather_rows<T>, dim3((n + 255) / 256, nrhs), dim3(256), 0, st, B, ldb, Pb, (int64_t)n, n, nrhs,
                     perm);
  if (nrhs >= SV_BLOCKED_RHS) {
    int rc = lu_solve_blocked<T>(A, n, lda, Pb, nrhs, st);
    if (rc) return rc;
    HIPCHECK(hipMemcpy2DAsync(B, ldb * sizeof(cplx<T>), Pb, n * sizeof(cplx<T>), n * sizeof(cplx<T>), nrhs,
                              hipMemcpyDeviceToDevice, st));
    return BFA_OK;
  }
  cplx<T> *Y = (cplx<T> *)((char *)Pb + vec);
  cplx<T> *X = (cplx<T> *)((char *)Y + vec);
  unsigned *tickets = (unsigned *)((char *)X + vec);
  const int nt = sv_nticket(nrhs);
  const size_t count = (size_t)n * nrhs;
  hipLaunchKernelGGL(k_sv_reset<T>, dim3((unsigned)((std::max(count, (size_t)nt) + 255) / 256)), dim3(256), 0, st, Y,
                     X, count, tickets, nt);
  const int nblk = (n + SV_KB - 1) / SV_KB;
  for (int j0 = 0, c = 0; j0 < nrhs; j0 += SV_RHS, c++) {
    const int nj = std::min(SV_RHS, nrhs - j0);
    const size_t o = (size_t)j0 * n;
    hipLaunchKernelGGL((k_trsv_stream<T, true>), dim3(nblk), dim3(256), 0, st, n, nj, A, lda, inv,
                       (const cplx<T> *)(Pb + o), Y + o, (int64_t)n, tickets + 2 * c, flag);
    hipLaunchKernelGGL((k_trsv_stream<T, false>), dim3(nblk), dim3(256), 0, st, n, nj, A, lda, inv,
                       (const cplx<T> *)(Y + o), X + o, (int64_t)n, tickets + 2 * c + 1, flag);
  }
  HIPCHECK(hipMemcpy2DAsync(B, ldb * sizeof(cplx<T>), X, n * sizeof(cplx<T>), n * sizeof(cplx<T>), nrhs,
                            hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

template <typename T>
int lu_solve(const cplx<T> *A, int n, int64_t lda, const int *ipiv, cplx<T> *B, int nrhs, int64_t ldb, hipStream_t st,
             void *scratch)
{
  int rc = lu_solve_prep<T>(A, n, lda, ipiv, st, scratch);
  if (rc == BFA_OK) rc = lu_solve_apply<T>(A, n, lda, B, nrhs, ldb, st, scratch);
  return rc;
}

size_t lu_solve_scratch_bytes(int n, int nrhs, size_t elem)
{
  // the blocked form needs the permuted right-hand sides only
  if (nrhs >= SV_BLOCKED_RHS) return sv_x_off(n, elem) + sv_align((size_t)n * nrhs * elem);
  return sv_x_off(n, elem) + 3 * sv_align((size_t)n * nrhs * elem) + sv_align((size_t)sv_nticket(nrhs) * 4);
}

// timeout flag of the streaming solves (nonzero: a wait gave up)
int lu_solve_check(const void *scratch, int n, hipStream_t st)
{
  unsigned f = 0;
  HIPCHECK(hipMemcpyAsync(&f, (const char *)scratch + sv_align((size_t)n * 4), 4, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return f ? fail(BFA_ERR_HIP, "triangular solve: dependency wait timed out") : BFA_OK;
}

int gemm_sub_z(int m, int n, int k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc,
               hipStream_t st)
{
  return gemm_sub<double>(m, n, k, (const cplx<double> *)A, lda, (const cplx<double> *)B, ldb, (cplx<double> *)C, ldc,
                          st);
}
int gemm_sub_c(int m, int n, int k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc,
               hipStream_t st)
{
  return gemm_sub<float>(m, n, k, (const cplx<float> *)A, lda, (const cplx<float> *)B, ldb, (cplx<float> *)C, ldc, st);
}
int lu_factor_z(void *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb)
{ return lu_factor<double>((cplx<double> *)A, n, lda, ipiv, info, st, nb); }

// pieces of the complex64 LU for the block-local factorisation (bfa_lu_local.hip)
bool profile_on() { return g_prof.on; }

int lu_local_begin(hipStream_t st)
{
  LUWork *lw = nullptr;
  int rc = lu_work(lw);
  if (rc) return rc;
  HIPCHECK(hipMemsetAsync(lw->pw.flags, 0, 8, st));
  return BFA_OK;
}
int lu_local_end(int *info, hipStream_t st)
{
  LUWork *lw = nullptr;
  int rc = lu_work(lw);
  if (rc) return rc;
  int flags[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(flags, lw->pw.flags, 8, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (flags[0]) return fail(BFA_ERR_HIP, "LU panel barrier timed out");
  if (info) *info = flags[1];
  return BFA_OK;
}
void laswp_c(void *A, int64_t lda, int c0, int c1, int k0, int kb, const int *ipiv, hipStream_t st)
{ laswp<float>((cplx<float> *)A, lda, c0, c1, k0, kb, ipiv, st); }
// Narrow column strips (2 right-hand sides per workgroup): these two serve
// the block-local LU's triangular inverses (256 right-hand sides), whose
// one-launch latency is the chain of eight 32-row blocks; a narrower strip
// makes each block's update shorter and the grid wider (U^-1 of a 256 x 256
// block: 146 us with 16 columns per workgroup, 86 with 4, 76 with 2)
void trsm_lower_c(int kb, int n, const void *L, int64_t ldl, void *B, int64_t ldb, hipStream_t st)
{
  if (kb <= 0 || n <= 0) return;
  if (kb > TF_MAXK) { trsm_lower<float>(kb, n, (const cplx<float> *)L, ldl, (cplx<float> *)B, ldb, st); return; }
  hipLaunchKernelGGL((k_trsm_lf<float, 2>), dim3((n + 1) / 2), dim3(256), 0, st, kb, n, (const cplx<float> *)L, ldl,
                     (cplx<float> *)B, ldb);
}
void trsm_upper_c(int kb, int n, const void *U, int64_t ldu, void *B, int64_t ldb, hipStream_t st)
{
  if (kb <= 0 || n <= 0) return;
  hipLaunchKernelGGL((k_trsm_uf<float, 2>), dim3((n + 1) / 2), dim3(256), 0, st, kb, n, (const cplx<float> *)U, ldu,
                     (cplx<float> *)B, ldb);
}
int *lu_local_flags()
{
  LUWork *lw = nullptr;
  return lu_work(lw) == BFA_OK ? lw->pw.flags : nullptr;
}
int lu_factor_c(void *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb)
{ return lu_factor<float>((cplx<float> *)A, n, lda, ipiv, info, st, nb); }
int lu_solve_z(const void *A, int n, int64_t lda, const int *ipiv, void *B, int nrhs, int64_t ldb, hipStream_t st, void *s)
{ return lu_solve<double>((const cplx<double> *)A, n, lda, ipiv, (cplx<double> *)B, nrhs, ldb, st, s); }
int lu_solve_c(const void *A, int n, int64_t lda, const int *ipiv, void *B, int nrhs, int64_t ldb, hipStream_t st, void *s)
{ return lu_solve<float>((const cplx<float> *)A, n, lda, ipiv, (cplx<float> *)B, nrhs, ldb, st, s); }
int lu_solve_prep_c(const void *A, int n, int64_t lda, const int *ipiv, hipStream_t st, void *s)
{ return lu_solve_prep<float>((const cplx<float> *)A, n, lda, ipiv, st, s); }
int lu_solve_apply_c(const void *A, int n, int64_t lda, void *B, int nrhs, int64_t ldb, hipStream_t st, void *s)
{ return lu_solve_apply<float>((const cplx<float> *)A, n, lda, (cplx<float> *)B, nrhs, ldb, st, s); }

void launch_sv_inv_g(const SvProb *d, int np, int nblk, hipStream_t st)
{
  if (np > 0 && nblk > 0) hipLaunchKernelGGL(kg_sv_inv, dim3(nblk, 2), dim3(256), 0, st, d, np);
}
void launch_sv_stream_g(bool lower, const SvProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (lower) hipLaunchKernelGGL(kg_sv_stream<true>, dim3(nblk), dim3(256), 0, st, d, np);
  else hipLaunchKernelGGL(kg_sv_stream<false>, dim3(nblk), dim3(256), 0, st, d, np);
}
int panel_sync_words() { return PB_WORDS; }
unsigned long long sv_tag_c64() { return (unsigned long long)SvTag<float>::bits * 0x100000001ull; }

void launch_gemm_g(bool f64, int tile, const GemmProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (tile != TILE_64) {   // the tile rows of tile_rows(): 64 for TILE_64, else 128
    if (f64) hipLaunchKernelGGL(kg_gemm128<double>, dim3(nblk), dim3(512), 0, st, d, np);
    else hipLaunchKernelGGL(kg_gemm128<float>, dim3(nblk), dim3(512), 0, st, d, np);
  } else {
    if (f64) hipLaunchKernelGGL(kg_gemm64<double>, dim3(nblk), dim3(256), 0, st, d, np);
    else hipLaunchKernelGGL(kg_gemm64<float>, dim3(nblk), dim3(256), 0, st, d, np);
  }
}

void launch_trsm_g(bool f64, bool lower, const TrsmProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (f64) {
    if (lower) hipLaunchKernelGGL((kg_trsm<double, true>), dim3(nblk), dim3(256), 0, st, d, np);
    else hipLaunchKernelGGL((kg_trsm<double, false>), dim3(nblk), dim3(256), 0, st, d, np);
  } else {
    if (lower) hipLaunchKernelGGL((kg_trsm<float, true>), dim3(nblk), dim3(256), 0, st, d, np);
    else hipLaunchKernelGGL((kg_trsm<float, false>), dim3(nblk), dim3(256), 0, st, d, np);
  }
}

int trsm_lf_cols(bool f64) { return f64 ? TfShape<double>::SC : TfShape<float>::SC; }
void launch_trsm_lf_g(bool f64, const TrsmProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (f64) hipLaunchKernelGGL(kg_trsm_lf<double>, dim3(nblk), dim3(256), 0, st, d, np);
  else hipLaunchKernelGGL(kg_trsm_lf<float>, dim3(nblk), dim3(256), 0, st, d, np);
}

void launch_swap_g(bool f64, const SwapProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (f64) hipLaunchKernelGGL(kg_swap<double>, dim3(nblk), dim3(128), 0, st, d, np);
  else hipLaunchKernelGGL(kg_swap<float>, dim3(nblk), dim3(128), 0, st, d, np);
}

size_t panel_lds_cap(bool f64)
{
  static int done[2] = {0, 0};
  const int i = f64 ? 0 : 1;
  const size_t cap = 150 * 1024;
  if (!done[i]) {
    hipError_t e = f64 ? hipFuncSetAttribute((const void *)kg_panel<double>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap)
                       : hipFuncSetAttribute((const void *)kg_panel<float>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap);
    done[i] = e == hipSuccess ? 1 : -1;
  }
  return done[i] > 0 ? cap : 64 * 1024;
}

int panel_max_resident(bool f64, size_t lds)
{
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  hipError_t e = f64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kg_panel<double>, 256, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kg_panel<float>, 256, lds);
  if (e != hipSuccess) return 0;
  return per * ncu;
}

void launch_panel_g(bool f64, const PanelProb *d, int np, int nblk, size_t lds, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (f64) hipLaunchKernelGGL(kg_panel<double>, dim3(nblk), dim3(256), lds, st, d, np);
  else hipLaunchKernelGGL(kg_panel<float>, dim3(nblk), dim3(256), lds, st, d, np);
}


}  // namespace bfa

// ---------------------------------------------------------------------------
extern "C" int bfa_lu_factor_z(void *d_A, int n, int64_t lda, int *d_ipiv, int *info, void *stream)
{
  if (!d_A || !d_ipiv || n <= 0 || lda < n) return fail(BFA_ERR_ARG, "bfa_lu_factor_z: bad argument");
  return bfa::lu_factor_z(d_A, n, lda, d_ipiv, info, (hipStream_t)stream, 256);
}

extern "C" int bfa_lu_solve_z(const void *d_A, int n, int64_t lda, const int *d_ipiv, void *d_B, int nrhs, int64_t ldb,
                              void *stream)
{
  if (!d_A || !d_ipiv || !d_B || n <= 0 || nrhs <= 0 || lda < n || ldb < n)
    return fail(BFA_ERR_ARG, "bfa_lu_solve_z: bad argument");
  void *scratch = nullptr;
  HIPCHECK(hipMalloc(&scratch, bfa::lu_solve_scratch_bytes(n, nrhs, 16)));
  int rc = bfa::lu_solve_z(d_A, n, lda, d_ipiv, d_B, nrhs, ldb, (hipStream_t)stream, scratch);
  if (rc == BFA_OK) rc = bfa::lu_solve_check(scratch, n, (hipStream_t)stream);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  hipFree(scratch);
  if (rc) return rc;
  if (e != hipSuccess) return fail(BFA_ERR_HIP, "bfa_lu_solve_z: %s", hipGetErrorString(e));
  return BFA_OK;
}

extern "C" int bfa_zgemm_sub(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B, int64_t ldb,
                             void *d_C, int64_t ldc, void *stream)
{
  if (m < 0 || n < 0 || k < 0 || (m > 0 && (lda < m || ldc < m)) || (k > 0 && ldb < k))
    return fail(BFA_ERR_ARG, "bfa_zgemm_sub: bad sizes");
  bfa::gemm_sub_z(m, n, k, d_A, lda, d_B, ldb, d_C, ldc, (hipStream_t)stream);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

extern "C" int bfa_cgemm_sub(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B, int64_t ldb,
                             void *d_C, int64_t ldc, void *stream)
{
  if (m < 0 || n < 0 || k < 0 || (m > 0 && (lda < m || ldc < m)) || (k > 0 && ldb < k))
    return fail(BFA_ERR_ARG, "bfa_cgemm_sub: bad sizes");
  bfa::gemm_sub_c(m, n, k, d_A, lda, d_B, ldb, d_C, ldc, (hipStream_t)stream);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

namespace bfa {
// Release every process-lifetime device object of the LU: work areas,
// streams and events of the look-ahead, instrumentation events, library
// handles.  A later LU call creates them again.
void lu_release()
{
  // the handles first: each is bound to a stream released below
  for (auto &e : g_blas) {
    if (e.h && hipSetDevice(e.dev) == hipSuccess) {
      (void)hipDeviceSynchronize();
      (void)rocblas_destroy_handle(e.h);
    }
    e = BlasEntry();
  }
  for (int d = 0; d < 16; d++) {
    LUWork &w = g_lu[d];
    if (!w.mem && !w.side && !w.s_pan) continue;
    if (hipSetDevice(w.device >= 0 ? w.device : d) != hipSuccess) continue;
    (void)hipDeviceSynchronize();
    if (w.mem) (void)hipFree(w.mem);
    if (w.side) (void)hipStreamDestroy(w.side);
    for (auto &e : w.ev) if (e) (void)hipEventDestroy(e);
    if (w.s_pan) (void)hipStreamDestroy(w.s_pan);
    if (w.s_bulk) (void)hipStreamDestroy(w.s_bulk);
    for (auto &e : w.evm) if (e) (void)hipEventDestroy(e);
    w = LUWork();
  }
  for (hipEvent_t e : g_prof.pool) (void)hipEventDestroy(e);
  g_prof = GemmProfile();
}

}  // namespace bfa

extern "C" int bfa_shutdown(void)
{
  int dev = 0;
  const bool have = hipGetDevice(&dev) == hipSuccess;
  bfa::lu_release();
  bfa::local_release();
  if (have) (void)hipSetDevice(dev);
  return BFA_OK;
}

extern "C" int bfa_lu_partition(int on)
{
  bfa::LUWork *lw = nullptr;
  int rc = bfa::lu_work(lw);
  if (rc) return rc;
  lw->mask_off = on == 0;
  return BFA_OK;
}

extern "C" int bfa_profile_enable(int on)
{
  g_prof.on = on != 0;
  return BFA_OK;
}

extern "C" int bfa_profile_read(int which, double *ms, double *flops, int *launches)
{
  // which: 0 / 1 = every complex128 / complex64 trailing-update launch;
  // 2 / 3 = the large updates (m, n >= 256) on all or the update CUs;
  // 4 / 5 = the large updates run as column slices on the panel CUs;
  // 6 = the fast form's split-FP16 launches (read and cleared); 7 = the
  // algorithmic bytes of the launches the last read of 6 returned, in *flops
  if (which == 6) return bfa::profile_read_split(ms, flops, launches);
  if (which == 7) {
    if (ms) *ms = 0.0;
    return bfa::profile_split_bytes(flops, launches);
  }
  if (which < 0 || which > 7) return fail(BFA_ERR_ARG, "bfa_profile_read: which must be 0..7");
  const int type = which & 1;
  const int lo = which < 2 ? 3 * type : which < 4 ? 3 * type + 1 : 3 * type + 2;
  const int hi = which < 2 ? 3 * type + 2 : lo;
  double t = 0.0, f = 0.0;
  int n = 0;
  for (int r = lo; r <= hi; r++) {
    for (auto &pr : g_prof.rec[r]) {
      HIPCHECK(hipEventSynchronize(pr.second));
      float x = 0.0f;
      HIPCHECK(hipEventElapsedTime(&x, pr.first, pr.second));
      t += x;
    }
    f += g_prof.flops[r];
    n += (int)g_prof.rec[r].size();
  }
  if (ms) *ms = t;
  if (flops) *flops = f;
  if (launches) *launches = n;
  // reading 0 / 1 clears the records of that type (every class); 2..5 keep them
  if (which >= 2) return BFA_OK;
  for (int r = 3 * type; r <= 3 * type + 2; r++) {
    g_prof.rec[r].clear();
    g_prof.flops[r] = 0.0;
  }
  bool empty = true;
  for (auto &v : g_prof.rec) empty = empty && v.empty();
  if (empty) g_prof.used = 0;
  return BFA_OK;
}
