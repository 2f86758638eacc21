// bfa_mlp.hip -- the small MLP of the learned near-field surrogate:
// training and inference kernels for an ensemble of E networks
//     x (nin) -> 256 -> 256 -> 256 -> nout,   GELU between layers,
// FP16 weights and activations feeding FP32-accumulating MFMA
// (v_mfma_f32_16x16x16_f16) in the forward pass; FP32 MFMA
// (v_mfma_f32_16x16x4_f32) for the backward products; FP32 master weights
// updated by Adam.  Ensemble members are the z dimension of every launch.
//
// Layout (per member, row-major, K padded to kp = nin rounded up to 16):
//   W1[h][kp] b1[h] W2[h][h] b2[h] W3[h][h] b3[h] W4[nout][h] b4[nout]
// Activations are [batch][features] row-major.
//
// No reference counterpart (the reference integrates every pair by
// cubature, src/libs/libbuff/Cubature.cc:419).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "bfa_internal.h"
#include "bfa_mlp.h"

using namespace bfa;

#define HIPCHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(e_ == hipErrorOutOfMemory ? BFA_ERR_OOM : BFA_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                \
  } while (0)

namespace {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z)
{
  // d/dz [z Phi(z)] = Phi(z) + z phi(z)
  const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
  return Phi + z * phi;
}

// ---------------------------------------------------------------------------
// Forward layer: Z = X W^T + b (FP32 accumulate), A = GELU(Z) as FP16, or
// Y = Z (last layer).  Tile 64 rows x 64 neurons, 4 waves, each wave 16 rows
// x 4 MFMA column tiles; K staged through LDS 128 at a time as 4-half
// vectors (row pitch 33 vectors: the 16 rows a wave reads fall on distinct
// banks).
#define FT 64
#define FKMAX 256
#define FKC 128
template <bool ACT>
__global__ __launch_bounds__(256) void k_fwd(const _Float16 *__restrict__ X, int64_t sx, const _Float16 *__restrict__ W,
                                             const float *__restrict__ bias, int64_t sw, int B, int K, int N,
                                             float *Z, _Float16 *A, float *Y, int64_t so)
{
  __shared__ h4 Xs[FT][FKC / 4 + 1];
  __shared__ h4 Ws[FT][FKC / 4 + 1];
  const int e = blockIdx.z;
  X += e * sx;
  W += e * sw;
  bias += e * sw;   // biases live in the FP32 parameter block, same member stride
  const int row0 = blockIdx.x * FT, col0 = blockIdx.y * FT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  f4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = (f4){0.f, 0.f, 0.f, 0.f};
  const h4 zero = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
  for (int kc = 0; kc < K; kc += FKC) {
    const int nk4 = min(FKC, K - kc) / 4;     // K is a multiple of 16
    for (int idx = tid; idx < FT * nk4; idx += blockDim.x) {
      const int r = idx / nk4, q = idx % nk4;
      Xs[r][q] = (row0 + r < B) ? *(const h4 *)(X + (int64_t)(row0 + r) * K + kc + 4 * q) : zero;
      Ws[r][q] = (col0 + r < N) ? *(const h4 *)(W + (int64_t)(col0 + r) * K + kc + 4 * q) : zero;
    }
    __syncthreads();
    for (int q0 = 0; q0 < nk4; q0 += 4) {
      const h4 a = Xs[wave * 16 + i16][q0 + q4];
#pragma unroll
      for (int c = 0; c < 4; c++)
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, Ws[c * 16 + i16][q0 + q4], acc[c], 0, 0, 0);
    }
    __syncthreads();
  }
  // lane holds D[4*(lane/16) + v][lane % 16] of each 16x16 tile
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int col = col0 + c * 16 + i16;
    if (col >= N) continue;
    const float bb = bias[col];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = row0 + wave * 16 + q4 * 4 + v;
      if (row >= B) continue;
      const float z = acc[c][v] + bb;
      const int64_t o = e * so + (int64_t)row * N + col;
      if (ACT) {
        if (Z) Z[o] = z;
        A[o] = (_Float16)gelu(z);
      } else {
        Y[o] = z;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Generic FP32-MFMA product for the backward pass:
//   C(m, n) = sum_k A(m, k) B(k, n),  A(m,k) = a[m*am + k*ak], B(k,n) = b[k*bk + n*bn]
// with per-member offsets (z dimension) and per-split K ranges (y2 = split).
// EPI 0: C stored; EPI 1: C *= GELU'(Zp(m, n)) (dX -> dZ of the previous
// layer, Zp row-major [m][ldz]).
template <typename TA, typename TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm32(int M, int N, int K, int ksplit, const TA *__restrict__ a, int64_t am,
                                                int64_t ak, int64_t sa, const TB *__restrict__ b, int64_t bk,
                                                int64_t bn, int64_t sb, float *c, int64_t ldc, int64_t sc,
                                                int64_t ssplit, const float *__restrict__ zp, int64_t ldz, int64_t sz)
{
  __shared__ float As[16][FT + 4];
  __shared__ float Bs[16][FT + 4];
  const int e = blockIdx.z;
  const int tiles_m = (M + FT - 1) / FT;
  const int tm = blockIdx.x % tiles_m, split = blockIdx.x / tiles_m;
  const int row0 = tm * FT, col0 = blockIdx.y * FT;
  const int kchunk = (K + ksplit - 1) / ksplit;
  const int kbeg = split * kchunk, kend = min(K, kbeg + kchunk);
  a += e * sa;
  b += e * sb;
  c += e * sc + split * ssplit;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;   // 2 x 2 waves of 32 x 32
  const int i16 = lane & 15, kq = lane >> 4;
  f4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    for (int idx = tid; idx < 16 * FT; idx += blockDim.x) {
      const int kk = idx / FT, r = idx % FT;
      const int k = k0 + kk;
      As[kk][r] = (k < kend && row0 + r < M) ? (float)a[(int64_t)(row0 + r) * am + (int64_t)k * ak] : 0.f;
      Bs[kk][r] = (k < kend && col0 + r < N) ? (float)b[(int64_t)k * bk + (int64_t)(col0 + r) * bn] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(As[kk + kq][wr * 32 + x * 16 + i16],
                                                           Bs[kk + kq][wc * 32 + y * 16 + i16], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  // f32 16x16x4: lane holds D[4*(lane/16) + v][lane % 16]
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) {
      const int col = col0 + wc * 32 + y * 16 + i16;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = row0 + wr * 32 + x * 16 + kq * 4 + v;
        if (row >= M) continue;
        float r = acc[x][y][v];
        if (EPI == 1) r *= gelu_grad(zp[e * sz + (int64_t)row * ldz + col]);
        c[(int64_t)row * ldc + col] = r;
      }
    }
}

// dst[e][i] = sum_s src[e][s][i]   (split-K partials of the weight gradients)
__global__ void k_sum_splits(const float *src, int nsplit, int64_t n, int64_t s_src, float *dst, int64_t s_dst)
{
  const int e = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float acc = 0.f;
    for (int s = 0; s < nsplit; s++) acc += src[e * s_src + s * n + i];
    dst[e * s_dst + i] = acc;
  }
}

// db[e][n] = sum_b dZ[e][b][n]; one thread per column, fixed order
__global__ void k_colsum(const float *dz, int B, int N, int64_t sdz, float *db, int64_t sdb)
{
  const int e = blockIdx.y;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float acc = 0.f;
  for (int b = 0; b < B; b++) acc += dz[e * sdz + (int64_t)b * N + n];
  db[e * sdb + n] = acc;
}

// (Weighted) MSE over all outputs: loss = sum w_b (Y - T)^2 / (B * nout);
// dY = 2 w_b (Y - T) / (B * nout); w = NULL means w_b = 1.
__global__ __launch_bounds__(256) void k_mse(const float *Y, const float *T, const float *w, int B, int nout,
                                             int64_t s, float *dY, double *loss)
{
  __shared__ double red[256];
  const int e = blockIdx.y;
  const int64_t n = (int64_t)B * nout;
  const float scale = 2.0f / (float)n;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float d = Y[e * s + i] - T[e * s + i];
    const float wb = w ? w[(int64_t)e * B + i / nout] : 1.0f;
    dY[e * s + i] = scale * wb * d;
    acc += (double)wb * d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[e * gridDim.x + blockIdx.x] = red[0] / (double)n;
}

// Adam on the FP32 master weights, FP16 copy refreshed
__global__ void k_adam(float *w, const float *g, float *m, float *v, _Float16 *w16, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = (_Float16)wi;
  }
}

__global__ void k_to_half(const float *src, _Float16 *dst, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (_Float16)src[i];
}

}  // namespace

namespace bfa {

MLPShape mlp_shape(int nin, int hidden, int nout)
{
  MLPShape s{};
  s.nin = nin;
  s.kp = (nin + 15) / 16 * 16;
  s.h = hidden;
  s.nout = nout;
  int64_t o = 0;
  s.oW[0] = o; o += (int64_t)hidden * s.kp; s.ob[0] = o; o += hidden;
  s.oW[1] = o; o += (int64_t)hidden * hidden; s.ob[1] = o; o += hidden;
  s.oW[2] = o; o += (int64_t)hidden * hidden; s.ob[2] = o; o += hidden;
  s.oW[3] = o; o += (int64_t)nout * hidden; s.ob[3] = o; o += nout;
  s.nparam = o;
  return s;
}

int64_t mlp_work_floats(const MLPShape &s, int B)
{
  // per member: Z1..Z3 (B*h each), Y (B*nout), dY, dZ (B*h, reused), dX (B*h)
  // + split-K partials (nsplit * h * h)
  return (int64_t)B * s.h * 5 + (int64_t)B * s.nout * 2 + (int64_t)MLP_KSPLIT * s.h * s.h;
}

int64_t mlp_work_halfs(const MLPShape &s, int B)
{
  return (int64_t)B * s.h * 3;   // A1..A3
}

// Forward pass of E members.  X16: [E][B][kp] (sx = 0 shares one input),
// W16: [E][nparam] FP16 (biases read from the FP32 master copy Wf when
// given, else converted on the fly is not needed: biases live in bias32).
int mlp_forward(const MLPShape &s, int E, int B, const _Float16 *X16, int64_t sx, const _Float16 *W16,
                const float *bias32, float *Y, float *Zsave, _Float16 *Abuf, hipStream_t st)
{
  if (s.kp > FKMAX || s.h > FKMAX || s.h % 16 || s.nout % 16) return fail(BFA_ERR_ARG, "mlp: unsupported shape");
  const int64_t np = s.nparam;
  const dim3 blk(256);
  const int tb = (B + FT - 1) / FT;
  const _Float16 *in = X16;
  int64_t sin = sx;
  int K = s.kp;
  for (int l = 0; l < 3; l++) {
    _Float16 *A = Abuf + (int64_t)l * E * B * s.h;
    float *Z = Zsave ? Zsave + (int64_t)l * E * B * s.h : nullptr;
    hipLaunchKernelGGL(k_fwd<true>, dim3(tb, (s.h + FT - 1) / FT, E), blk, 0, st, in, sin, W16 + s.oW[l],
                       bias32 + s.ob[l], np, B, K, s.h, Z, A, (float *)nullptr, (int64_t)B * s.h);
    in = A;
    sin = (int64_t)B * s.h;
    K = s.h;
  }
  hipLaunchKernelGGL(k_fwd<false>, dim3(tb, (s.nout + FT - 1) / FT, E), blk, 0, st, in, sin, W16 + s.oW[3],
                     bias32 + s.ob[3], np, B, K, s.nout, (float *)nullptr, (_Float16 *)nullptr, Y,
                     (int64_t)B * s.nout);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

// Backward pass given dY [E][B][nout]: gradients [E][nparam] (FP32).
int mlp_backward(const MLPShape &s, int E, int B, const _Float16 *X16, int64_t sx, const _Float16 *W16,
                 const float *Zsave, const _Float16 *Abuf, const float *dY, float *grad, float *work, hipStream_t st)
{
  const int64_t np = s.nparam;
  const int64_t BH = (int64_t)B * s.h;
  float *dZ = work;                          // [E][B][h]
  float *dX = work + (int64_t)E * BH;        // [E][B][h]
  float *part = work + (int64_t)2 * E * BH;  // [E][ksplit][h*h]
  const int64_t spart = (int64_t)MLP_KSPLIT * s.h * s.h;
  const dim3 blk(256);
  const float *dcur = dY;
  int ncur = s.nout;
  int64_t sdcur = (int64_t)B * s.nout;
  for (int l = 3; l >= 0; l--) {
    const int Kin = l == 0 ? s.kp : s.h;
    const _Float16 *Xin = l == 0 ? X16 : Abuf + (int64_t)(l - 1) * E * BH;
    const int64_t sxin = l == 0 ? sx : BH;
    // dW_l[n][k] = sum_b dcur[b][n] Xin[b][k]   (split over b, then summed)
    {
      const int tm = (ncur + FT - 1) / FT, tn = (Kin + FT - 1) / FT;
      hipLaunchKernelGGL((k_gemm32<float, _Float16, 0>), dim3(tm * MLP_KSPLIT, tn, E), blk, 0, st, ncur, Kin, B,
                         MLP_KSPLIT, dcur, (int64_t)1, (int64_t)ncur, sdcur, Xin, (int64_t)Kin, (int64_t)1, sxin,
                         part, (int64_t)Kin, spart, (int64_t)ncur * Kin, (const float *)nullptr, (int64_t)0,
                         (int64_t)0);
      const int64_t n = (int64_t)ncur * Kin;
      hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024), E), blk, 0, st,
                         (const float *)part, MLP_KSPLIT, n, spart, grad + s.oW[l], np);
      hipLaunchKernelGGL(k_colsum, dim3((ncur + 255) / 256, E), blk, 0, st, dcur, B, ncur, sdcur, grad + s.ob[l], np);
    }
    if (l == 0) break;
    // dX[b][k] = sum_n dcur[b][n] W[n][k], then dZ_{l-1} = dX * GELU'(Z_{l-1})
    {
      const int tm = (B + FT - 1) / FT, tn = (s.h + FT - 1) / FT;
      float *dst = (dcur == dZ) ? dX : dZ;
      hipLaunchKernelGGL((k_gemm32<float, _Float16, 1>), dim3(tm, tn, E), blk, 0, st, B, s.h, ncur, 1, dcur,
                         (int64_t)ncur, (int64_t)1, sdcur, W16 + s.oW[l], (int64_t)Kin, (int64_t)1, np, dst,
                         (int64_t)s.h, BH, (int64_t)0, Zsave + (int64_t)(l - 1) * E * BH, (int64_t)s.h, BH);
      dcur = dst;
      ncur = s.h;
      sdcur = BH;
    }
  }
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

int mlp_mse(int E, int B, int nout, const float *Y, const float *T, const float *w, float *dY, double *loss_parts,
            int nparts, hipStream_t st)
{
  hipLaunchKernelGGL(k_mse, dim3(nparts, E), dim3(256), 0, st, Y, T, w, B, nout, (int64_t)B * nout, dY, loss_parts);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

int mlp_adam(int64_t n, float *w, const float *g, float *m, float *v, _Float16 *w16, float lr, int step,
             hipStream_t st)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, st, w, g, m, v,
                     w16, n, lr, b1, b2, eps, c1, c2);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

int to_half(const float *src, _Float16 *dst, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(k_to_half, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, st, src, dst, n);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

}  // namespace bfa

// ---------------------------------------------------------------------------
// One-shot C entry points (tests, tooling)

extern "C" int bfa_mlp_param_count(int nin, int hidden, int nout, int64_t *count, int *kp)
{
  if (nin <= 0 || hidden <= 0 || nout <= 0) return fail(BFA_ERR_ARG, "bfa_mlp_param_count: bad shape");
  MLPShape s = mlp_shape(nin, hidden, nout);
  if (count) *count = s.nparam;
  if (kp) *kp = s.kp;
  return BFA_OK;
}

// Loss, gradients and outputs of E members on one batch (host arrays):
// params [E][nparam] (FP32 master; the forward uses their FP16 rounding),
// X [E][B][nin] (FP32, rounded to FP16 on input), T and Y [E][B][nout].
extern "C" int bfa_mlp_grad(int nin, int hidden, int nout, int E, const float *params, int B, const float *X,
                            const float *T, double *loss, float *grad, float *Y)
{
  if (E <= 0 || B <= 0 || !params || !X || !T) return fail(BFA_ERR_ARG, "bfa_mlp_grad: bad argument");
  MLPShape s = mlp_shape(nin, hidden, nout);
  if (s.kp > 256 || hidden > 256 || hidden % 64 || nout % 16) return fail(BFA_ERR_ARG, "bfa_mlp_grad: unsupported shape");
  const int64_t np = s.nparam;
  std::vector<_Float16> xh((size_t)E * B * s.kp, (_Float16)0.0f);
  for (int e = 0; e < E; e++)
    for (int b = 0; b < B; b++)
      for (int k = 0; k < nin; k++) xh[((size_t)e * B + b) * s.kp + k] = (_Float16)X[((size_t)e * B + b) * nin + k];
  const int nparts = 64;
  size_t bytes = (size_t)E * np * (4 + 2 + 4) + xh.size() * 2 + (size_t)E * B * nout * 4 * 3 +
                 (size_t)E * mlp_work_floats(s, B) * 4 + (size_t)E * mlp_work_halfs(s, B) * 2 + (size_t)E * B * s.h * 3 * 4 +
                 (size_t)E * nparts * 8 + 16 * 256;
  char *mem = nullptr;
  HIPCHECK(hipMalloc(&mem, bytes));
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = (o + 255) / 256 * 256; o = r + b; return mem + r; };
  float *W = (float *)take((size_t)E * np * 4);
  _Float16 *W16 = (_Float16 *)take((size_t)E * np * 2);
  float *G = (float *)take((size_t)E * np * 4);
  _Float16 *X16 = (_Float16 *)take(xh.size() * 2);
  float *Td = (float *)take((size_t)E * B * nout * 4);
  float *Yd = (float *)take((size_t)E * B * nout * 4);
  float *dY = (float *)take((size_t)E * B * nout * 4);
  float *Z = (float *)take((size_t)E * B * s.h * 3 * 4);
  _Float16 *A = (_Float16 *)take((size_t)E * mlp_work_halfs(s, B) * 2);
  float *work = (float *)take((size_t)E * mlp_work_floats(s, B) * 4);
  double *lp = (double *)take((size_t)E * nparts * 8);
  hipStream_t st = nullptr;
  int rc = BFA_OK;
  auto go = [&]() -> int {
    HIPCHECK(hipMemcpy(W, params, (size_t)E * np * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(X16, xh.data(), xh.size() * 2, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(Td, T, (size_t)E * B * nout * 4, hipMemcpyHostToDevice));
    int r = to_half(W, W16, (int64_t)E * np, st);
    if (r) return r;
    r = mlp_forward(s, E, B, X16, (int64_t)B * s.kp, W16, W, Yd, Z, A, st);
    if (r) return r;
    r = mlp_mse(E, B, nout, Yd, Td, nullptr, dY, lp, nparts, st);
    if (r) return r;
    r = mlp_backward(s, E, B, X16, (int64_t)B * s.kp, W16, Z, A, dY, G, work, st);
    if (r) return r;
    HIPCHECK(hipDeviceSynchronize());
    if (grad) HIPCHECK(hipMemcpy(grad, G, (size_t)E * np * 4, hipMemcpyDeviceToHost));
    if (Y) HIPCHECK(hipMemcpy(Y, Yd, (size_t)E * B * nout * 4, hipMemcpyDeviceToHost));
    if (loss) {
      std::vector<double> h((size_t)E * nparts);
      HIPCHECK(hipMemcpy(h.data(), lp, h.size() * 8, hipMemcpyDeviceToHost));
      for (int e = 0; e < E; e++) {
        double sacc = 0.0;
        for (int p = 0; p < nparts; p++) sacc += h[(size_t)e * nparts + p];
        loss[e] = sacc;
      }
    }
    return BFA_OK;
  };
  rc = go();
  (void)hipFree(mem);
  return rc;
}
