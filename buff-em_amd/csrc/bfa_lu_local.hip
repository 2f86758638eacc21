// bfa_lu_local.hip -- complex64 LU with block-local pivoting and trailing
// updates on the FP16 matrix cores, the factorisation behind the fast
// refinement form.
//
// No reference counterpart: the reference factors with LAPACK zgetrf
// (HMatrix::LUFactorize, buff-scatter.cc:285).  This factorisation is only
// a preconditioner for FP64 iterative refinement against the resident FP64
// matrix (bfa_solve_refine), which restores FP64 accuracy; it trades the
// partial-pivoting search over all rows -- one grid-wide barrier per column,
// the latency-bound critical path of the complex64 LU -- for two changes:
//
//  - Pivoting within each 256-row diagonal block.  Step j factors the
//    diagonal block with partial pivoting over its own rows (the panel
//    recursion of the partial-pivoting LU, run on jb rows), swaps those rows
//    in the other columns, forms U12 = L11^-1 A12 (unit-lower TRSM) and
//    L21 = A21 U11^-1 (U11^-1 by a one-launch upper TRSM, then a product on
//    the matrix cores) -- no synchronisation across workgroups anywhere.
//    The diagonal block's own factorisation is one launch of one workgroup
//    (k_diag_lu).  The VIE matrices factor stably this way (growth ~1;
//    partial pivoting leaves >99.5 % of the rows in place on them).
//  - The trailing update A22 -= L21 U12 in split FP16: every FP32 operand x
//    (scaled by a power of two per row of L21 / column of U12 so that the
//    row's or column's largest magnitude lies in [2^14, 2^15)) is written as
//    hi + lo with hi = fp16(x), lo = fp16(x - hi), and each real product is
//    hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 with FP32
//    accumulation (4 real products per complex product, 12 MFMA per 16x16x32
//    block).  hi + lo carries 22 significant bits; the dropped lo*lo term is
//    below 2^-22 relative; entries under 2^-17 of their row's (column's)
//    maximum lose relative precision in lo but stay within 2^-38 of that
//    maximum absolutely.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <utility>
#include <vector>

#include "bfa_internal.h"
#include "bfa_group.h"
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

struct c2 { float re, im; };
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define LB 256          // block width of the factorisation (= max k of the split operands)
#define HT 128          // GEMM workgroup tile (rows and columns)
#define HK 32           // k per stage = K of one MFMA

// power-of-two scale bringing max into [2^14, 2^15): returns e with max * 2^e there
__device__ __forceinline__ int split_exp(float mx)
{
  if (!(mx > 0.0f) || !isfinite(mx)) return 0;
  return 14 - ilogbf(mx);
}

__device__ __forceinline__ void split_one(float x, _Float16 &hi, _Float16 &lo)
{
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// L21 (m x k, column-major) -> planes P[4][mp][kp] (re hi, re lo, im hi,
// im lo; row-major, k contiguous) scaled by 2^e per row, rfac[r] = 2^-e.
// Rows in [m, mp) and k in [k, kp) are zero.  Block of 32 rows and KC
// chunks of 32 k (k <= 32 KC); thread (row = t % 32, chunk = t / 32) holds
// 32 consecutive k of its row.
template <int KC>
__device__ __forceinline__ void split_rows_body(const c2 *__restrict__ A, int64_t lda, int m, int k, int kp,
                                                int64_t mp, _Float16 *P, float *rfac, int blk, float (&smax)[KC][32])
{
  const int row = threadIdx.x & 31, kc = threadIdx.x >> 5;
  const int r = blk * 32 + row;
  float re[32], im[32];
  float mx = 0.0f;
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int kk = kc * 32 + j;
    c2 v = {0.0f, 0.0f};
    if (r < m && kk < k) v = A[(size_t)kk * lda + r];
    re[j] = v.re;
    im[j] = v.im;
    mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
  }
  smax[kc][row] = mx;
  __syncthreads();
  mx = 0.0f;
#pragma unroll
  for (int i = 0; i < KC; i++) mx = fmaxf(mx, smax[i][row]);
  const int e = split_exp(mx);
  if (kc == 0 && r < m) rfac[r] = ldexpf(1.0f, -e);
  if (kc * 32 >= kp) return;
  const float s = ldexpf(1.0f, e);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      _Float16 a, b;
      split_one(re[q * 8 + j] * s, a, b);
      rh[j] = a; rl[j] = b;
      split_one(im[q * 8 + j] * s, a, b);
      ih[j] = a; il[j] = b;
    }
    const size_t o = (size_t)r * kp + kc * 32 + q * 8;
    *(h8 *)(P + o) = rh;
    *(h8 *)(P + (size_t)mp * kp + o) = rl;
    *(h8 *)(P + 2 * (size_t)mp * kp + o) = ih;
    *(h8 *)(P + 3 * (size_t)mp * kp + o) = il;
  }
}

template <int KC>
__global__ __launch_bounds__(32 * KC) void k_split_rows(const c2 *__restrict__ A, int64_t lda, int m, int k, int kp,
                                                    int64_t mp, _Float16 *P, float *rfac)
{
  __shared__ float smax[KC][32];
  split_rows_body<KC>(A, lda, m, k, kp, mp, P, rfac, blockIdx.x, smax);
}

// U12 (k x n, column-major, k contiguous) -> planes P[4][np][kp] scaled by
// 2^e per column, cfac[c] = 2^-e.  Thread (column = t / KC, chunk = t % KC)
// holds 32 consecutive k of its column (k <= 32 KC); the column maximum by
// shuffles.  (Reading the block's columns along k into LDS first, for
// fully contiguous loads, left the factorisation's time unchanged.)
template <int KC>
__device__ __forceinline__ void split_cols_body(const c2 *__restrict__ B, int64_t ldb, int k, int n, int kp,
                                                int64_t np, _Float16 *P, float *cfac, int blk)
{
  const int kc = threadIdx.x % KC;
  const int c = blk * 32 + threadIdx.x / KC;
  float re[32], im[32];
  float mx = 0.0f;
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int kk = kc * 32 + j;
    c2 v = {0.0f, 0.0f};
    if (c < n && kk < k) v = B[(size_t)c * ldb + kk];
    re[j] = v.re;
    im[j] = v.im;
    mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
  }
  mx = fmaxf(mx, __shfl_xor(mx, 1));
  mx = fmaxf(mx, __shfl_xor(mx, 2));
  mx = fmaxf(mx, __shfl_xor(mx, 4));
  if (KC > 8) mx = fmaxf(mx, __shfl_xor(mx, 8));
  const int e = split_exp(mx);
  if (kc == 0 && c < n) cfac[c] = ldexpf(1.0f, -e);
  if (kc * 32 >= kp) return;
  const float s = ldexpf(1.0f, e);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    h8 rh, rl, ih, il;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      _Float16 a, b;
      split_one(re[q * 8 + j] * s, a, b);
      rh[j] = a; rl[j] = b;
      split_one(im[q * 8 + j] * s, a, b);
      ih[j] = a; il[j] = b;
    }
    const size_t o = (size_t)c * kp + kc * 32 + q * 8;
    *(h8 *)(P + o) = rh;
    *(h8 *)(P + (size_t)np * kp + o) = rl;
    *(h8 *)(P + 2 * (size_t)np * kp + o) = ih;
    *(h8 *)(P + 3 * (size_t)np * kp + o) = il;
  }
}

template <int KC>
__global__ __launch_bounds__(32 * KC) void k_split_cols(const c2 *__restrict__ B, int64_t ldb, int k, int n, int kp,
                                                    int64_t np, _Float16 *P, float *cfac)
{
  split_cols_body<KC>(B, ldb, k, n, kp, np, P, cfac, blockIdx.x);
}

// ---------------------------------------------------------------------------
// C(m x n, complex64) -= A B (or C = A B) from the split planes.  Workgroup
// tile 128 x 128 in 16 x 16 MFMA blocks; k staged 32 at a time through LDS
// ([plane][row][32 halves], 16-byte chunks XOR-swizzled by row so that a
// fragment read hits distinct banks), the next stage prefetched into
// registers while the current one computes.  Fragment layout of v_mfma_f32_16x16x32_f16:
// lane l supplies row l % 16 of its first operand and column l % 16 of its
// second, 8 halves each at k offset 8 (l / 16) of the stage (any k order
// common to both gives the same sum); the result block holds
// D(4 (l / 16) + v, l % 16) in register v.  Here the first operand is B's
// fragment (column c of the product), the second A's (row r), so D is the
// product block transposed.
//
// Tile order: super-tiles of G x G tiles, dealt round-robin to the 8 XCDs
// (the dispatcher deals workgroups to XCDs round-robin, so workgroup b runs
// on XCD b % 8): XCD x works through super-tiles x, x + 8, ..., each one's
// G^2 tiles together, so a super-tile's row slab of A and column slab of B
// (G x 128 rows / columns, 1 MB each at G = 4) stay in that XCD's L2 while
// its CUs work through them.  Tiles outside the matrix exit at once.
__device__ __forceinline__ bool super_tile(int bid, int tiles_m, int tiles_n, int G, int &ti, int &tj)
{
  const int xcd = bid & 7, l = bid >> 3, gg = G * G;
  const int s = (l / gg) * 8 + xcd, w = l % gg;
  const int sm = (tiles_m + G - 1) / G, sn = (tiles_n + G - 1) / G;
  if (s >= sm * sn) return false;
  ti = (s % sm) * G + w % G;
  tj = (s / sm) * G + w / G;
  return ti < tiles_m && tj < tiles_n;
}

__device__ __forceinline__ h8 hneg(h8 x)
{
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  u4 v = __builtin_bit_cast(u4, x);
  v ^= 0x80008000u;
  return __builtin_bit_cast(h8, v);
}

// NW = 8 waves in 2 x 4 of 64 x 32 each, two LDS buffers (128 KB, one
// barrier per stage), one workgroup of two waves per SIMD (181 VGPRs, no
// spill).  Measured at 19,616 x 19,360 x 256 (the largest trailing update)
// with G = 4: 2.86 ms, 818 TF/s of FP16 products (33 % of the 2.5 PF/s
// peak).  Variants measured there: without the C loads 2.58 ms; without
// operand loads after the first stage 2.84 ms; without LDS fragment reads
// 2.95 ms; row blocks interleaved in pairs 2.90 ms; four waves of 64 x 64
// fed by direct global-to-LDS loads (one wave per SIMD) 3.61 ms -- neither
// LDS nor operand traffic limits it; the C read-modify-write (16 B per
// entry per 24 x 256 FP16 products, near the 312 flop/B ridge) and the
// unoverlapped epilogue of one workgroup per CU do.  Issuing half of the C
// loads during the last k stage (229 VGPRs; all of them spill at 256)
// left the factorisation's time unchanged.
template <int NW>
__device__ __forceinline__ void gemm_hs_tile(int m, int n, int kp, const _Float16 *__restrict__ PA, int64_t mp,
                                             const _Float16 *__restrict__ PB, int64_t np,
                                             const float *__restrict__ rfac, const float *__restrict__ cfac, c2 *C,
                                             int64_t ldc, int row0, int col0, int store,
                                             _Float16 (*As)[4][HT][HK], _Float16 (*Bs)[4][HT][HK])
{
  constexpr bool DB = NW == 8;
  constexpr int CB = NW == 4 ? 4 : 2;           // 16-column blocks per wave
  constexpr int NT = 64 * NW;
  constexpr int NL = 2048 / NT;                 // 16-byte chunks per thread per operand and stage
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / (NW / 2), wc = wave % (NW / 2);
  const int l16 = lane & 15, lq = lane >> 4;
  const size_t pa = (size_t)mp * kp, pb = (size_t)np * kp;

  h8 ga[NL], gb[NL];
  auto gload = [&](int s) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int q = tid + NT * i;
      const int p = q >> 9, r = (q >> 2) & 127, ch = q & 3;
      ga[i] = *(const h8 *)(PA + p * pa + (size_t)(row0 + r) * kp + s * HK + ch * 8);
      gb[i] = *(const h8 *)(PB + p * pb + (size_t)(col0 + r) * kp + s * HK + ch * 8);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int q = tid + NT * i;
      const int p = q >> 9, r = (q >> 2) & 127, ch = q & 3;
      const int sw = ch ^ ((r >> 2) & 3);
      *(h8 *)&As[buf][p][r][sw * 8] = ga[i];
      *(h8 *)&Bs[buf][p][r][sw * 8] = gb[i];
    }
  };

  f4 ar[4][CB], ai[4][CB];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < CB; y++)
#pragma unroll
      for (int v = 0; v < 4; v++) ar[x][y][v] = ai[x][y][v] = 0.0f;

  const int nst = kp / HK;
  gload(0);
  if (DB) {
    sstore(0);
    __syncthreads();
  }
  for (int s = 0; s < nst; s++) {
    const int buf = DB ? (s & 1) : 0;
    if (!DB) {
      sstore(0);
      __syncthreads();
    }
    if (s + 1 < nst) gload(s + 1);
    h8 b[CB][4];
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int c = wc * (16 * CB) + cb * 16 + l16;
      const int sw = lq ^ ((c >> 2) & 3);
#pragma unroll
      for (int p = 0; p < 4; p++) b[cb][p] = *(const h8 *)&Bs[buf][p][c][sw * 8];
    }
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      const int r = wr * 64 + rb * 16 + l16;
      const int sw = lq ^ ((r >> 2) & 3);
      h8 a[4];
#pragma unroll
      for (int p = 0; p < 4; p++) a[p] = *(const h8 *)&As[buf][p][r][sw * 8];
      const h8 n2 = hneg(a[2]), n3 = hneg(a[3]);
      // Re += Ar Br - Ai Bi ; Im += Ar Bi + Ai Br, each real product as
      // hi*hi + hi*lo + lo*hi; alternate accumulators between dependent MFMAs
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][0], a[0], ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][2], a[0], ai[rb][cb], 0, 0, 0);
      }
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][1], a[0], ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][3], a[0], ai[rb][cb], 0, 0, 0);
      }
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][0], a[1], ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][2], a[1], ai[rb][cb], 0, 0, 0);
      }
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][2], n2, ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][0], a[2], ai[rb][cb], 0, 0, 0);
      }
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][3], n2, ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][1], a[2], ai[rb][cb], 0, 0, 0);
      }
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][2], n3, ar[rb][cb], 0, 0, 0);
        ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[cb][0], a[3], ai[rb][cb], 0, 0, 0);
      }
    }
    if (DB && s + 1 < nst) sstore(buf ^ 1);
    __syncthreads();
  }
  // C -= (A B) 2^-(e_row + e_col), or C = (A B) 2^-(e_row + e_col) (store).
  // The accumulators hold the product transposed (B fragment as the MFMA's
  // first operand): register v of block (rb, cb) is row l16, column
  // 4 lq + v, so 16 lanes cover 16 consecutive rows of one column -- 128
  // contiguous bytes per column and instruction.  All loads of C are issued
  // before the first store.
  float rf[4], cf[CB][4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int row = row0 + wr * 64 + rb * 16 + l16;
    rf[rb] = row < m ? rfac[row] : 0.0f;
  }
#pragma unroll
  for (int cb = 0; cb < CB; cb++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int col = col0 + wc * (16 * CB) + cb * 16 + 4 * lq + v;
      cf[cb][v] = col < n ? cfac[col] : 0.0f;
    }
  c2 x[4][CB][4];
  if (!store) {
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
      for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = row0 + wr * 64 + rb * 16 + l16;
          const int col = col0 + wc * (16 * CB) + cb * 16 + 4 * lq + v;
          x[rb][cb][v] = (row < m && col < n) ? C[(size_t)col * ldc + row] : c2{0.0f, 0.0f};
        }
  }
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = row0 + wr * 64 + rb * 16 + l16;
        const int col = col0 + wc * (16 * CB) + cb * 16 + 4 * lq + v;
        if (row < m && col < n) {
          const float f = rf[rb] * cf[cb][v];
          c2 y = x[rb][cb][v];
          if (store) {
            y.re = ar[rb][cb][v] * f;
            y.im = ai[rb][cb][v] * f;
          } else {
            y.re -= ar[rb][cb][v] * f;
            y.im -= ai[rb][cb][v] * f;
          }
          C[(size_t)col * ldc + row] = y;
        }
      }
}

__global__ __launch_bounds__(512, 1) void k_gemm_hs(int m, int n, int kp, const _Float16 *__restrict__ PA, int64_t mp,
                                                     const _Float16 *__restrict__ PB, int64_t np,
                                                     const float *__restrict__ rfac, const float *__restrict__ cfac,
                                                     c2 *C, int64_t ldc, int tiles_m, int tiles_n, int store, int G)
{
  __shared__ __align__(16) _Float16 As[2][4][HT][HK];
  __shared__ __align__(16) _Float16 Bs[2][4][HT][HK];
  int ti, tj;
  if (!super_tile(blockIdx.x, tiles_m, tiles_n, G, ti, tj)) return;
  gemm_hs_tile<8>(m, n, kp, PA, mp, PB, np, rfac, cfac, C, ldc, ti * HT, tj * HT, store, As, Bs);
}

// ---------------------------------------------------------------------------
// The jb x jb diagonal block (jb <= 256) with partial pivoting over its rows
// in one workgroup of 1024 threads.  The block stays in global memory (512
// KB, L2-resident); panels of 32 columns, each:
//   leaf    rows c0..jb-1 of the panel.  Thread t keeps the panel part of
//           row t (its data row) in registers and tracks the row's position:
//           an interchange only swaps two positions.  The columns run in
//           windows of four held at x[0..3] (the window's loop body is the
//           same for every window and stays small; x shifts by four between
//           windows).  Per column d: the candidate |re| + |im| of each row at
//           a position >= d as the key (bits of |re| + |im|) << 32 |
//           ~position, whose maximum -- DPP within each wave, then the four
//           rows' waves through LDS -- is the izamax pivot (largest
//           magnitude, first position among equals); the pivot row's values
//           go to every thread through LDS, the rows below scale and update.
//           A zero pivot is recorded (first one, 1-based global column) and
//           the column is left as it is, as LAPACK's zgetf2 does.  Each
//           finished column value goes to LDS by data row; at the end every
//           row lands at its final position in A and in LDS (L11 on top of
//           L21), and perm[position] = data row;
//   swaps   that permutation applied to the block's other columns for the
//           positions the interchanges touched (all loads, a barrier, all
//           stores);
//   U12     L11^-1 A12 for the panel's rows in four blocks of eight rows,
//           into A and LDS;
//   update  L21 and U12 split into FP16 hi + lo planes (each row of L21 /
//           column of U12 scaled by a power of two, as in the trailing
//           update), then A22 -= L21 U12 in 16 x 16 tiles, twelve
//           v_mfma_f32_16x16x32_f16 per tile (U12's fragment as the first
//           operand, so that register v of lane l holds row l % 16, column
//           4 (l / 16) + v of the tile), four tiles of a wave at a time.
// Barriers separate the phases; a workgroup barrier also orders the global
// stores of one phase before the loads of the next.  Pivots are written
// absolute (piv_off + local row).
#define LEAF_W 32
#define DL_NT 1024
#define DL_NW 16
#define DL_M 256
// One max-reduction step of a 64-bit key across lanes: DPP control CTRL on
// the rows enabled by RM; a lane whose source is disabled or out of range
// takes 0 (no candidate).
template <int CTRL, int RM>
__device__ __forceinline__ void key_max_step(unsigned &lo, unsigned &hi)
{
  const unsigned ol = __builtin_amdgcn_update_dpp(0u, lo, CTRL, RM, 0xF, false);
  const unsigned oh = __builtin_amdgcn_update_dpp(0u, hi, CTRL, RM, 0xF, false);
  const unsigned long long o = ((unsigned long long)oh << 32) | ol, k = ((unsigned long long)hi << 32) | lo;
  const bool take = o > k;
  lo = take ? ol : lo;
  hi = take ? oh : hi;
}

// maximum of k over the wavefront (all lanes active), returned uniform:
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and row_mirror reduce
// each row of 16 lanes; row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2,
// 3) carry the rows' maxima up to lane 63
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long k)
{
  unsigned lo = (unsigned)k, hi = (unsigned)(k >> 32);
  key_max_step<0xB1, 0xF>(lo, hi);
  key_max_step<0x4E, 0xF>(lo, hi);
  key_max_step<0x141, 0xF>(lo, hi);
  key_max_step<0x140, 0xF>(lo, hi);
  key_max_step<0x142, 0xA>(lo, hi);
  key_max_step<0x143, 0xC>(lo, hi);
  lo = __builtin_amdgcn_readlane(lo, 63);
  hi = __builtin_amdgcn_readlane(hi, 63);
  return ((unsigned long long)hi << 32) | lo;
}

// 16-byte chunk position of a split plane row (64 bytes = 4 chunks):
// chunk ^ swizzle(row), so that 16 lanes reading one chunk of 16
// consecutive rows hit distinct banks
__device__ __forceinline__ int diag_swz(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }

struct DiagShared {
  union {
    c2 L[DL_M][LEAF_W + 1];   // the panel by position, rows c0..jb-1 (stride 33: 16 lanes on distinct banks)
    _Float16 LP[4][DL_M - LEAF_W][LEAF_W];   // then L21's split planes (re hi, re lo, im hi, im lo), 16-byte chunks swizzled
  };
  union {
    c2 U[LEAF_W][DL_M];       // U12, columns counted from c0 + 32
    _Float16 UP[4][DL_M - LEAF_W][LEAF_W];   // then U12's split planes by column
  };
  float rfs[DL_M], cfs[DL_M];   // 2^-e of L21's rows and U12's columns
  unsigned long long wkey[2][4];  // argmax keys of the column's four waves (slots alternate between columns)
  c2 colv[2][DL_M];           // the current column by data row
  int piv[LEAF_W];            // the panel's pivots (local positions)
  c2 prow[2][LEAF_W];         // pivot row
  int perm[DL_M];
  int mv[2 * LEAF_W], nmv;    // positions touched by the panel's interchanges
};

// Column 8W + U of the leaf: thread (row r, quarter q) holds columns
// 8q..8q+7 of data row r in x; the four waves of quarter W hold the column.
// Those waves form the argmax keys of the rows at positions >= d (a wave
// maximum by DPP, one key per wave to LDS) and put the column's values in
// LDS by data row; after the first barrier every thread takes the pivot,
// the pivot row's threads put it in LDS, and after the second the rows
// below scale column d and update their later columns.  No global memory
// is touched (a barrier would wait for it); the pivots collect in LDS.
// The slots alternate with the column's parity.
template <int U>
__device__ __forceinline__ void leaf_col(c2 (&x)[8], int &pos, int &nmv, bool act, int q, int r, int t, int lane,
                                         int wv, int c0, int W, int *flags, int piv_off, DiagShared &sh)
{
  const int C = 8 * W + U, d = c0 + C, b = U & 1;
  if (q == W) {
    unsigned long long kk = 0ull;
    if (act && pos >= d)
      kk = ((unsigned long long)__float_as_uint(fabsf(x[U].re) + fabsf(x[U].im)) << 32) | (unsigned)~pos;
    kk = wave_max_key(kk);
    if (lane == 0) sh.wkey[b][wv & 3] = kk;
    sh.colv[b][r] = x[U];
  }
  __syncthreads();
  unsigned long long kk = sh.wkey[b][0];
#pragma unroll
  for (int i = 1; i < 4; i++) kk = sh.wkey[b][i] > kk ? sh.wkey[b][i] : kk;
  const float best = __uint_as_float((unsigned)(kk >> 32));
  const int bp = (int)~(unsigned)kk;
  if (t == 0) {
    sh.piv[C] = bp;
    if (best > 0.0f && bp != d) {
      sh.mv[nmv] = d;
      sh.mv[nmv + 1] = bp;
      nmv += 2;
    }
  }
  if (best > 0.0f) {
    if (act && pos == bp) {
#pragma unroll
      for (int j = 0; j < 8; j++) sh.prow[b][8 * q + j] = x[j];
      pos = d;
    } else if (act && pos == d) {
      pos = bp;
    }
    __syncthreads();
    if (q >= W && act && pos > d) {
      const c2 cand = sh.colv[b][r];
      const c2 dv = sh.prow[b][C];
      const float rr = 1.0f / (dv.re * dv.re + dv.im * dv.im);
      const c2 inv = {dv.re * rr, -dv.im * rr};
      const c2 l = {cand.re * inv.re - cand.im * inv.im, cand.re * inv.im + cand.im * inv.re};
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const c2 u = sh.prow[b][8 * q + j];
        const c2 y = {fmaf(l.im, u.im, fmaf(-l.re, u.re, x[j].re)), fmaf(-l.im, u.re, fmaf(-l.re, u.im, x[j].im))};
        if (q > W || j > U) x[j] = y;
        else if (j == U) x[j] = l;
      }
    }
  } else {
    if (t == 0 && flags[1] == 0) flags[1] = piv_off + d + 1;
    __syncthreads();
  }
}

__global__ __launch_bounds__(DL_NT) void k_diag_lu(c2 *A, int64_t lda, int jb, int *ipiv, int piv_off, int *flags)
{
  __shared__ DiagShared sh;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  for (int c0 = 0; c0 < jb; c0 += LEAF_W) {
    const int pw = min(LEAF_W, jb - c0);
    // leaf: thread (row r = t % 256, quarter q = t / 256) holds columns
    // 8q..8q+7 of data row r; windows of eight columns, one quarter at a time
    {
      const int r = t & (DL_M - 1), q = t >> 8;
      const bool act = r >= c0 && r < jb;
      int pos = r, nmv = 0;          // nmv: thread 0's count of the moved list
      c2 x[8];
#pragma unroll
      for (int j = 0; j < 8; j++)
        x[j] = (act && 8 * q + j < pw) ? A[(size_t)(c0 + 8 * q + j) * lda + r] : c2{0.0f, 0.0f};
#pragma unroll 1
      for (int W = 0; 8 * W < pw; W++) {
        const int left = pw - 8 * W;
        leaf_col<0>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 1) leaf_col<1>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 2) leaf_col<2>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 3) leaf_col<3>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 4) leaf_col<4>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 5) leaf_col<5>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 6) leaf_col<6>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
        if (left > 7) leaf_col<7>(x, pos, nmv, act, q, r, t, lane, wv, c0, W, flags, piv_off, sh);
      }
      // every row to its final position, in A and in LDS; the pivots out
      if (act) {
        if (q == 0) sh.perm[pos] = r;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          sh.L[pos - c0][8 * q + j] = x[j];
          if (8 * q + j < pw) A[(size_t)(c0 + 8 * q + j) * lda + pos] = x[j];
        }
      }
      if (t < pw) ipiv[c0 + t] = piv_off + sh.piv[t];
      if (t == 0) sh.nmv = nmv;
    }
    __syncthreads();
    // swaps: each position in the moved list (both rows of every interchange,
    // repeats harmless) takes row perm[r]'s data in the columns outside the
    // panel (cc < c0 left of it, cc >= c0 right of it); lane = list entry,
    // wave = column group, so that one load touches one column
    {
      const int li = lane, cg = wv, ncc = jb - pw;
      const int r = li < sh.nmv ? sh.mv[li] : 0;
      const int src = li < sh.nmv ? sh.perm[r] : 0;
      const bool mov = li < sh.nmv && src != r;
      c2 v[14];
#pragma unroll
      for (int i = 0; i < 14; i++) {
        const int cc = cg + 16 * i;
        const int col = cc < c0 ? cc : cc + pw;
        v[i] = (mov && cc < ncc) ? A[(size_t)col * lda + src] : c2{0.0f, 0.0f};
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 14; i++) {
        const int cc = cg + 16 * i;
        const int col = cc < c0 ? cc : cc + pw;
        if (mov && cc < ncc) A[(size_t)col * lda + r] = v[i];
      }
    }
    __syncthreads();
    const int nr = jb - c0 - LEAF_W;  // rows and columns of A22 (a full panel when nr > 0)
    if (nr <= 0) break;
    // U12 = L11^-1 A12: four threads per column, thread q (= wave / 4, so
    // that the branches and L's reads are uniform per wave) holding rows
    // 8q..8q+7; block q solved by its thread once blocks < q have been
    // subtracted, then handed to the threads below through LDS
    {
      const int cl = t & (DL_M - 1), q = t >> 8;
      const bool on = cl < nr;
      c2 *col = A + (size_t)(c0 + LEAF_W + cl) * lda + c0 + 8 * q;
      c2 a[8];
#pragma unroll
      for (int i = 0; i < 8; i++) a[i] = on ? col[i] : c2{0.0f, 0.0f};
#pragma unroll
      for (int b = 0; b < 4; b++) {
        if (on && q == b) {
#pragma unroll
          for (int i = 1; i < 8; i++)
#pragma unroll
            for (int k = 0; k < i; k++) {
              const c2 l = sh.L[8 * b + i][8 * b + k];
              a[i].re -= l.re * a[k].re - l.im * a[k].im;
              a[i].im -= l.re * a[k].im + l.im * a[k].re;
            }
#pragma unroll
          for (int i = 0; i < 8; i++) sh.U[8 * b + i][cl] = a[i];
        }
        __syncthreads();
        if (on && q > b) {
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const c2 u = sh.U[8 * b + k][cl];
#pragma unroll
            for (int i = 0; i < 8; i++) {
              const c2 l = sh.L[8 * q + i][8 * b + k];
              a[i].re -= l.re * u.re - l.im * u.im;
              a[i].im -= l.re * u.im + l.im * u.re;
            }
          }
        }
      }
      if (on)
#pragma unroll
        for (int i = 0; i < 8; i++) col[i] = a[i];
    }
    __syncthreads();
    // L21 and U12 to split FP16 planes (each row of L21 / column of U12
    // scaled by a power of two that brings its largest magnitude into
    // [2^14, 2^15), x = hi + lo), written over their FP32 copies: four
    // threads per row / column, eight k each, the maximum over the four by
    // DPP within the quad
    {
      const int rr = t >> 2, kc = t & 3;
      const bool lon = rr < nr;
      c2 lv[8], uv[8];
      float lm = 0.0f, um = 0.0f;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        lv[i] = lon ? sh.L[LEAF_W + rr][8 * kc + i] : c2{0.0f, 0.0f};
        uv[i] = lon ? sh.U[8 * kc + i][rr] : c2{0.0f, 0.0f};
        lm = fmaxf(lm, fmaxf(fabsf(lv[i].re), fabsf(lv[i].im)));
        um = fmaxf(um, fmaxf(fabsf(uv[i].re), fabsf(uv[i].im)));
      }
      lm = fmaxf(lm, __shfl_xor(lm, 1));
      lm = fmaxf(lm, __shfl_xor(lm, 2));
      um = fmaxf(um, __shfl_xor(um, 1));
      um = fmaxf(um, __shfl_xor(um, 2));
      const int le = split_exp(lm), ue = split_exp(um);
      __syncthreads();
      if (rr < DL_M - LEAF_W) {
        const float ls = ldexpf(1.0f, le), us = ldexpf(1.0f, ue);
        if (kc == 0) {
          sh.rfs[rr] = ldexpf(1.0f, -le);
          sh.cfs[rr] = ldexpf(1.0f, -ue);
        }
        h8 l0, l1, l2, l3, u0, u1, u2, u3;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          _Float16 h, o;
          split_one(lv[i].re * ls, h, o); l0[i] = h; l1[i] = o;
          split_one(lv[i].im * ls, h, o); l2[i] = h; l3[i] = o;
          split_one(uv[i].re * us, h, o); u0[i] = h; u1[i] = o;
          split_one(uv[i].im * us, h, o); u2[i] = h; u3[i] = o;
        }
        const int pos = 8 * (kc ^ diag_swz(rr));
        *(h8 *)&sh.LP[0][rr][pos] = l0;
        *(h8 *)&sh.LP[1][rr][pos] = l1;
        *(h8 *)&sh.LP[2][rr][pos] = l2;
        *(h8 *)&sh.LP[3][rr][pos] = l3;
        *(h8 *)&sh.UP[0][rr][pos] = u0;
        *(h8 *)&sh.UP[1][rr][pos] = u1;
        *(h8 *)&sh.UP[2][rr][pos] = u2;
        *(h8 *)&sh.UP[3][rr][pos] = u3;
      }
    }
    __syncthreads();
    // A22 -= L21 U12 in 16 x 16 tiles, each k = 32 on twelve
    // v_mfma_f32_16x16x32_f16 (hi*hi + hi*lo + lo*hi per real product), two
    // tiles of a wave at a time: their loads of A22 first, then the products
    const int nt = (nr + 15) / 16;
    for (int tile0 = wv * 2; tile0 < nt * nt; tile0 += DL_NW * 2) {
      c2 cv[2][4];
#pragma unroll
      for (int g = 0; g < 2; g++) {
        const int tile = tile0 + g, tr = tile % nt, tc = tile / nt;
        const int row = tr * 16 + l16;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int cl = tc * 16 + 4 * lq + v;
          cv[g][v] = (tile < nt * nt && row < nr && cl < nr)
                         ? A[(size_t)(c0 + LEAF_W + cl) * lda + c0 + LEAF_W + row] : c2{0.0f, 0.0f};
        }
      }
#pragma unroll
      for (int g = 0; g < 2; g++) {
        const int tile = tile0 + g, tr = tile % nt, tc = tile / nt;
        if (tile >= nt * nt) break;
        const int row = tr * 16 + l16, cu = tc * 16 + l16;   // this lane's L21 row and U12 column
        h8 a[4], bb[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          a[q] = *(const h8 *)&sh.LP[q][row][8 * (lq ^ diag_swz(row))];
          bb[q] = *(const h8 *)&sh.UP[q][cu][8 * (lq ^ diag_swz(cu))];
        }
        const h8 n2 = hneg(a[2]), n3 = hneg(a[3]);
        f4 pr = {0.0f, 0.0f, 0.0f, 0.0f}, pi = {0.0f, 0.0f, 0.0f, 0.0f};
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[0], a[0], pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[2], a[0], pi, 0, 0, 0);
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[1], a[0], pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[3], a[0], pi, 0, 0, 0);
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[0], a[1], pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[2], a[1], pi, 0, 0, 0);
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[2], n2, pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[0], a[2], pi, 0, 0, 0);
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[3], n2, pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[1], a[2], pi, 0, 0, 0);
        pr = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[2], n3, pr, 0, 0, 0);
        pi = __builtin_amdgcn_mfma_f32_16x16x32_f16(bb[0], a[3], pi, 0, 0, 0);
        const float rf = row < nr ? sh.rfs[row] : 0.0f;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int cl = tc * 16 + 4 * lq + v;
          if (row < nr && cl < nr) {
            const float f = rf * sh.cfs[cl];
            A[(size_t)(c0 + LEAF_W + cl) * lda + c0 + LEAF_W + row] = {cv[g][v].re - pr[v] * f,
                                                                         cv[g][v].im - pi[v] * f};
          }
        }
      }
    }
    __syncthreads();
  }
}

// X = I (n x n, leading dimension n)
__global__ void k_eye(c2 *X, int n)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r < n) X[(size_t)c * n + r] = {r == c ? 1.0f : 0.0f, 0.0f};
}

// Row interchanges of one block (rows k0..k0+kb, kb <= LB, every pivot
// inside the block, as the diagonal-block kernel chooses them) in columns
// c0..c1.  LAPACK's sequence of interchanges is composed into one
// permutation of the block's rows -- thread 0 walks only the rows that
// were interchanged, found by ballots -- and each workgroup then moves the
// rows that changed place in its 32 columns: all loads, a barrier, all
// stores.  The columns of one workgroup are its own, so nothing crosses
// workgroups.
#define LSW_COLS 32
__global__ __launch_bounds__(256) void k_laswp_local(c2 *A, int64_t lda, int c0, int c1, int k0, int kb,
                                                     const int *__restrict__ ipiv)
{
  __shared__ int perm[LB], pv[LB];
  __shared__ unsigned long long mask[LB / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int p = t < kb ? ipiv[k0 + t] - k0 : t;
  if (p < 0 || p >= kb) p = t;                          // never for the block-local pivots
  perm[t] = t;
  pv[t] = p;
  const unsigned long long b = __ballot(p != t);
  if (lane == 0) mask[w] = b;
  __syncthreads();
  if (t == 0) {
    for (int q = 0; q < LB / 64; q++) {
      unsigned long long m = mask[q];
      while (m) {
        const int i = q * 64 + __builtin_ctzll(m);
        m &= m - 1;
        const int pp = pv[i], x = perm[i];
        perm[i] = perm[pp];
        perm[pp] = x;
      }
    }
  }
  __syncthreads();
  const int src = perm[t];
  const bool mov = t < kb && src != t;
  const int cb = c0 + blockIdx.x * LSW_COLS;
  c2 v[LSW_COLS];
#pragma unroll
  for (int i = 0; i < LSW_COLS; i++)
    v[i] = (mov && cb + i < c1) ? A[(size_t)(cb + i) * lda + k0 + src] : c2{0.0f, 0.0f};
  __syncthreads();
#pragma unroll
  for (int i = 0; i < LSW_COLS; i++)
    if (mov && cb + i < c1) A[(size_t)(cb + i) * lda + k0 + t] = v[i];
}

void laswp_local(c2 *A, int64_t lda, int c0, int c1, int k0, int kb, const int *ipiv, hipStream_t st)
{
  if (c1 <= c0 || kb <= 0) return;
  hipLaunchKernelGGL(k_laswp_local, dim3((unsigned)((c1 - c0 + LSW_COLS - 1) / LSW_COLS)), dim3(256), 0, st, A, lda,
                     c0, c1, k0, kb, ipiv);
}

// ---------------------------------------------------------------------------
// Split planes and scale factors of one product's operands
struct SplitWS {
  _Float16 *PA = nullptr, *PB = nullptr;
  float *rf = nullptr, *cf = nullptr;
  int64_t mp = 0;                  // rows of each plane (padded to the GEMM tile)
};

// split planes of L11^-1 (the A operand of U12 = L11^-1 A12), two of them
// so that the look-ahead streams fill one while the caller's stream reads
// the other
struct LinvPlanes {
  _Float16 *P = nullptr;
  float *rf = nullptr;
};

// per-device workspace: two sets of split planes (the trailing update's on
// the caller's stream, the block column's L21 product on the look-ahead
// stream), U11^-1 and L11^-1, L11^-1's planes, the two look-ahead streams
// and their events
struct LocalWork {
  void *mem = nullptr;
  size_t bytes = 0;
  hipStream_t side = nullptr, side2 = nullptr;
  hipEvent_t ev[12] = {};
};
LocalWork g_lw[64];

struct LocalPtrs {
  SplitWS bulk, pan;
  SplitWS small;                   // planes of <= 256 x 256 operands (mp = LB)
  c2 *Uinv = nullptr, *Linv = nullptr;
  LinvPlanes lp[2];
  LinvPlanes up[2];                // U11^-1's column planes (the B operand of L21)
  LocalWork *lw = nullptr;
};

size_t al256(size_t x) { return (x + 255) / 256 * 256; }

int local_work(int n, LocalPtrs *o)
{
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(BFA_ERR_ARG, "device index %d unsupported", dev);
  const int64_t p = ((int64_t)n + HT - 1) / HT * HT;
  // the caller's planes hold up to 2 LB of k (two block steps' update in one
  // product), the side's LB
  const size_t planes = al256((size_t)4 * p * LB * sizeof(_Float16)), fac = al256((size_t)p * 4);
  const size_t set = 2 * planes + 2 * fac, bset = 4 * planes + 2 * fac, tri = al256((size_t)LB * LB * sizeof(c2));
  const size_t lpl = al256((size_t)4 * LB * LB * sizeof(_Float16)) + al256((size_t)LB * 4);
  const size_t sset = 2 * lpl;
  const size_t need = bset + set + 2 * tri + 4 * lpl + sset;
  LocalWork &w = g_lw[dev];
  if (w.bytes < need) {
    if (w.mem) {
      HIPCHECK(hipDeviceSynchronize());
      HIPCHECK(hipFree(w.mem));
      w.mem = nullptr;
      w.bytes = 0;
    }
    HIPCHECK(hipMalloc(&w.mem, need));
    w.bytes = need;
  }
  if (!w.side) {
    HIPCHECK(hipStreamCreateWithFlags(&w.side, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&w.side2, hipStreamNonBlocking));
    for (auto &e : w.ev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  char *b = (char *)w.mem;
  o->bulk.PA = (_Float16 *)b;
  o->bulk.PB = (_Float16 *)(b + 2 * planes);
  o->bulk.rf = (float *)(b + 4 * planes);
  o->bulk.cf = (float *)(b + 4 * planes + fac);
  o->bulk.mp = p;
  b += bset;
  o->pan.PA = (_Float16 *)b;
  o->pan.PB = (_Float16 *)(b + planes);
  o->pan.rf = (float *)(b + 2 * planes);
  o->pan.cf = (float *)(b + 2 * planes + fac);
  o->pan.mp = p;
  b += set;
  o->Uinv = (c2 *)b;
  o->Linv = (c2 *)(b + tri);
  b += 2 * tri;
  for (LinvPlanes *l : {&o->lp[0], &o->lp[1], &o->up[0], &o->up[1]}) {
    l->P = (_Float16 *)b;
    l->rf = (float *)(b + al256((size_t)4 * LB * LB * sizeof(_Float16)));
    b += lpl;
  }
  o->small.PA = (_Float16 *)b;
  o->small.rf = (float *)(b + al256((size_t)4 * LB * LB * sizeof(_Float16)));
  o->small.PB = (_Float16 *)(b + lpl);
  o->small.cf = (float *)(b + lpl + al256((size_t)4 * LB * LB * sizeof(_Float16)));
  o->small.mp = LB;
  o->lw = &w;
  return BFA_OK;
}

// HIP events around the split GEMM launches while bfa_profile_enable is on
struct SplitProfile {
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rec;
  double flops = 0.0;
  double bytes = 0.0;       // algorithmic: A and B planes once, C read + written (or written when stored)
  size_t used = 0;
};
SplitProfile g_sp;
double g_sp_last_bytes = 0.0;
int g_sp_last_launches = 0;

hipEvent_t sp_event(hipStream_t st)
{
  if (g_sp.used >= g_sp.pool.size()) {
    for (int i = 0; i < 256; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) == hipSuccess) g_sp.pool.push_back(e);
    }
  }
  if (g_sp.used >= g_sp.pool.size()) return nullptr;
  hipEvent_t e = g_sp.pool[g_sp.used++];
  (void)hipEventRecord(e, st);
  return e;
}

// super-tile edge G (4, or smaller while fewer than 16 super-tiles would
// leave XCDs idle) and the padded grid of super_tile().  Measured at
// 19,616 x 19,360 x 256: G = 1, 2 or 4 2.86 ms, G = 8 3.03 ms, G = 16
// 3.16 ms (an 8 x 8 super-tile's operand slabs fill an XCD's 4 MB L2); at
// k = 512, alone: G = 1, 2, 4, 8 4.43, 4.28, 4.37, 4.72 ms, but G = 2
// fetched 10.1 GB against 7.8 GB for G = 4 and was no faster inside the
// factorisation, where other kernels share the bandwidth.
void super_grid(int tm, int tn, int *G, size_t *nblk)
{
  int g = 4;
  while (g > 1 && (size_t)((tm + g - 1) / g) * ((tn + g - 1) / g) < 16) g /= 2;
  const size_t ns = (size_t)((tm + g - 1) / g) * ((tn + g - 1) / g);
  *G = g;
  *nblk = (ns + 7) / 8 * 8 * g * g;
}

// split planes of A (m x k) by rows and B (k x n) by columns, k <= 512 (the
// caller's planes; the side's hold k <= 256)
void split_ops(int m, int n, int k, const c2 *A, int64_t lda, const c2 *B, int64_t ldb, const SplitWS &w,
               hipStream_t st)
{
  const int kp = (k + HK - 1) / HK * HK;
  const dim3 grid((unsigned)(w.mp / 32));
  if (k <= LB) {
    if (m > 0) hipLaunchKernelGGL(k_split_rows<8>, grid, dim3(256), 0, st, A, lda, m, k, kp, w.mp, w.PA, w.rf);
    if (n > 0) hipLaunchKernelGGL(k_split_cols<8>, grid, dim3(256), 0, st, B, ldb, k, n, kp, w.mp, w.PB, w.cf);
  } else {
    if (m > 0) hipLaunchKernelGGL(k_split_rows<16>, grid, dim3(512), 0, st, A, lda, m, k, kp, w.mp, w.PA, w.rf);
    if (n > 0) hipLaunchKernelGGL(k_split_cols<16>, grid, dim3(512), 0, st, B, ldb, k, n, kp, w.mp, w.PB, w.cf);
  }
}

// C(m x n) -= A B (store: C = A B) from split planes: A's planes PA (rows
// padded to mpA) with row factors rf, B's planes PB (columns padded to npB)
// with column factors cf, the product's columns starting at column c0 of B
void gemm_ops(int m, int n, int k, c2 *C, int64_t ldc, const _Float16 *PA, const float *rf, int64_t mpA,
              const _Float16 *PB, const float *cf, int64_t npB, int c0, hipStream_t st, int store)
{
  if (m <= 0 || n <= 0 || k <= 0) return;
  const int kp = (k + HK - 1) / HK * HK;
  const int tm = (m + HT - 1) / HT, tn = (n + HT - 1) / HT;
  const bool prof = bfa::profile_on();
  hipEvent_t e0 = prof ? sp_event(st) : nullptr;
  int G = 1;
  size_t nblk = 0;
  super_grid(tm, tn, &G, &nblk);
  hipLaunchKernelGGL(k_gemm_hs, dim3((unsigned)nblk), dim3(512), 0, st, m, n, kp, PA, mpA, PB + (size_t)c0 * kp, npB,
                     rf, cf + c0, C, ldc, tm, tn, store, G);
  if (prof) {
    hipEvent_t e1 = sp_event(st);
    if (e0 && e1) {
      g_sp.rec.push_back({e0, e1});
      g_sp.flops += 8.0 * m * (double)n * k;
      g_sp.bytes += 8.0 * k * ((double)m + n) + (store ? 8.0 : 16.0) * m * (double)n;
    }
  }
}

// the same with both operands in one set of planes
void gemm_planes(int m, int n, int k, c2 *C, int64_t ldc, const SplitWS &w, int c0, hipStream_t st, int store)
{
  gemm_ops(m, n, k, C, ldc, w.PA, w.rf, w.mp, w.PB, w.cf, w.mp, c0, st, store);
}

// step j's block column on the look-ahead stream st: the jb x jb diagonal
// block with partial pivoting over its rows, then L21 = A21 U11^-1 (U11^-1
// by the one-launch upper TRSM on the identity, the product on the matrix
// cores stored over A21).  Beside it, on st2 once the diagonal block is
// done: A21's planes, then L11^-1 (the unit-lower TRSM on the identity)
// split into planes lp, the A operand of the caller's U12 = L11^-1 A12.
// ev_diag and ev_linv are scratch events; st waits for st2 at the end.
int block_column(c2 *A, int n, int64_t lda, int j, int jb, int *ipiv, int *flags, const LocalPtrs &P,
                 const LinvPlanes &lp, hipStream_t st, hipStream_t st2, hipEvent_t ev_diag, hipEvent_t ev_linv)
{
  c2 *Ajj = A + (size_t)j * lda + j;
  hipLaunchKernelGGL(k_diag_lu, dim3(1), dim3(DL_NT), 0, st, Ajj, lda, jb, ipiv + j, j, flags);
  const int below = n - j - jb;
  if (below <= 0) return BFA_OK;
  HIPCHECK(hipEventRecord(ev_diag, st));
  HIPCHECK(hipStreamWaitEvent(st2, ev_diag, 0));
  const int kp = (jb + HK - 1) / HK * HK;
  // st2: A21's planes (they need only the diagonal block's row swaps, which
  // do not reach below it), then L11^-1 and its planes
  hipLaunchKernelGGL(k_split_rows<8>, dim3((unsigned)(P.pan.mp / 32)), dim3(256), 0, st2, (const c2 *)(Ajj + jb), lda,
                     below, jb, kp, P.pan.mp, P.pan.PA, P.pan.rf);
  HIPCHECK(hipEventRecord(ev_diag, st2));                 // A21's planes done
  hipLaunchKernelGGL(k_eye, dim3((jb + 255) / 256, jb), dim3(256), 0, st2, P.Linv, jb);
  bfa::trsm_lower_c(jb, jb, Ajj, lda, P.Linv, jb, st2);
  hipLaunchKernelGGL(k_split_rows<8>, dim3(LB / 32), dim3(256), 0, st2, (const c2 *)P.Linv, (int64_t)jb, jb, jb,
                     kp, (int64_t)LB, lp.P, lp.rf);
  HIPCHECK(hipEventRecord(ev_linv, st2));
  // st: U11^-1, its planes, L21 = A21 U11^-1 over A21 once st2's planes of A21 exist
  hipLaunchKernelGGL(k_eye, dim3((jb + 255) / 256, jb), dim3(256), 0, st, P.Uinv, jb);
  bfa::trsm_upper_c(jb, jb, Ajj, lda, P.Uinv, jb, st);
  hipLaunchKernelGGL(k_split_cols<8>, dim3((unsigned)(P.pan.mp / 32)), dim3(256), 0, st, (const c2 *)P.Uinv,
                     (int64_t)jb, jb, jb, kp, P.pan.mp, P.pan.PB, P.pan.cf);
  HIPCHECK(hipStreamWaitEvent(st, ev_diag, 0));
  gemm_planes(below, jb, jb, Ajj + jb, lda, P.pan, 0, st, 1);
  HIPCHECK(hipStreamWaitEvent(st, ev_linv, 0));
  return BFA_OK;
}

// Grouped forms over the problems of a batch (descriptors sorted by blk0,
// each workgroup's problem found by binary search): the split planes and
// the split-FP16 product of every instance's trailing update in one launch
// each.
template <typename Q> __device__ __forceinline__ int hs_find(const Q *__restrict__ d, int np, int bid)
{
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].blk0 <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool COLS>
__global__ __launch_bounds__(256) void kg_split(const HsSplitProb *__restrict__ d, int np)
{
  __shared__ float smax[8][32];
  const HsSplitProb q = d[hs_find(d, np, blockIdx.x)];
  const int b = blockIdx.x - q.blk0;
  if (COLS)
    split_cols_body<8>((const c2 *)q.X, q.ldx, q.k, q.nv, q.kp, q.pr, (_Float16 *)q.P, (float *)q.fac, b);
  else
    split_rows_body<8>((const c2 *)q.X, q.ldx, q.nv, q.k, q.kp, q.pr, (_Float16 *)q.P, (float *)q.fac, b, smax);
}

__global__ __launch_bounds__(512, 1) void kg_gemm_hs(const HsGemmProb *__restrict__ d, int np)
{
  __shared__ __align__(16) _Float16 As[2][4][HT][HK];
  __shared__ __align__(16) _Float16 Bs[2][4][HT][HK];
  const HsGemmProb q = d[hs_find(d, np, blockIdx.x)];
  int ti, tj;
  if (!super_tile(blockIdx.x - q.blk0, q.tiles_m, q.tiles_n, q.G, ti, tj)) return;
  gemm_hs_tile<8>(q.m, q.n, q.kp, (const _Float16 *)q.PA, q.mp, (const _Float16 *)q.PB, q.np, (const float *)q.rf,
                  (const float *)q.cf, (c2 *)q.C, q.ldc, ti * HT, tj * HT, 0, As, Bs);
}

}  // namespace

void bfa::launch_split_g(bool cols, const HsSplitProb *d, int np, int nblk, hipStream_t st)
{
  if (np <= 0 || nblk <= 0) return;
  if (cols) hipLaunchKernelGGL(kg_split<true>, dim3(nblk), dim3(256), 0, st, d, np);
  else hipLaunchKernelGGL(kg_split<false>, dim3(nblk), dim3(256), 0, st, d, np);
}

void bfa::launch_gemm_hs_g(const HsGemmProb *d, int np, int nblk, hipStream_t st)
{
  if (np > 0 && nblk > 0) hipLaunchKernelGGL(kg_gemm_hs, dim3(nblk), dim3(512), 0, st, d, np);
}

void bfa::hs_grid(int m, int n, int *G, int *nblk)
{
  size_t nb = 0;
  super_grid((m + HT - 1) / HT, (n + HT - 1) / HT, G, &nb);
  *nblk = (int)nb;
}

void bfa::local_release()
{
  for (int d = 0; d < 64; d++) {
    LocalWork &w = g_lw[d];
    if (!w.mem && !w.side) continue;
    if (hipSetDevice(d) == hipSuccess) {
      (void)hipDeviceSynchronize();
      if (w.mem) (void)hipFree(w.mem);
      if (w.side) (void)hipStreamDestroy(w.side);
      if (w.side2) (void)hipStreamDestroy(w.side2);
      for (auto &e : w.ev) if (e) (void)hipEventDestroy(e);
    }
    w = LocalWork();
  }
  for (hipEvent_t e : g_sp.pool) (void)hipEventDestroy(e);
  g_sp = SplitProfile();
}

int bfa::profile_read_split(double *ms, double *flops, int *launches)
{
  double t = 0.0;
  for (auto &pr : g_sp.rec) {
    HIPCHECK(hipEventSynchronize(pr.second));
    float x = 0.0f;
    HIPCHECK(hipEventElapsedTime(&x, pr.first, pr.second));
    t += x;
  }
  if (ms) *ms = t;
  if (flops) *flops = g_sp.flops;
  if (launches) *launches = (int)g_sp.rec.size();
  g_sp_last_bytes = g_sp.bytes;
  g_sp_last_launches = (int)g_sp.rec.size();
  g_sp.rec.clear();
  g_sp.flops = 0.0;
  g_sp.bytes = 0.0;
  g_sp.used = 0;
  return BFA_OK;
}

int bfa::profile_split_bytes(double *bytes, int *launches)
{
  if (bytes) *bytes = g_sp_last_bytes;
  if (launches) *launches = g_sp_last_launches;
  return BFA_OK;
}

// ---------------------------------------------------------------------------
extern "C" int bfa_cgemm_sub_split(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B, int64_t ldb,
                                   void *d_C, int64_t ldc, void *stream)
{
  if (m < 0 || n < 0 || k < 0 || k > 2 * LB || (m > 0 && (lda < m || ldc < m)) || (k > 0 && ldb < k))
    return fail(BFA_ERR_ARG, "bfa_cgemm_sub_split: bad sizes (k <= %d)", 2 * LB);
  if (m == 0 || n == 0 || k == 0) return BFA_OK;
  LocalPtrs P;
  int rc = local_work(m > n ? m : n, &P);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  split_ops(m, n, k, (const c2 *)d_A, lda, (const c2 *)d_B, ldb, P.bulk, st);
  gemm_planes(m, n, k, (c2 *)d_C, ldc, P.bulk, 0, st, 0);
  HIPCHECK(hipGetLastError());
  return BFA_OK;
}

namespace {

// Schedule 1 (look-ahead): the caller's stream st carries, per step, the row
// swaps, U12 = L11^-1 A12 (A12 split, the product stored over A12), the
// splits and the trailing update -- the next block column first; the side
// streams then factor that block column (diagonal block, L21, L11^-1's
// planes) while st updates the remaining columns.  They touch disjoint
// columns (the side streams only the next block column, st everything right
// of it) and each has its own split planes; L11^-1's planes alternate
// between two sets.
int sched_lookahead(c2 *A, int n, int64_t lda, int *d_ipiv, int *flags, const LocalPtrs &P, hipStream_t st)
{
  int rc = BFA_OK;
  LocalWork *lw = P.lw;
  hipStream_t sp = lw->side, sp2 = lw->side2;
  hipEvent_t ev_in = lw->ev[0], ev_col = lw->ev[1], ev_next = lw->ev[2], ev_diag = lw->ev[3], ev_linv = lw->ev[4];
  // the side streams start after everything queued on st (the matrix itself)
  HIPCHECK(hipEventRecord(ev_in, st));
  HIPCHECK(hipStreamWaitEvent(sp, ev_in, 0));
  HIPCHECK(hipStreamWaitEvent(sp2, ev_in, 0));
  rc = block_column(A, n, lda, 0, n < LB ? n : LB, d_ipiv, flags, P, P.lp[0], sp, sp2, ev_diag, ev_linv);
  if (rc) return rc;
  HIPCHECK(hipEventRecord(ev_col, sp));
  for (int j = 0; j < n; j += LB) {
    const int jb = n - j < LB ? n - j : LB;
    HIPCHECK(hipStreamWaitEvent(st, ev_col, 0));         // block column j factored
    const int rest = n - j - jb;
    if (rest <= 0) {
      laswp_local(A, lda, 0, j, j, jb, d_ipiv, st);
      break;
    }
    laswp_local(A, lda, j + jb, n, j, jb, d_ipiv, st);
    c2 *A12 = A + (size_t)(j + jb) * lda + j;
    const LinvPlanes &lp = P.lp[(j / LB) & 1];
    const int kp = (jb + HK - 1) / HK * HK;
    hipLaunchKernelGGL(k_split_cols<8>, dim3((unsigned)(P.bulk.mp / 32)), dim3(256), 0, st, (const c2 *)A12, lda, jb,
                       rest, kp, P.bulk.mp, P.bulk.PB, P.bulk.cf);
    gemm_ops(jb, rest, jb, A12, lda, lp.P, lp.rf, LB, P.bulk.PB, P.bulk.cf, P.bulk.mp, 0, st, 1);   // U12
    const c2 *L21 = A + (size_t)j * lda + j + jb;
    split_ops(rest, rest, jb, L21, lda, A12, lda, P.bulk, st);
    const int nb2 = rest < LB ? rest : LB;
    gemm_planes(rest, nb2, jb, A12 + jb, lda, P.bulk, 0, st, 0);   // the next block column
    HIPCHECK(hipEventRecord(ev_next, st));
    HIPCHECK(hipStreamWaitEvent(sp, ev_next, 0));
    rc = block_column(A, n, lda, j + jb, nb2, d_ipiv, flags, P, P.lp[((j + jb) / LB) & 1], sp, sp2, ev_diag,
                      ev_linv);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(ev_col, sp));
    // the left columns' interchanges are off the look-ahead chain
    laswp_local(A, lda, 0, j, j, jb, d_ipiv, st);
    gemm_planes(rest, rest - nb2, jb, A12 + (size_t)nb2 * lda + jb, lda, P.bulk, nb2, st, 0);
  }
  HIPCHECK(hipEventRecord(ev_col, sp));
  HIPCHECK(hipStreamWaitEvent(st, ev_col, 0));
  return BFA_OK;
}

// Schedule 2 (diagonal chain first).  The critical path of the look-ahead is
// the chain from one diagonal block's factorisation to the next's; here it
// carries only 256-row pieces.  With blocks j, j+1, j+2 (rows/columns
// starting at j, j1, j2), per step:
//   side2:  L11^-1 of block j and its row planes; block j's interchanges
//           in block column j+1 and U12 there, A(j, j1) = L11^-1 A(j, j1);
//           L21 below block j+1: A(j2.., j) = A(j2.., j) U11^-1.
//   side:   U11^-1 and its column planes; L21 of block j+1's rows,
//           A(j1, j) = A(j1, j) U11^-1 (beside side2's U12); the diagonal
//           block A(j1, j1) -= A(j1, j) A(j, j1); its factorisation.
//   caller: block j's interchanges right of block j+1 and U12 there; the
//           splits of L21 (all rows below block j) and U12 (all columns
//           right of it); block column j+1 below its diagonal block; block
//           column j+2; the left interchanges; the rest of the update.
// The side stream waits for the caller's updates of block columns j+1 and
// j+2 (events gc, nn) and the caller for the side's L21 and U12 pieces
// (u1, l1, lr).  The pieces the streams work on at the same time are
// disjoint: the side streams work in block columns j+1 (rows below block
// j+1 on side2) and j+2 or, for the interchanges, rows of block j+1; the
// caller in rows of block j and columns right of block j+2.  The planes of
// L11^-1 and U11^-1 alternate between two sets by block parity; each set's
// next writer is ordered after its last reader through these events.
int sched_chain(c2 *A, int n, int64_t lda, int *ipiv, int *flags, const LocalPtrs &P, hipStream_t st)
{
  LocalWork *lw = P.lw;
  hipStream_t sp = lw->side, sp2 = lw->side2;
  hipEvent_t ev_in = lw->ev[0], ev_diag = lw->ev[1], ev_linv = lw->ev[2], ev_uinv = lw->ev[3], ev_u1 = lw->ev[4],
             ev_lr = lw->ev[5], ev_l1 = lw->ev[6], ev_gc = lw->ev[7], ev_nn = lw->ev[8], ev_end = lw->ev[9];
  HIPCHECK(hipEventRecord(ev_in, st));
  HIPCHECK(hipStreamWaitEvent(sp, ev_in, 0));
  HIPCHECK(hipStreamWaitEvent(sp2, ev_in, 0));
  hipLaunchKernelGGL(k_diag_lu, dim3(1), dim3(DL_NT), 0, sp, A, lda, n < LB ? n : LB, ipiv, 0, flags);
  HIPCHECK(hipEventRecord(ev_diag, sp));
  const dim3 sgrid(LB / 32), bgrid((unsigned)(P.bulk.mp / 32)), pgrid((unsigned)(P.pan.mp / 32));
  static int fuse = -1;
  if (fuse < 0) {
    const char *e = getenv("BFA_LU_LOCAL_FUSE");
    fuse = !(e && strcmp(e, "0") == 0);
  }
  bool deferred = false;
  for (int j = 0; j < n; j += LB) {
    const int jb = n - j < LB ? n - j : LB;
    const int j1 = j + jb, jb1 = n - j1 < LB ? n - j1 : LB;
    if (jb1 <= 0) {                                      // the last block
      HIPCHECK(hipStreamWaitEvent(st, ev_diag, 0));
      laswp_local(A, lda, 0, j, j, jb, ipiv, st);
      break;
    }
    const int j2 = j1 + jb1, jb2 = n - j2 <= 0 ? 0 : n - j2 < LB ? n - j2 : LB;
    const int kp = (jb + HK - 1) / HK * HK;
    const LinvPlanes &lp = P.lp[(j / LB) & 1], &up = P.up[(j / LB) & 1];
    c2 *Ajj = A + (size_t)j * lda + j;
    c2 *Aj_1 = A + (size_t)j1 * lda + j;                 // block (j, j+1)
    c2 *Aj1_j = A + (size_t)j * lda + j1;                // block (j+1, j)
    c2 *Aj1_1 = A + (size_t)j1 * lda + j1;               // block (j+1, j+1)
    // side2: L11^-1 and its planes, then U12 of block column j+1 (after the
    // caller's update of that column by step j-1): block j's interchanges
    // there and A(j, j1) = L11^-1 A(j, j1)
    HIPCHECK(hipStreamWaitEvent(sp2, ev_diag, 0));
    hipLaunchKernelGGL(k_eye, dim3((jb + 255) / 256, jb), dim3(256), 0, sp2, P.Linv, jb);
    bfa::trsm_lower_c(jb, jb, Ajj, lda, P.Linv, jb, sp2);
    hipLaunchKernelGGL(k_split_rows<8>, sgrid, dim3(256), 0, sp2, (const c2 *)P.Linv, (int64_t)jb, jb, jb, kp,
                       (int64_t)LB, lp.P, lp.rf);
    HIPCHECK(hipEventRecord(ev_linv, sp2));
    if (j > 0) HIPCHECK(hipStreamWaitEvent(sp2, ev_nn, 0));
    laswp_local(A, lda, j1, j2, j, jb, ipiv, sp2);
    hipLaunchKernelGGL(k_split_cols<8>, sgrid, dim3(256), 0, sp2, (const c2 *)Aj_1, lda, jb, jb1, kp, (int64_t)LB,
                       P.small.PB, P.small.cf);
    gemm_ops(jb, jb1, jb, Aj_1, lda, lp.P, lp.rf, LB, P.small.PB, P.small.cf, LB, 0, sp2, 1);
    HIPCHECK(hipEventRecord(ev_u1, sp2));
    // side: U11^-1 and its planes
    hipLaunchKernelGGL(k_eye, dim3((jb + 255) / 256, jb), dim3(256), 0, sp, P.Uinv, jb);
    bfa::trsm_upper_c(jb, jb, Ajj, lda, P.Uinv, jb, sp);
    hipLaunchKernelGGL(k_split_cols<8>, sgrid, dim3(256), 0, sp, (const c2 *)P.Uinv, (int64_t)jb, jb, jb, kp,
                       (int64_t)LB, up.P, up.rf);
    HIPCHECK(hipEventRecord(ev_uinv, sp));
    // side2: L21 below block j+1 (after the caller's update of block column
    // j below block j by step j-1)
    const int below2 = n - j2;
    if (j > 0) HIPCHECK(hipStreamWaitEvent(sp2, ev_gc, 0));
    if (below2 > 0) {
      c2 *Aj2_j = A + (size_t)j * lda + j2;
      hipLaunchKernelGGL(k_split_rows<8>, pgrid, dim3(256), 0, sp2, (const c2 *)Aj2_j, lda, below2, jb, kp,
                         P.pan.mp, P.pan.PA, P.pan.rf);
      HIPCHECK(hipStreamWaitEvent(sp2, ev_uinv, 0));
      gemm_ops(below2, jb, jb, Aj2_j, lda, P.pan.PA, P.pan.rf, P.pan.mp, up.P, up.rf, LB, 0, sp2, 1);
    }
    HIPCHECK(hipEventRecord(ev_lr, sp2));
    // side: L21 of block j+1's rows, block j+1's diagonal block once U12 of
    // block column j+1 is there, its factorisation
    if (j > 0) HIPCHECK(hipStreamWaitEvent(sp, ev_gc, 0));
    hipLaunchKernelGGL(k_split_rows<8>, sgrid, dim3(256), 0, sp, (const c2 *)Aj1_j, lda, jb1, jb, kp, (int64_t)LB,
                       P.small.PA, P.small.rf);
    gemm_ops(jb1, jb, jb, Aj1_j, lda, P.small.PA, P.small.rf, LB, up.P, up.rf, LB, 0, sp, 1);
    HIPCHECK(hipEventRecord(ev_l1, sp));
    hipLaunchKernelGGL(k_split_rows<8>, sgrid, dim3(256), 0, sp, (const c2 *)Aj1_j, lda, jb1, jb, kp, (int64_t)LB,
                       P.small.PA, P.small.rf);
    HIPCHECK(hipStreamWaitEvent(sp, ev_u1, 0));
    hipLaunchKernelGGL(k_split_cols<8>, sgrid, dim3(256), 0, sp, (const c2 *)Aj_1, lda, jb, jb1, kp, (int64_t)LB,
                       P.small.PB, P.small.cf);
    gemm_ops(jb1, jb1, jb, Aj1_1, lda, P.small.PA, P.small.rf, LB, P.small.PB, P.small.cf, LB, 0, sp, 0);
    hipLaunchKernelGGL(k_diag_lu, dim3(1), dim3(DL_NT), 0, sp, Aj1_1, lda, jb1, ipiv + j1, j1, flags);
    HIPCHECK(hipEventRecord(ev_diag, sp));
    // caller: U12 right of block j+1
    HIPCHECK(hipStreamWaitEvent(st, ev_linv, 0));
    const int rest1 = n - j2;
    if (rest1 > 0) {
      c2 *Aj_2 = A + (size_t)j2 * lda + j;
      laswp_local(A, lda, j2, n, j, jb, ipiv, st);
      hipLaunchKernelGGL(k_split_cols<8>, bgrid, dim3(256), 0, st, (const c2 *)Aj_2, lda, jb, rest1, kp, P.bulk.mp,
                         P.bulk.PB, P.bulk.cf);
      gemm_ops(jb, rest1, jb, Aj_2, lda, lp.P, lp.rf, LB, P.bulk.PB, P.bulk.cf, P.bulk.mp, 0, st, 1);
    }
    // caller: the splits of L21 and U12, then the updates
    HIPCHECK(hipStreamWaitEvent(st, ev_u1, 0));
    HIPCHECK(hipStreamWaitEvent(st, ev_l1, 0));
    HIPCHECK(hipStreamWaitEvent(st, ev_lr, 0));
    const int below = n - j1;
    split_ops(below, below, jb, Aj1_j, lda, Aj_1, lda, P.bulk, st);
    if (below2 > 0)                                      // block column j+1 below its diagonal block
      gemm_ops(below2, jb1, jb, A + (size_t)j1 * lda + j2, lda, P.bulk.PA + (size_t)jb1 * kp, P.bulk.rf + jb1,
               P.bulk.mp, P.bulk.PB, P.bulk.cf, P.bulk.mp, 0, st, 0);
    HIPCHECK(hipEventRecord(ev_gc, st));
    // Pairs of steps share one trailing update of k = 2 LB (half the C
    // traffic per flop): the first step of a pair updates block columns
    // j+2 and j+3 and the rows of block j+1 only, and leaves rows below
    // block j+1 right of block j+3 to the second step, which updates them
    // with both steps' L21 and U12 in one product.  Nothing reads that
    // region in between: the side streams work in block columns j+1..j+3
    // and the caller's next step in rows of block j+1 and block column j+3.
    const int j3 = j2 + jb2, jb3 = n - j3 <= 0 ? 0 : n - j3 < LB ? n - j3 : LB;
    // (while the rest is wide enough that the caller's stream, not the side
    // chain, sets the pace: measured at N = 19,872, pairs with fewer than
    // 6,144 columns right of them were slower fused)
    const bool first = fuse && !deferred && jb == LB && jb1 == LB && jb2 == LB && jb3 > 0 && n - j3 - jb3 >= 24 * LB;
    const int nn = first ? jb2 + jb3 : jb2;
    if (nn > 0)                                          // block column j+2 (and j+3)
      gemm_ops(below, nn, jb, A + (size_t)j2 * lda + j1, lda, P.bulk.PA, P.bulk.rf, P.bulk.mp, P.bulk.PB,
               P.bulk.cf, P.bulk.mp, jb1, st, 0);
    HIPCHECK(hipEventRecord(ev_nn, st));
    if (j > 0) laswp_local(A, lda, 0, j, j, jb, ipiv, st);
    if (first) {                                         // rows of block j+1 right of block j+3
      const int e = j3 + jb3;
      gemm_ops(jb1, n - e, jb, A + (size_t)e * lda + j1, lda, P.bulk.PA, P.bulk.rf, P.bulk.mp, P.bulk.PB,
               P.bulk.cf, P.bulk.mp, jb1 + nn, st, 0);
      deferred = true;
      continue;
    }
    const int rest2 = n - j2 - jb2;
    if (deferred) {                                      // both steps' update below block j+1, right of j+2
      const int a = j - LB;
      split_ops(below, rest2, 2 * LB, A + (size_t)a * lda + j1, lda, A + (size_t)(j2 + jb2) * lda + a, lda, P.bulk,
                st);
      gemm_planes(below, rest2, 2 * LB, A + (size_t)(j2 + jb2) * lda + j1, lda, P.bulk, 0, st, 0);
      deferred = false;
    } else if (rest2 > 0)
      gemm_ops(below, rest2, jb, A + (size_t)(j2 + jb2) * lda + j1, lda, P.bulk.PA, P.bulk.rf, P.bulk.mp,
               P.bulk.PB, P.bulk.cf, P.bulk.mp, jb1 + jb2, st, 0);
  }
  HIPCHECK(hipEventRecord(ev_end, sp));
  HIPCHECK(hipStreamWaitEvent(st, ev_end, 0));
  HIPCHECK(hipEventRecord(ev_end, sp2));
  HIPCHECK(hipStreamWaitEvent(st, ev_end, 0));
  return BFA_OK;
}

}  // namespace

// BFA_LU_LOCAL_SCHED=lookahead selects schedule 1; the default is the
// diagonal chain (schedule 2)
extern "C" int bfa_lu_factor_c_local(void *d_A, int n, int64_t lda, int *d_ipiv, int *info, void *stream)
{
  if (!d_A || !d_ipiv || n <= 0 || lda < n) return fail(BFA_ERR_ARG, "bfa_lu_factor_c_local: bad argument");
  hipStream_t st = (hipStream_t)stream;
  c2 *A = (c2 *)d_A;
  LocalPtrs P;
  int rc = local_work(n, &P);
  if (rc) return rc;
  rc = bfa::lu_local_begin(st);
  if (rc) return rc;
  int *flags = bfa::lu_local_flags();
  if (!flags) return fail(BFA_ERR_STATE, "bfa_lu_factor_c_local: LU work area");
  static int sched = -1;
  if (sched < 0) {
    const char *e = getenv("BFA_LU_LOCAL_SCHED");
    sched = e && strcmp(e, "lookahead") == 0 ? 1 : 2;
  }
  rc = sched == 1 ? sched_lookahead(A, n, lda, d_ipiv, flags, P, st) : sched_chain(A, n, lda, d_ipiv, flags, P, st);
  if (rc) return rc;
  HIPCHECK(hipGetLastError());
  return bfa::lu_local_end(info, st);
}
