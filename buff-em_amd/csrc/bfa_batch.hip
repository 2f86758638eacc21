// bfa_batch.hip -- the batching form: many independent VIE problems of
// uneven size, swept over frequencies, in grouped launches captured as one
// hipGraph per sweep step.
//
// Per frequency (all instances at once, every launch grouped):
//   assembly   far field (every instance's tetrahedron-pair moments into
//              its own table, then the BF tiles gathering them; one pass
//              when the tables would not fit), touching pairs, RHS
//              (complex128)
//   convert    M64 -> complex64 copy padded to a multiple of 256 with an
//              identity block (the padding never pivots: its rows are zero
//              in the original columns)
//   factor     blocked right-looking LU, nb = 256, recursive panel down to
//              32-column LDS-resident panels (one grid barrier per column,
//              per instance) -- the same kernels as the single-matrix LU,
//              driven by descriptor tables
//   solve      FP64 iterative refinement: d = LU32^-1 r (gathered, complex64
//              triangular sweeps), x += d, r = b - M64 x; a fixed number of
//              rounds in which converged instances idle (device flags), so
//              the graph is static
//   power      EMT absorbed / scattered power of each instance
// The frequency lives in device memory (FreqState), written by a one-thread
// kernel before each replay.  The trailing update of each size bucket is
// one of three measured variants (bfa_batch_tune): complex64 GEMM tiles of
// 64 x 64 or 128 x 128 on FP32 MFMA, or the split-FP16 product (hi + lo
// planes per instance, then one grouped 128 x 128 FP16-MFMA launch).
//
// No reference counterpart: the reference runs one problem per process and
// one frequency at a time (src/applications/buff-scatter/buff-scatter.cc:
// 236-363); every per-instance result matches that loop's.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "bfa_group.h"
#include "bfa_internal.h"

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

#define NB 256          // LU block
#define PWID 32         // LDS-resident panel width
#define SVB 64          // triangular-solve block
#define RCH 256         // residual column chunk
#define NBUCKET 16
#define CSTRIDE 768   // u32 per panel barrier: 8 counters 128 B apart, then 8 x 32 keys

// per-instance solve state on the device
struct SolveInst {
  const z2 *M;          // complex128 matrix, N x N used, ld = ld
  const z2 *B;          // right-hand sides [nrhs][ld]
  z2 *X, *R;            // solution, residual [nrhs][ld]
  c2 *D;                // correction (complex64) [nrhs][ld]
  z2 *P;                // residual partials [nchunk][nrhs][N]
  double *norms;        // [nrhs][nblk][2]
  const int *perm;      // [Np]
  int *conv, *iters;    // flags
  double *resid;
  int64_t ld;
  int N, Np, nblk, nchunk;
  int row0, res0;       // grouped launch offsets: row blocks (256), residual blocks
  int col0, pad;        // column offsets for the padded conversion (one column per workgroup)
  c2 *LU;               // complex64 LU, Np x Np, ld = Np
  const int *ipiv;
  int *permw;           // writable perm
  int *flags;           // LU flags [2]
  c2 *Y, *Xs;           // streaming-sweep outputs (tag-filled each round)
  unsigned *tick;       // [2] sweep tickets
  unsigned long long tag;
};

__device__ __forceinline__ int sfirst(const SolveInst &q, int which)
{
  return which == 0 ? q.row0 : which == 1 ? q.res0 : q.col0;
}
__device__ __forceinline__ int find_solve(const SolveInst *__restrict__ d, int n, int bid, int which)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sfirst(d[mid], which) <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_set_state(FreqState *fs, double wre, double wim, int fidx)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) { fs->omega_re = wre; fs->omega_im = wim; fs->fidx = fidx; }
}

__global__ void k_fill_u32(unsigned *p, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// LU32(:, c) = complex64(M64(:, c)) for rows < N, identity on the padding;
// one workgroup per column of the padded matrix
__global__ __launch_bounds__(256) void kg_convert(const SolveInst *__restrict__ d, int ni)
{
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 2)];
  const int c = blockIdx.x - q.col0;
  if (c >= q.Np) return;
  c2 *out = q.LU + (size_t)c * q.Np;
  const z2 *in = q.M + (size_t)c * q.ld;
  for (int r = threadIdx.x; r < q.Np; r += blockDim.x) {
    c2 v = {r == c ? 1.0f : 0.0f, 0.0f};
    if (r < q.N && c < q.N) {
      const z2 m = in[r];
      v = {(float)m.re, (float)m.im};
    }
    out[r] = v;
  }
}

// row permutation of the whole factorisation: perm[i] = source row of row i
// after all interchanges (one workgroup per instance, sequential in LDS)
__global__ __launch_bounds__(64) void kg_perm(const SolveInst *__restrict__ d, int ni)
{
  extern __shared__ int pl[];
  const SolveInst &q = d[blockIdx.x];
  for (int i = threadIdx.x; i < q.Np; i += blockDim.x) pl[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < q.Np; i++) {
      const int p = q.ipiv[i];
      if (p != i && p >= 0 && p < q.Np) { const int t = pl[i]; pl[i] = pl[p]; pl[p] = t; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < q.Np; i += blockDim.x) q.permw[i] = pl[i];
}

// X = 0, R = B (padding rows 0), conv = iters = 0
__global__ __launch_bounds__(256) void kg_refine_init(const SolveInst *__restrict__ d, int ni, int nrhs)
{
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 0)];
  const int r = (blockIdx.x - q.row0) * blockDim.x + threadIdx.x;
  if (r == 0) { *q.conv = 0; *q.iters = 0; *q.resid = -1.0; }
  if (r >= q.Np) return;
  for (int j = 0; j < nrhs; j++) {
    q.X[(size_t)j * q.ld + r] = {0.0, 0.0};
    q.R[(size_t)j * q.ld + r] = r < q.N ? q.B[(size_t)j * q.ld + r] : z2{0.0, 0.0};
  }
}

// D(i, j) = complex64(R(perm[i], j)); the sweep outputs Y, Xs are filled
// with the tag and the sweep tickets cleared
__global__ __launch_bounds__(256) void kg_gather(const SolveInst *__restrict__ d, int ni, int nrhs)
{
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 0)];
  if (*q.conv) return;
  const int r = (blockIdx.x - q.row0) * blockDim.x + threadIdx.x;
  if (r == 0) { q.tick[0] = 0u; q.tick[1] = 0u; }
  if (r >= q.Np) return;
  const int s = q.perm[r];
  for (int j = 0; j < nrhs; j++) {
    const z2 v = q.R[(size_t)j * q.ld + s];
    q.D[(size_t)j * q.ld + r] = {(float)v.re, (float)v.im};
    ((unsigned long long *)q.Y)[(size_t)j * q.ld + r] = q.tag;
    ((unsigned long long *)q.Xs)[(size_t)j * q.ld + r] = q.tag;
  }
}

// X += (LU)^-1 D (the backward sweep's output) on rows < N
__global__ __launch_bounds__(256) void kg_add(const SolveInst *__restrict__ d, int ni, int nrhs)
{
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 0)];
  if (*q.conv) return;
  const int r = (blockIdx.x - q.row0) * blockDim.x + threadIdx.x;
  if (r >= q.N) return;
  for (int j = 0; j < nrhs; j++) {
    const c2 v = q.Xs[(size_t)j * q.ld + r];
    z2 &x = q.X[(size_t)j * q.ld + r];
    x.re += (double)v.re;
    x.im += (double)v.im;
  }
}

// P[chunk][j][r] = sum_{c in chunk} M(r, c) X(c, j): workgroup = (row block,
// column chunk) of one instance, M streamed once, X chunk in LDS
__global__ __launch_bounds__(256) void kg_residual_part(const SolveInst *__restrict__ d, int ni, int nrhs)
{
  __shared__ z2 xs[BFA_MAXPW_BATCH][RCH];
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 1)];
  if (*q.conv) return;
  const int t = blockIdx.x - q.res0;
  const int rb = t % q.nblk, ch = t / q.nblk;
  const int r = rb * 256 + threadIdx.x;
  const int c0 = ch * RCH, nc = min(RCH, q.N - c0);
  for (int idx = threadIdx.x; idx < nrhs * RCH; idx += blockDim.x) {
    const int j = idx / RCH, c = idx % RCH;
    xs[j][c] = c < nc ? q.X[(size_t)j * q.ld + c0 + c] : z2{0, 0};
  }
  __syncthreads();
  if (r >= q.N) return;
  double ar[BFA_MAXPW_BATCH], ai[BFA_MAXPW_BATCH];
#pragma unroll
  for (int j = 0; j < BFA_MAXPW_BATCH; j++) { ar[j] = 0.0; ai[j] = 0.0; }
  const z2 *col = q.M + (size_t)c0 * q.ld + r;
#pragma unroll 8
  for (int c = 0; c < nc; c++) {
    const z2 m = col[(size_t)c * q.ld];
#pragma unroll
    for (int j = 0; j < BFA_MAXPW_BATCH; j++) {
      if (j < nrhs) {
        const z2 x = xs[j][c];
        ar[j] += m.re * x.re - m.im * x.im;
        ai[j] += m.re * x.im + m.im * x.re;
      }
    }
  }
  for (int j = 0; j < nrhs; j++) q.P[((size_t)ch * nrhs + j) * q.N + r] = {ar[j], ai[j]};
}

// R = B - sum P ; per-block partial norms of R and B
__global__ __launch_bounds__(256) void kg_residual_final(const SolveInst *__restrict__ d, int ni, int nrhs)
{
  __shared__ double sr[256], sb[256];
  const SolveInst &q = d[find_solve(d, ni, blockIdx.x, 0)];
  if (*q.conv) return;
  const int rb = blockIdx.x - q.row0, j = blockIdx.y;
  if (rb >= q.nblk) return;
  const int r = rb * 256 + threadIdx.x;
  double nr = 0.0, nb = 0.0;
  if (r < q.N) {
    double re = 0.0, im = 0.0;
    for (int k = 0; k < q.nchunk; k++) {
      const z2 p = q.P[((size_t)k * nrhs + j) * q.N + r];
      re += p.re; im += p.im;
    }
    const z2 b = q.B[(size_t)j * q.ld + r];
    const z2 v = {b.re - re, b.im - im};
    q.R[(size_t)j * q.ld + r] = v;
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
    q.norms[2 * ((size_t)j * q.nblk + rb)] = sr[0];
    q.norms[2 * ((size_t)j * q.nblk + rb) + 1] = sb[0];
  }
}

// one workgroup per instance: resid = max_j ||r_j|| / ||b_j||, convergence
__global__ __launch_bounds__(64) void kg_refine_check(const SolveInst *__restrict__ d, int ni, int nrhs, double tol)
{
  const SolveInst &q = d[blockIdx.x];
  if (*q.conv || threadIdx.x != 0) return;
  double worst = 0.0;
  for (int j = 0; j < nrhs; j++) {
    double s = 0.0, b = 0.0;
    for (int k = 0; k < q.nblk; k++) { s += q.norms[2 * ((size_t)j * q.nblk + k)]; b += q.norms[2 * ((size_t)j * q.nblk + k) + 1]; }
    worst = fmax(worst, b > 0.0 ? sqrt(s / b) : sqrt(s));
  }
  *q.iters += 1;
  *q.resid = worst;
  if (worst <= tol) *q.conv = 1;
}

}  // namespace

// ===========================================================================
// Batch object
// ===========================================================================
struct bfa_batch {
  int count = 0, npw = 0, device = -1;
  std::vector<bfa_geometry *> geo;
  std::vector<bfa_material> mats;
  std::vector<int> N, Np;
  PWBatch pws{};
  // large buffers, one allocation per instance
  std::vector<void *> M64, LU32;
  // split-FP16 planes of each instance's trailing update (allocated when a
  // tuned bucket first uses them): A's planes, B's, row and column factors
  std::vector<void *> HS;
  // two-pass far field: each instance's tetrahedron-pair moment table
  // (allocated at the first plan when all of them fit in a quarter of the
  // free memory; otherwise the single-pass far field)
  std::vector<void *> TP;
  int tp_state = 0;              // 0 not tried, 1 tables, -1 single pass
  // everything else in one allocation; per-instance byte offsets
  void *small = nullptr;
  struct Off { size_t ipiv, perm, B, X, R, D, Y, Xs, inv, tick, P, norms, part, sdata, sj; };
  std::vector<Off> off;
  size_t o_fs = 0, o_conv = 0, o_iters = 0, o_resid = 0, o_counters = 0, o_flags = 0;
  size_t o_hist_it = 0, o_hist_res = 0, o_hist_info = 0;
  size_t ncounters = 0, ncounters_max = 0;
  double *power = nullptr;       // [nfreq][count][npw][2], grown on demand
  int power_nfreq = 0;
  std::vector<PanelSlots> slots;
  // grouped descriptors and the launch sequence of one frequency
  std::vector<unsigned char> hdesc;
  void *ddesc = nullptr;
  std::vector<std::function<void(hipStream_t)>> ops;
  std::vector<int> op_phase;     // 0 assembly, 1 convert+factor, 2 solve, 3 power
  // tuning: GEMM tile variant per size bucket (floor(log2 m)), -1 = default rule
  int tile[NBUCKET];
  double tune_ms[NBUCKET][NUM_TILES];
  bool planned = false;
  int plan_maxit = -1;
  double plan_tol = 0.0;
  // graph of one frequency
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr, graph_p = nullptr;
  hipGraphExec_t exec = nullptr, exec_p = nullptr;
  double phase_ms[4] = {0, 0, 0, 0};
  double run_ms = 0.0, power_ms = 0.0;

  template <typename T> T *at(size_t o) const { return (T *)((char *)small + o); }
  int *ipiv(int i) const { return at<int>(off[i].ipiv); }
  int *perm(int i) const { return at<int>(off[i].perm); }
  z2 *Bp(int i) const { return at<z2>(off[i].B); }
  z2 *Xp(int i) const { return at<z2>(off[i].X); }
  z2 *Rp(int i) const { return at<z2>(off[i].R); }
  c2 *Dptr(int i) const { return at<c2>(off[i].D); }
  FreqState *fs() const { return at<FreqState>(o_fs); }
  int *conv() const { return at<int>(o_conv); }
  int *iters() const { return at<int>(o_iters); }
  double *resid() const { return at<double>(o_resid); }
  unsigned *counters() const { return at<unsigned>(o_counters); }
  int *flags() const { return at<int>(o_flags); }
  void reset_plan()
  {
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    if (exec_p) { (void)hipGraphExecDestroy(exec_p); exec_p = nullptr; }
    if (graph_p) { (void)hipGraphDestroy(graph_p); graph_p = nullptr; }
    if (ddesc) { (void)hipFree(ddesc); ddesc = nullptr; }
    hdesc.clear();
    ops.clear();
    op_phase.clear();
    ncounters = 0;
    planned = false;
  }
  ~bfa_batch();
};

namespace {

struct Arena {
  std::vector<unsigned char> &h;
  explicit Arena(std::vector<unsigned char> &v) : h(v) {}
  template <typename T> size_t put(const std::vector<T> &v)
  {
    size_t off = (h.size() + 255) / 256 * 256;
    h.resize(off + v.size() * sizeof(T));
    if (!v.empty()) memcpy(h.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

template <typename T> const T *dp(const bfa_batch *B, size_t off) { return (const T *)((char *)B->ddesc + off); }

int bucket_of(int m)
{
  int b = 0;
  while ((2 << b) <= m && b < NBUCKET - 1) b++;
  return b;
}

int tile_for(const bfa_batch *B, int m, int n)
{
  const int b = bucket_of(std::max(1, m));
  if (B->tile[b] >= 0) return B->tile[b];
  return (m >= 256 && n >= 256) ? TILE_128 : TILE_64;
}

// the split-FP16 option of the trailing update: plane rows for an m-row
// update, the per-instance plane buffer (for Np - NB rows at most) and the
// three grouped launches' descriptors of one step
int64_t hs_rows(int m) { return ((int64_t)m + 127) / 128 * 128; }
size_t hs_bytes(int Np) { return Np > NB ? (size_t)hs_rows(Np - NB) * (4 * 2 * NB * 2 + 8) : 0; }

int hs_ensure(bfa_batch *B)
{
  if (B->HS.size() != (size_t)B->count) B->HS.assign(B->count, nullptr);
  for (int i = 0; i < B->count; i++) {
    if (B->HS[i] || hs_bytes(B->Np[i]) == 0) continue;
    HIPCHECK(hipMalloc(&B->HS[i], hs_bytes(B->Np[i])));
  }
  return BFA_OK;
}

struct HsStep {
  std::vector<HsSplitProb> rows, cols;
  std::vector<HsGemmProb> gemm;
  int brows = 0, bcols = 0, bgemm = 0;   // workgroups of each launch
  // instance i's update C -= A B with A = L21 (m x NB), B = U12 (NB x n)
  void add(const bfa_batch *B, int i, const GemmProb &g)
  {
    const int64_t pr = hs_rows(B->Np[i] - NB), mp = hs_rows(g.m), np = hs_rows(g.n);
    char *base = (char *)B->HS[i];
    void *PA = base, *PB = base + (size_t)pr * 4 * NB * 2;
    void *rf = base + (size_t)pr * 8 * NB * 2, *cf = (char *)rf + (size_t)pr * 4;
    HsSplitProb r{};
    r.X = g.A; r.ldx = g.lda; r.P = PA; r.fac = rf; r.pr = mp; r.nv = g.m; r.k = NB; r.kp = NB; r.blk0 = brows;
    rows.push_back(r);
    brows += (int)(mp / 32);
    HsSplitProb c{};
    c.X = g.B; c.ldx = g.ldb; c.P = PB; c.fac = cf; c.pr = np; c.nv = g.n; c.k = NB; c.kp = NB; c.blk0 = bcols;
    cols.push_back(c);
    bcols += (int)(np / 32);
    HsGemmProb q{};
    q.PA = PA; q.PB = PB; q.rf = rf; q.cf = cf; q.C = g.C; q.mp = mp; q.np = np; q.ldc = g.ldc;
    q.m = g.m; q.n = g.n; q.kp = NB; q.tiles_m = (int)(mp / 128); q.tiles_n = (int)(np / 128);
    int G = 1, nb = 0;
    hs_grid(g.m, g.n, &G, &nb);
    q.G = G; q.blk0 = bgemm;
    gemm.push_back(q);
    bgemm += nb;
  }
  void launch(const HsSplitProb *dr, const HsSplitProb *dc, const HsGemmProb *dg, hipStream_t st) const
  {
    launch_split_g(false, dr, (int)rows.size(), brows, st);
    launch_split_g(true, dc, (int)cols.size(), bcols, st);
    launch_gemm_hs_g(dg, (int)gemm.size(), bgemm, st);
  }
};

// Builder of the grouped launch sequence.  Launch closures refer to their
// descriptors by offset; the arena is uploaded once after planning.
struct Planner {
  bfa_batch *B;
  Arena ar;
  int phase = 0;
  explicit Planner(bfa_batch *b) : B(b), ar(b->hdesc) {}
  void op(std::function<void(hipStream_t)> f)
  {
    B->ops.push_back(std::move(f));
    B->op_phase.push_back(phase);
  }

  void gemm(std::vector<GemmProb> probs, bool f64)
  {
    std::vector<GemmProb> by[NUM_TILES];
    for (auto &p : probs) {
      if (p.m <= 0 || p.n <= 0 || p.k <= 0) continue;
      // the split-FP16 variant serves only the k = NB trailing updates (lu());
      // the other products of a bucket tuned to it take the 128 x 128 tile
      const int t = tile_for(B, p.m, p.n);
      by[t == TILE_HS ? TILE_128 : t].push_back(p);
    }
    for (int t = 0; t < NUM_TILES; t++) {
      if (by[t].empty()) continue;
      const int ts = tile_rows(t);
      int blk = 0;
      for (auto &p : by[t]) {
        p.tiles_m = (p.m + ts - 1) / ts;
        p.blk0 = blk;
        blk += p.tiles_m * ((p.n + ts - 1) / ts);
      }
      const size_t off = ar.put(by[t]);
      const int np = (int)by[t].size(), nblk = blk;
      op([Bp = B, off, np, nblk, t, f64](hipStream_t st) { launch_gemm_g(f64, t, dp<GemmProb>(Bp, off), np, nblk, st); });
    }
  }
  void trsm_diag(std::vector<TrsmProb> probs, bool lower, bool f64)
  {
    int blk = 0;
    std::vector<TrsmProb> v;
    for (auto &p : probs) {
      if (p.kb <= 0 || p.n <= 0) continue;
      p.blk0 = blk;
      blk += (p.n + 63) / 64;
      v.push_back(p);
    }
    if (v.empty()) return;
    const size_t off = ar.put(v);
    const int np = (int)v.size(), nblk = blk;
    op([Bp = B, off, np, nblk, lower, f64](hipStream_t st) { launch_trsm_g(f64, lower, dp<TrsmProb>(Bp, off), np, nblk, st); });
  }
  void swap(std::vector<SwapProb> probs, bool f64)
  {
    int blk = 0;
    std::vector<SwapProb> v;
    for (auto &p : probs) {
      if (p.c1 <= p.c0 || p.kb <= 0) continue;
      p.blk0 = blk;
      blk += (p.c1 - p.c0 + 127) / 128;
      v.push_back(p);
    }
    if (v.empty()) return;
    const size_t off = ar.put(v);
    const int np = (int)v.size(), nblk = blk;
    op([Bp = B, off, np, nblk, f64](hipStream_t st) { launch_swap_g(f64, dp<SwapProb>(Bp, off), np, nblk, st); });
  }

  // one LU sub-panel per instance: A (top-left of the panel), m rows, w cols
  struct Pan { c2 *A; int64_t lda; int m; int *ipiv; int piv_off; int inst; };

  int panel_leaf(const std::vector<Pan> &ps, int w)
  {
    // at most ~31 KB of rows per workgroup: more workgroups per panel are
    // faster even alone (m = 19,872: 256 us per 32 columns at 124 rows per
    // workgroup, 313 us at 600)
    const size_t cap = std::min(panel_lds_cap(false), (size_t)31 * 1024);
    const int row_bytes = w * (int)sizeof(c2);
    const int rmax = std::max(16, (int)(cap / row_bytes));
    const int maxres = panel_max_resident(false, (size_t)rmax * row_bytes);
    if (maxres <= 0) return fail(BFA_ERR_HIP, "batch: panel occupancy query failed");
    std::vector<PanelProb> chunk;
    int used = 0, rpw_max = 0;
    auto flush = [&]() {
      if (chunk.empty()) return;
      const size_t off = ar.put(chunk);
      const int np = (int)chunk.size(), nblk = used;
      const size_t lds = (size_t)rpw_max * row_bytes;
      op([Bp = B, off, np, nblk, lds](hipStream_t st) { launch_panel_g(false, dp<PanelProb>(Bp, off), np, nblk, lds, st); });
      chunk.clear();
      used = 0;
      rpw_max = 0;
    };
    for (const Pan &p : ps) {
      int nwg = (p.m + rmax - 1) / rmax;
      int rpw = (p.m + nwg - 1) / nwg;
      nwg = (p.m + rpw - 1) / rpw;
      if (nwg > maxres || nwg > 256) return fail(BFA_ERR_ARG, "batch: panel of %d rows too tall", p.m);
      if (used + nwg > maxres) flush();
      PanelProb q{};
      q.A = p.A; q.ipiv = p.ipiv; q.lda = p.lda; q.m = p.m; q.w = w; q.rpw = rpw; q.blk0 = used; q.nwg = nwg;
      q.piv_off = p.piv_off;
      q.slots = B->slots[p.inst];
      if (B->ncounters >= B->ncounters_max) return fail(BFA_ERR_ARG, "batch: barrier counter pool exhausted");
      // per panel: barrier counters, then the column keys (64-bit aligned)
      unsigned *c = B->counters() + (size_t)CSTRIDE * B->ncounters++;
      q.slots.bar = c;
      q.slots.key = (unsigned long long *)(c + 8 * 32);
      chunk.push_back(q);
      used += nwg;
      rpw_max = std::max(rpw_max, rpw);
    }
    flush();
    return BFA_OK;
  }

  void trsm_lower_blocked(const std::vector<Pan> &ps, const std::vector<c2 *> &Bm, const std::vector<int> &ncols, int kb)
  {
    // B(kb x n) := L^-1 B, L unit lower (kb x kb at the top of the panel)
    if (kb <= 256) {
      const int sc = trsm_lf_cols(false);
      std::vector<TrsmProb> v;
      int blk = 0;
      for (size_t k = 0; k < ps.size(); k++) {
        if (ncols[k] <= 0) continue;
        TrsmProb t{};
        t.L = ps[k].A; t.B = Bm[k]; t.ldl = ps[k].lda; t.ldb = ps[k].lda; t.kb = kb; t.n = ncols[k]; t.blk0 = blk;
        blk += (ncols[k] + sc - 1) / sc;
        v.push_back(t);
      }
      if (v.empty()) return;
      const size_t off = ar.put(v);
      const int np = (int)v.size(), nblk = blk;
      op([Bp = B, off, np, nblk](hipStream_t st) { launch_trsm_lf_g(false, dp<TrsmProb>(Bp, off), np, nblk, st); });
      return;
    }
    for (int i0 = 0; i0 < kb; i0 += 32) {
      const int ib = std::min(32, kb - i0);
      std::vector<TrsmProb> tp;
      std::vector<GemmProb> gp;
      for (size_t k = 0; k < ps.size(); k++) {
        if (ncols[k] <= 0) continue;
        const Pan &p = ps[k];
        TrsmProb t{};
        t.L = p.A + (size_t)i0 * p.lda + i0; t.B = Bm[k] + i0; t.ldl = p.lda; t.ldb = p.lda; t.kb = ib; t.n = ncols[k];
        tp.push_back(t);
        if (i0 + ib < kb) {
          GemmProb g{};
          g.A = p.A + (size_t)i0 * p.lda + i0 + ib; g.lda = p.lda;
          g.B = Bm[k] + i0; g.ldb = p.lda;
          g.C = Bm[k] + i0 + ib; g.ldc = p.lda;
          g.m = kb - i0 - ib; g.n = ncols[k]; g.k = ib;
          gp.push_back(g);
        }
      }
      trsm_diag(tp, true, false);
      gemm(gp, false);
    }
  }

  int panel_rec(const std::vector<Pan> &ps, int w)
  {
    if (w <= PWID) return panel_leaf(ps, w);
    const int w1 = w / 2, w2 = w - w1;
    int rc = panel_rec(ps, w1);
    if (rc) return rc;
    std::vector<SwapProb> sw;
    std::vector<c2 *> Ar;
    std::vector<int> nc;
    std::vector<GemmProb> gp;
    std::vector<Pan> right;
    for (const Pan &p : ps) {
      c2 *R = p.A + (size_t)w1 * p.lda;
      SwapProb s{};
      s.A = R - p.piv_off; s.ipiv = p.ipiv - p.piv_off; s.lda = p.lda; s.c0 = 0; s.c1 = w2; s.k0 = p.piv_off; s.kb = w1;
      sw.push_back(s);
      Ar.push_back(R);
      nc.push_back(w2);
      GemmProb g{};
      g.A = p.A + w1; g.lda = p.lda; g.B = R; g.ldb = p.lda; g.C = R + w1; g.ldc = p.lda;
      g.m = p.m - w1; g.n = w2; g.k = w1;
      gp.push_back(g);
      right.push_back({R + w1, p.lda, p.m - w1, p.ipiv + w1, p.piv_off + w1, p.inst});
    }
    swap(sw, false);
    trsm_lower_blocked(ps, Ar, nc, w1);
    gemm(gp, false);
    rc = panel_rec(right, w2);
    if (rc) return rc;
    std::vector<SwapProb> sl;
    for (const Pan &p : ps) {
      SwapProb s{};
      s.A = p.A - p.piv_off; s.ipiv = p.ipiv - p.piv_off; s.lda = p.lda; s.c0 = 0; s.c1 = w1;
      s.k0 = p.piv_off + w1; s.kb = w2;
      sl.push_back(s);
    }
    swap(sl, false);
    return BFA_OK;
  }

  int lu()
  {
    int maxNp = 0;
    for (int i = 0; i < B->count; i++) maxNp = std::max(maxNp, B->Np[i]);
    for (int j = 0; j < maxNp; j += NB) {
      std::vector<Pan> ps;
      for (int i = 0; i < B->count; i++) {
        if (B->Np[i] <= j) continue;
        c2 *A = (c2 *)B->LU32[i];
        const int64_t ld = B->Np[i];
        ps.push_back({A + (size_t)j * ld + j, ld, B->Np[i] - j, B->ipiv(i) + j, j, i});
      }
      int rc = panel_rec(ps, NB);
      if (rc) return rc;
      std::vector<SwapProb> sw;
      std::vector<c2 *> Bm;
      std::vector<int> nc;
      std::vector<GemmProb> gp;
      HsStep hs;
      for (const Pan &p : ps) {
        const int i = p.inst;
        c2 *A = (c2 *)B->LU32[i];
        const int *ip = B->ipiv(i);
        SwapProb s{};
        s.A = A; s.ipiv = ip; s.lda = p.lda; s.c0 = 0; s.c1 = j; s.k0 = j; s.kb = NB;
        sw.push_back(s);
        s.c0 = j + NB; s.c1 = B->Np[i];
        sw.push_back(s);
        const int rest = B->Np[i] - j - NB;
        Bm.push_back(p.A + (size_t)NB * p.lda);
        nc.push_back(rest);
        if (rest > 0) {
          GemmProb g{};
          g.A = p.A + NB; g.lda = p.lda; g.B = p.A + (size_t)NB * p.lda; g.ldb = p.lda;
          g.C = p.A + (size_t)NB * p.lda + NB; g.ldc = p.lda;
          g.m = rest; g.n = rest; g.k = NB;
          if (tile_for(B, rest, rest) == TILE_HS) {
            if ((rc = hs_ensure(B))) return rc;
            hs.add(B, i, g);
          } else {
            gp.push_back(g);
          }
        }
      }
      swap(sw, false);
      trsm_lower_blocked(ps, Bm, nc, NB);
      gemm(gp, false);
      if (!hs.gemm.empty()) {
        const size_t o_r = ar.put(hs.rows), o_c = ar.put(hs.cols), o_g = ar.put(hs.gemm);
        op([Bp = B, hs, o_r, o_c, o_g](hipStream_t st) {
          hs.launch(dp<HsSplitProb>(Bp, o_r), dp<HsSplitProb>(Bp, o_c), dp<HsGemmProb>(Bp, o_g), st);
        });
      }
    }
    return BFA_OK;
  }

  // streaming triangular solves of the complex64 factors: D -> Y -> Xs
  std::vector<SvProb> sv_descs(int nrhs, int &nblk)
  {
    std::vector<SvProb> v;
    nblk = 0;
    for (int i = 0; i < B->count; i++) {
      const bfa_batch::Off &o = B->off[i];
      SvProb q{};
      q.LU = B->LU32[i];
      q.inv = B->at<void>(o.inv);
      q.Bin = B->Dptr(i);
      q.Y = B->at<void>(o.Y);
      q.X = B->at<void>(o.Xs);
      q.tick = B->at<unsigned>(o.tick);
      q.flag = B->flags() + 2 * i;
      q.skip = B->conv() + i;
      q.ld = B->Np[i];
      q.ldx = B->Np[i];
      q.n = B->Np[i];
      q.nrhs = nrhs;
      q.blk0 = nblk;
      nblk += B->Np[i] / SVB;
      v.push_back(q);
    }
    return v;
  }
  // once per factorisation: inverted diagonal blocks of every instance
  void solve_inverses(int nrhs)
  {
    int nblk = 0;
    const size_t off = ar.put(sv_descs(nrhs, nblk));
    const int np = B->count;
    op([Bp = B, off, np, nblk](hipStream_t st) { launch_sv_inv_g(dp<SvProb>(Bp, off), np, nblk, st); });
    sv_off = off;
    sv_nblk = nblk;
  }
  // forward and backward sweeps (idle for converged instances)
  void solve_sweeps()
  {
    const size_t off = sv_off;
    const int np = B->count, nblk = sv_nblk;
    op([Bp = B, off, np, nblk](hipStream_t st) {
      launch_sv_stream_g(true, dp<SvProb>(Bp, off), np, nblk, st);
      launch_sv_stream_g(false, dp<SvProb>(Bp, off), np, nblk, st);
    });
  }
  size_t sv_off = 0;
  int sv_nblk = 0;
};

}  // namespace


bfa_batch::~bfa_batch()
{
  reset_plan();
  if (cap) (void)hipStreamDestroy(cap);
  for (void *p : M64) if (p) (void)hipFree(p);
  for (void *p : LU32) if (p) (void)hipFree(p);
  for (void *p : HS) if (p) (void)hipFree(p);
  for (void *p : TP) if (p) (void)hipFree(p);
  if (small) (void)hipFree(small);
  if (power) (void)hipFree(power);
}

namespace {

// Per-instance solve descriptors (grouped offsets filled in)
std::vector<SolveInst> solve_descs(const bfa_batch *B, int &nrow, int &nres, int &ncol)
{
  std::vector<SolveInst> v(B->count);
  nrow = nres = ncol = 0;
  for (int i = 0; i < B->count; i++) {
    SolveInst &q = v[i];
    const bfa_batch::Off &o = B->off[i];
    q.M = (const z2 *)B->M64[i];
    q.B = B->Bp(i);
    q.X = B->Xp(i);
    q.R = B->Rp(i);
    q.D = B->Dptr(i);
    q.P = B->at<z2>(o.P);
    q.norms = B->at<double>(o.norms);
    q.perm = B->perm(i);
    q.permw = B->perm(i);
    q.conv = B->conv() + i;
    q.iters = B->iters() + i;
    q.resid = B->resid() + i;
    q.ld = B->Np[i];
    q.N = B->N[i];
    q.Np = B->Np[i];
    q.nblk = (B->N[i] + 255) / 256;
    q.nchunk = (B->N[i] + RCH - 1) / RCH;
    q.row0 = nrow;
    nrow += (B->Np[i] + 255) / 256;
    q.res0 = nres;
    nres += q.nblk * q.nchunk;
    q.col0 = ncol;
    ncol += B->Np[i];
    q.LU = (c2 *)B->LU32[i];
    q.ipiv = B->ipiv(i);
    q.flags = B->flags() + 2 * i;
    q.Y = B->at<c2>(o.Y);
    q.Xs = B->at<c2>(o.Xs);
    q.tick = B->at<unsigned>(o.tick);
    q.tag = sv_tag_c64();
  }
  return v;
}

// the per-instance moment tables of the two-pass far field, or none
void tp_tables(bfa_batch *B)
{
  if (B->tp_state != 0) return;
  B->tp_state = -1;
  const char *e = getenv("BFA_FAR_TABLE");
  if (e && strcmp(e, "0") == 0) return;
  size_t need = 0, fr = 0, tot = 0;
  for (int i = 0; i < B->count; i++) need += tp_table_bytes(B->geo[i]->nt);
  if (hipMemGetInfo(&fr, &tot) != hipSuccess || need > fr / 4) return;
  B->TP.assign(B->count, nullptr);
  for (int i = 0; i < B->count; i++)
    if (hipMalloc(&B->TP[i], tp_table_bytes(B->geo[i]->nt)) != hipSuccess) {
      (void)hipGetLastError();
      for (void *&p : B->TP) {
        if (p) (void)hipFree(p);
        p = nullptr;
      }
      B->TP.clear();
      return;
    }
  B->tp_state = 1;
}

std::vector<AsmInst> asm_descs(const bfa_batch *B, bool solution, int &nfar, int &nnear, int &nrow, int &nemt,
                               int &ntab)
{
  std::vector<AsmInst> v(B->count);
  nfar = nnear = nrow = nemt = ntab = 0;
  for (int i = 0; i < B->count; i++) {
    AsmInst &q = v[i];
    const bfa_geometry *g = B->geo[i];
    q.g = g->dg;
    q.mat = B->mats[i];
    q.table = (const double *)g->d.table;
    q.M = B->M64[i];
    q.ldm = B->Np[i];
    q.B = solution ? (void *)B->Xp(i) : (void *)B->Bp(i);
    q.ldb = B->Np[i];
    q.Q = (double *)B->LU32[i];   // the complex64 LU is dead once the solve is done
    q.part = B->at<double>(B->off[i].part);
    const int N = B->N[i];
    q.ntiles = (N + FT2 - 1) / FT2;
    q.far0 = nfar;
    nfar += q.ntiles * (q.ntiles + 1) / 2;
    q.near0 = nnear;
    nnear += (int)((g->dg.nnz + 255) / 256);
    q.row0 = nrow;
    q.nrowblk = (N + 255) / 256;
    nrow += q.nrowblk;
    q.emt0 = nemt;
    nemt += q.ntiles * (q.ntiles + 1) / 2;
    q.tpt = B->tp_state == 1 ? (double *)B->TP[i] : nullptr;
    q.ntb = (g->nt + 31) / 32;
    q.tab0 = ntab;
    ntab += q.ntb * (q.ntb + 1) / 2;
  }
  return v;
}

int build_plan(bfa_batch *B, double tol, int maxit)
{
  B->reset_plan();
  Planner P(B);
  const int ni = B->count, npw = B->npw;
  // descriptor tables
  int nfar, nnear, nrow, nemt, ntab, nfar2, nnear2, nrow2, nemt2, ntab2, srow, sres, scol;
  tp_tables(B);
  const size_t o_asm = P.ar.put(asm_descs(B, false, nfar, nnear, nrow, nemt, ntab));
  const size_t o_emt = P.ar.put(asm_descs(B, true, nfar2, nnear2, nrow2, nemt2, ntab2));
  const size_t o_sol = P.ar.put(solve_descs(B, srow, sres, scol));
  if (nemt < 0 || nfar < 0) return fail(BFA_ERR_ARG, "batch: problem too large for one grouped launch");
  const PWBatch pws = B->pws;
  FreqState *fs = B->fs();
  // --- assembly
  P.phase = 0;
  // bfa_batch_run accepts positive real frequencies only (k_set_state stores
  // omega_im = 0), so the far field runs its real-k kernel
  if (B->tp_state == 1) {
    P.op([B, o_asm, ni, ntab, fs](hipStream_t st) { launch_tp_table_g(dp<AsmInst>(B, o_asm), ni, ntab, fs, true, st); });
    P.op([B, o_asm, ni, nfar, fs](hipStream_t st) { launch_far_gather_g(dp<AsmInst>(B, o_asm), ni, nfar, fs, st); });
  } else {
    P.op([B, o_asm, ni, nfar, fs](hipStream_t st) { launch_far_g(dp<AsmInst>(B, o_asm), ni, nfar, fs, true, st); });
  }
  bool expr = false;
  for (const bfa_material &m : B->mats) expr = expr || m.kind == 4;
  P.op([B, o_asm, ni, nnear, fs, expr](hipStream_t st) {
    launch_near_g(dp<AsmInst>(B, o_asm), ni, nnear, fs, expr, st);
  });
  P.op([B, o_asm, ni, nrow, fs, pws, npw](hipStream_t st) {
    launch_rhs_g(dp<AsmInst>(B, o_asm), ni, nrow, fs, pws, npw, st);
  });
  // --- convert + factor
  P.phase = 1;
  const size_t ncnt = B->ncounters_max * CSTRIDE;
  unsigned *cnt = B->counters();
  int *flags0 = B->flags();
  const size_t nflags = (size_t)2 * ni;
  P.op([cnt, ncnt, flags0, nflags](hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, st, cnt, ncnt);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (unsigned *)flags0, nflags);
  });
  P.op([B, o_sol, ni, scol](hipStream_t st) {
    hipLaunchKernelGGL(kg_convert, dim3(scol), dim3(256), 0, st, dp<SolveInst>(B, o_sol), ni);
  });
  int rc = P.lu();
  if (rc) return rc;
  // --- solve with refinement
  P.phase = 2;
  int maxNp = 0;
  for (int i = 0; i < ni; i++) maxNp = std::max(maxNp, B->Np[i]);
  const size_t perm_lds = (size_t)maxNp * sizeof(int);
  if (perm_lds > 64 * 1024)
    HIPCHECK(hipFuncSetAttribute((const void *)kg_perm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)perm_lds));
  P.op([B, o_sol, ni, perm_lds](hipStream_t st) {
    hipLaunchKernelGGL(kg_perm, dim3(ni), dim3(64), perm_lds, st, dp<SolveInst>(B, o_sol), ni);
  });
  P.op([B, o_sol, ni, srow, npw](hipStream_t st) {
    hipLaunchKernelGGL(kg_refine_init, dim3(srow), dim3(256), 0, st, dp<SolveInst>(B, o_sol), ni, npw);
  });
  P.solve_inverses(npw);
  for (int it = 0; it < maxit; it++) {
    P.op([B, o_sol, ni, srow, npw](hipStream_t st) {
      hipLaunchKernelGGL(kg_gather, dim3(srow), dim3(256), 0, st, dp<SolveInst>(B, o_sol), ni, npw);
    });
    P.solve_sweeps();
    P.op([B, o_sol, ni, srow, sres, npw, tol](hipStream_t st) {
      const SolveInst *d = dp<SolveInst>(B, o_sol);
      hipLaunchKernelGGL(kg_add, dim3(srow), dim3(256), 0, st, d, ni, npw);
      hipLaunchKernelGGL(kg_residual_part, dim3(sres), dim3(256), 0, st, d, ni, npw);
      hipLaunchKernelGGL(kg_residual_final, dim3(srow, npw), dim3(256), 0, st, d, ni, npw);
      hipLaunchKernelGGL(kg_refine_check, dim3(ni), dim3(64), 0, st, d, ni, npw, tol);
    });
  }
  // --- power (the complex64 LU storage becomes the EMT scratch)
  P.phase = 3;
  P.op([B, o_emt, ni, nemt, nnear2, nrow, fs, pws, npw](hipStream_t st) {
    const AsmInst *d = dp<AsmInst>(B, o_emt);
    launch_emt_q_g(d, ni, nemt, fs, st);
    launch_emt_near_g(d, ni, nnear2, fs, st);
    launch_emt_rows_g(d, ni, nrow, fs, pws, npw, st);
    launch_emt_final_g(d, ni, fs, npw, B->power, st);
  });
  // upload descriptors
  if (!B->hdesc.empty()) {
    HIPCHECK(hipMalloc(&B->ddesc, B->hdesc.size()));
    HIPCHECK(hipMemcpy(B->ddesc, B->hdesc.data(), B->hdesc.size(), hipMemcpyHostToDevice));
  }
  B->planned = true;
  B->plan_tol = tol;
  B->plan_maxit = maxit;
  return BFA_OK;
}

}  // namespace

#define BATCH_MAXFREQ 256

namespace {

// per-frequency record of the refinement and LU status
__global__ void k_record(const FreqState *fs, int ni, const int *iters, const double *resid, const int *flags,
                         int *hit, double *hres, int *hinfo)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const size_t o = (size_t)fs->fidx * ni + i;
  hit[o] = iters[i];
  hres[o] = resid[i];
  hinfo[o] = flags[2 * i] ? -1 : flags[2 * i + 1];
}

}  // namespace

extern "C" int bfa_batch_create(int count, bfa_geometry *const *geoms, const bfa_material *mats,
                                const bfa_planewave *pw, int npw, bfa_batch **out)
{
  if (!out || count <= 0 || !geoms || !mats || !pw || npw <= 0) return fail(BFA_ERR_ARG, "bfa_batch_create: bad argument");
  if (npw > BFA_MAXPW_BATCH) return fail(BFA_ERR_ARG, "bfa_batch_create: at most %d plane waves", BFA_MAXPW_BATCH);
  *out = nullptr;
  bfa_batch *B = new bfa_batch();
  auto bail = [&](int rc) { delete B; return rc; };
  B->count = count;
  B->npw = npw;
  for (int j = 0; j < npw; j++) B->pws.pw[j] = pw[j];
  for (int b = 0; b < NBUCKET; b++) {
    B->tile[b] = -1;
    for (int t = 0; t < NUM_TILES; t++) B->tune_ms[b][t] = 0.0;
  }
  if (hipGetDevice(&B->device) != hipSuccess) return bail(fail(BFA_ERR_HIP, "bfa_batch_create: no device"));
  if (panel_sync_words() != CSTRIDE) return bail(fail(BFA_ERR_STATE, "bfa_batch_create: panel barrier layout mismatch"));
  size_t o = 0;
  auto take = [&o](size_t bytes) { size_t r = (o + 255) / 256 * 256; o = r + bytes; return r; };
  for (int i = 0; i < count; i++) {
    bfa_geometry *g = geoms[i];
    if (!g || !g->uploaded || !g->static_built)
      return bail(fail(BFA_ERR_STATE, "bfa_batch_create: instance %d not uploaded / static table not built", i));
    B->geo.push_back(g);
    B->mats.push_back(mats[i]);
    const int N = g->nbf, Np = (N + NB - 1) / NB * NB;
    if (Np > 38400) return bail(fail(BFA_ERR_ARG, "bfa_batch_create: instance %d too large (N=%d)", i, N));
    B->N.push_back(N);
    B->Np.push_back(Np);
    bfa_batch::Off f{};
    const int nchunk = (N + RCH - 1) / RCH, nblk = (N + 255) / 256;
    f.ipiv = take((size_t)Np * 4);
    f.perm = take((size_t)Np * 4);
    f.B = take((size_t)npw * Np * 16);
    f.X = take((size_t)npw * Np * 16);
    f.R = take((size_t)npw * Np * 16);
    f.D = take((size_t)npw * Np * 8);
    f.Y = take((size_t)npw * Np * 8);
    f.Xs = take((size_t)npw * Np * 8);
    f.inv = take((size_t)2 * (Np / SVB) * SVB * SVB * 8);
    f.tick = take(8);
    f.P = take((size_t)nchunk * npw * N * 16);
    f.norms = take((size_t)npw * nblk * 16);
    f.part = take((size_t)2 * npw * nblk * 8);
    f.sdata = take((size_t)2 * 256 * PWID * 16);
    f.sj = take(2 * PWID * 16);
    B->off.push_back(f);
    B->ncounters_max += (size_t)(Np / NB) * (NB / PWID);
  }
  B->o_fs = take(sizeof(FreqState));
  B->o_conv = take((size_t)count * 4);
  B->o_iters = take((size_t)count * 4);
  B->o_resid = take((size_t)count * 8);
  B->o_flags = take((size_t)count * 8);
  B->o_counters = take(B->ncounters_max * CSTRIDE * 4);
  B->o_hist_it = take((size_t)BATCH_MAXFREQ * count * 4);
  B->o_hist_res = take((size_t)BATCH_MAXFREQ * count * 8);
  B->o_hist_info = take((size_t)BATCH_MAXFREQ * count * 4);
  hipError_t e = hipMalloc(&B->small, o);
  if (e != hipSuccess) return bail(fail(BFA_ERR_OOM, "bfa_batch_create: %zu bytes of work space", o));
  if (hipMemset(B->small, 0, o) != hipSuccess) return bail(fail(BFA_ERR_HIP, "bfa_batch_create: memset"));
  e = hipMalloc((void **)&B->power, (size_t)BATCH_MAXFREQ * count * npw * 16);
  if (e != hipSuccess) return bail(fail(BFA_ERR_OOM, "bfa_batch_create: power buffer"));
  for (int i = 0; i < count; i++) {
    const int N = B->N[i], Np = B->Np[i];
    void *m = nullptr, *l = nullptr;
    if (hipMalloc(&m, (size_t)Np * N * 16) != hipSuccess)
      return bail(fail(BFA_ERR_OOM, "bfa_batch_create: instance %d matrix (%d x %d complex128)", i, N, N));
    B->M64.push_back(m);
    if (hipMalloc(&l, (size_t)Np * Np * 8) != hipSuccess)
      return bail(fail(BFA_ERR_OOM, "bfa_batch_create: instance %d complex64 LU", i));
    B->LU32.push_back(l);
    PanelSlots ps{};
    ps.rowdata = B->at<void>(B->off[i].sdata);
    ps.rowj = B->at<void>(B->off[i].sj);
    ps.flags = B->flags() + 2 * i;
    ps.bar = nullptr;
    B->slots.push_back(ps);
  }
  if (hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(BFA_ERR_HIP, "bfa_batch_create: stream"));
  *out = B;
  return BFA_OK;
}

extern "C" int bfa_batch_destroy(bfa_batch *b)
{
  if (b) {
    (void)hipStreamSynchronize(b->cap);
    delete b;
  }
  return BFA_OK;
}

extern "C" int bfa_batch_info(const bfa_batch *b, int *count, int *npw, int *N, int *Np)
{
  if (!b) return fail(BFA_ERR_ARG, "bfa_batch_info: null batch");
  if (count) *count = b->count;
  if (npw) *npw = b->npw;
  for (int i = 0; i < b->count; i++) {
    if (N) N[i] = b->N[i];
    if (Np) Np[i] = b->Np[i];
  }
  return BFA_OK;
}

// Autotuner: for every size bucket (floor(log2 m) of the trailing-update
// dimension) that the batch's LU visits, time each GEMM tile variant on the
// batch's own grouped trailing updates of the step that has the most
// problems in that bucket, and keep the fastest.
extern "C" int bfa_batch_tune(bfa_batch *B, int reps)
{
  if (!B) return fail(BFA_ERR_ARG, "bfa_batch_tune: null batch");
  if (reps <= 0) reps = 3;
  hipStream_t st = B->cap;
  for (int i = 0; i < B->count; i++)
    HIPCHECK(hipMemsetAsync(B->LU32[i], 0, (size_t)B->Np[i] * B->Np[i] * 8, st));
  int maxNp = 0;
  for (int i = 0; i < B->count; i++) maxNp = std::max(maxNp, B->Np[i]);
  const int nsteps = maxNp / NB;
  std::vector<std::vector<int>> cnt(NBUCKET, std::vector<int>(nsteps + 1, 0));
  for (int j = 0; j < nsteps; j++)
    for (int i = 0; i < B->count; i++) {
      const int rest = B->Np[i] - (j + 1) * NB;
      if (rest > 0) cnt[bucket_of(rest)][j]++;
    }
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  void *dbuf = nullptr;
  const size_t hs_r = 0, hs_c = ((size_t)B->count * sizeof(HsSplitProb) + 255) / 256 * 256,
               hs_g = 2 * hs_c;
  HIPCHECK(hipMalloc(&dbuf, std::max((size_t)B->count * sizeof(GemmProb), hs_g + B->count * sizeof(HsGemmProb)) + 256));
  int rc = BFA_OK;
  for (int b = 0; b < NBUCKET && rc == BFA_OK; b++) {
    int jbest = -1, best = 0;
    for (int j = 0; j < nsteps; j++)
      if (cnt[b][j] > best) { best = cnt[b][j]; jbest = j; }
    if (jbest < 0) continue;
    double tbest = 1e30;
    int vbest = -1;
    for (int t = 0; t < NUM_TILES; t++) {
      std::vector<GemmProb> v;
      HsStep hs;
      const int ts = tile_rows(t);
      int blk = 0;
      if (t == TILE_HS && (rc = hs_ensure(B))) break;
      for (int i = 0; i < B->count; i++) {
        const int j = jbest, rest = B->Np[i] - (j + 1) * NB;
        if (rest <= 0 || bucket_of(rest) != b) continue;
        c2 *A = (c2 *)B->LU32[i];
        const int64_t ld = B->Np[i];
        c2 *Ajj = A + (size_t)j * NB * ld + (size_t)j * NB;
        GemmProb g{};
        g.A = Ajj + NB; g.lda = ld; g.B = Ajj + (size_t)NB * ld; g.ldb = ld; g.C = Ajj + (size_t)NB * ld + NB; g.ldc = ld;
        g.m = rest; g.n = rest; g.k = NB;
        if (t == TILE_HS) {
          hs.add(B, i, g);
          continue;
        }
        g.tiles_m = (rest + ts - 1) / ts;
        g.blk0 = blk;
        blk += g.tiles_m * ((rest + ts - 1) / ts);
        v.push_back(g);
      }
      // the split variant is timed with its two plane launches
      char *db = (char *)dbuf;
      const bool up = t == TILE_HS
                          ? hipMemcpy(db + hs_r, hs.rows.data(), hs.rows.size() * sizeof(HsSplitProb),
                                      hipMemcpyHostToDevice) == hipSuccess &&
                                hipMemcpy(db + hs_c, hs.cols.data(), hs.cols.size() * sizeof(HsSplitProb),
                                          hipMemcpyHostToDevice) == hipSuccess &&
                                hipMemcpy(db + hs_g, hs.gemm.data(), hs.gemm.size() * sizeof(HsGemmProb),
                                          hipMemcpyHostToDevice) == hipSuccess
                          : hipMemcpy(dbuf, v.data(), v.size() * sizeof(GemmProb), hipMemcpyHostToDevice) == hipSuccess;
      if (!up) {
        rc = fail(BFA_ERR_HIP, "bfa_batch_tune: upload");
        break;
      }
      auto go = [&]() {
        if (t == TILE_HS)
          hs.launch((const HsSplitProb *)(db + hs_r), (const HsSplitProb *)(db + hs_c), (const HsGemmProb *)(db + hs_g),
                    st);
        else
          launch_gemm_g(false, t, (const GemmProb *)dbuf, (int)v.size(), blk, st);
      };
      go();
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps; r++) go();
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(BFA_ERR_HIP, "bfa_batch_tune: timing run failed"); break; }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      B->tune_ms[b][t] = ms / reps;
      if (ms < tbest) { tbest = ms; vbest = t; }
    }
    B->tile[b] = vbest;
  }
  (void)hipFree(dbuf);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == BFA_OK && hipGetLastError() != hipSuccess) rc = fail(BFA_ERR_HIP, "bfa_batch_tune: kernel launch failed");
  B->reset_plan();
  return rc;
}

// the tile variant per bucket, set by the caller (a stored tuning, or one
// variant everywhere); the next run plans again
extern "C" int bfa_batch_set_tiles(bfa_batch *B, const int *tile)
{
  if (!B || !tile) return fail(BFA_ERR_ARG, "bfa_batch_set_tiles: null argument");
  for (int b = 0; b < NBUCKET; b++)
    if (tile[b] < -1 || tile[b] >= NUM_TILES) return fail(BFA_ERR_ARG, "bfa_batch_set_tiles: tile %d", tile[b]);
  for (int b = 0; b < NBUCKET; b++) B->tile[b] = tile[b];
  B->reset_plan();
  return BFA_OK;
}

extern "C" int bfa_batch_tuning(const bfa_batch *B, int *tile, double *ms)
{
  if (!B) return fail(BFA_ERR_ARG, "bfa_batch_tuning: null batch");
  for (int b = 0; b < NBUCKET; b++) {
    if (tile) tile[b] = B->tile[b];
    for (int t = 0; t < NUM_TILES; t++)
      if (ms) ms[b * NUM_TILES + t] = B->tune_ms[b][t];
  }
  return BFA_OK;
}

extern "C" int bfa_batch_run(bfa_batch *B, const double *omega, int nfreq, int use_graph, double tol, int maxit,
                             double *power, int *iters, double *resid, int *info)
{
  if (!B || !omega || nfreq <= 0) return fail(BFA_ERR_ARG, "bfa_batch_run: bad argument");
  if (nfreq > BATCH_MAXFREQ) return fail(BFA_ERR_ARG, "bfa_batch_run: at most %d frequencies per call", BATCH_MAXFREQ);
  for (int f = 0; f < nfreq; f++)
    if (!(omega[f] > 0.0)) return fail(BFA_ERR_ARG, "bfa_batch_run: frequencies must be positive");
  if (maxit <= 0) maxit = 6;
  if (!(tol > 0.0)) tol = 1e-10;
  int dev = -1;
  HIPCHECK(hipGetDevice(&dev));
  if (dev != B->device) return fail(BFA_ERR_STATE, "bfa_batch_run: batch was created on device %d", B->device);
  if (!B->planned || B->plan_maxit != maxit || B->plan_tol != tol) {
    int rc = build_plan(B, tol, maxit);
    if (rc) return rc;
  }
  hipStream_t st = B->cap;
  const int ni = B->count;
  int *hit = B->at<int>(B->o_hist_it), *hinfo = B->at<int>(B->o_hist_info);
  double *hres = B->at<double>(B->o_hist_res);
  auto record = [&](hipStream_t s) {
    hipLaunchKernelGGL(k_record, dim3((ni + 63) / 64), dim3(64), 0, s, (const FreqState *)B->fs(), ni,
                       (const int *)B->iters(), (const double *)B->resid(), (const int *)B->flags(), hit, hres, hinfo);
  };
  if (use_graph && !B->exec) {
    // graph 1: assembly, convert, factor, solve, status record; graph 2: power
    for (int part = 0; part < 2; part++) {
      HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (size_t k = 0; k < B->ops.size(); k++)
        if ((B->op_phase[k] == 3) == (part == 1)) B->ops[k](st);
      if (part == 0) record(st);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (e != hipSuccess) return fail(BFA_ERR_HIP, "bfa_batch_run: graph capture failed: %s", hipGetErrorString(e));
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(BFA_ERR_HIP, "bfa_batch_run: graph instantiation failed: %s", hipGetErrorString(e));
      }
      if (part == 0) { B->graph = g; B->exec = x; }
      else { B->graph_p = g; B->exec_p = x; }
    }
  }
  // events per frequency: [start, after solve, after power]
  std::vector<hipEvent_t> ev(3 * (size_t)nfreq + 5);
  for (auto &x : ev) HIPCHECK(hipEventCreate(&x));
  hipEvent_t *ph_ev = ev.data() + 3 * (size_t)nfreq;   // eager phase marks [4] + end
  double ph[4] = {0, 0, 0, 0};
  for (int f = 0; f < nfreq; f++) {
    hipEvent_t *e3 = ev.data() + 3 * (size_t)f;
    HIPCHECK(hipEventRecord(e3[0], st));
    hipLaunchKernelGGL(k_set_state, dim3(1), dim3(1), 0, st, B->fs(), omega[f], 0.0, f);
    if (use_graph) {
      HIPCHECK(hipGraphLaunch(B->exec, st));
      HIPCHECK(hipEventRecord(e3[1], st));
      if (power) HIPCHECK(hipGraphLaunch(B->exec_p, st));
      HIPCHECK(hipEventRecord(e3[2], st));
    } else {
      int cur = -1;
      for (size_t k = 0; k < B->ops.size(); k++) {
        if (B->op_phase[k] != cur) {
          cur = B->op_phase[k];
          HIPCHECK(hipEventRecord(ph_ev[cur], st));
          if (cur == 3) {
            record(st);
            HIPCHECK(hipEventRecord(e3[1], st));
            if (!power) break;
          }
        }
        B->ops[k](st);
      }
      if (cur != 3) {   // no power ops ran
        HIPCHECK(hipEventRecord(ph_ev[3], st));
      }
      HIPCHECK(hipEventRecord(e3[2], st));
      HIPCHECK(hipEventRecord(ph_ev[4], st));
      HIPCHECK(hipEventSynchronize(ph_ev[4]));
      float ms;
      for (int p = 0; p < 4; p++) {
        HIPCHECK(hipEventElapsedTime(&ms, ph_ev[p], ph_ev[p + 1]));
        ph[p] += ms;
      }
    }
  }
  HIPCHECK(hipEventSynchronize(ev[3 * (size_t)nfreq - 1]));
  double tsolve = 0.0, tpower = 0.0;
  for (int f = 0; f < nfreq; f++) {
    float a = 0.f, b = 0.f;
    HIPCHECK(hipEventElapsedTime(&a, ev[3 * f], ev[3 * f + 1]));
    HIPCHECK(hipEventElapsedTime(&b, ev[3 * f + 1], ev[3 * f + 2]));
    tsolve += a;
    tpower += b;
  }
  for (auto &x : ev) (void)hipEventDestroy(x);
  HIPCHECK(hipGetLastError());
  B->run_ms = tsolve;
  B->power_ms = tpower;
  for (int p = 0; p < 4; p++) B->phase_ms[p] = use_graph ? 0.0 : ph[p];
  if (power) HIPCHECK(hipMemcpy(power, B->power, (size_t)nfreq * ni * B->npw * 16, hipMemcpyDeviceToHost));
  if (iters) HIPCHECK(hipMemcpy(iters, hit, (size_t)nfreq * ni * 4, hipMemcpyDeviceToHost));
  if (resid) HIPCHECK(hipMemcpy(resid, hres, (size_t)nfreq * ni * 8, hipMemcpyDeviceToHost));
  std::vector<int> inf((size_t)nfreq * ni);
  HIPCHECK(hipMemcpy(inf.data(), hinfo, inf.size() * 4, hipMemcpyDeviceToHost));
  if (info) memcpy(info, inf.data(), inf.size() * 4);
  for (int v : inf)
    if (v < 0) return fail(BFA_ERR_HIP, "bfa_batch_run: LU panel barrier timed out");
  return BFA_OK;
}

extern "C" int bfa_batch_timing(const bfa_batch *B, double *ms)
{
  if (!B || !ms) return fail(BFA_ERR_ARG, "bfa_batch_timing: bad argument");
  ms[0] = B->run_ms;
  for (int p = 0; p < 4; p++) ms[1 + p] = B->phase_ms[p];
  ms[5] = B->power_ms;
  return BFA_OK;
}

extern "C" int bfa_batch_solution(const bfa_batch *B, int inst, void *J)
{
  if (!B || !J || inst < 0 || inst >= B->count) return fail(BFA_ERR_ARG, "bfa_batch_solution: bad argument");
  HIPCHECK(hipMemcpy2D(J, (size_t)B->N[inst] * 16, B->Xp(inst), (size_t)B->Np[inst] * 16, (size_t)B->N[inst] * 16,
                       B->npw, hipMemcpyDeviceToHost));
  return BFA_OK;
}

extern "C" int bfa_batch_matrix(const bfa_batch *B, int inst, void *M)
{
  if (!B || !M || inst < 0 || inst >= B->count) return fail(BFA_ERR_ARG, "bfa_batch_matrix: bad argument");
  HIPCHECK(hipMemcpy2D(M, (size_t)B->N[inst] * 16, B->M64[inst], (size_t)B->Np[inst] * 16, (size_t)B->N[inst] * 16,
                       B->N[inst], hipMemcpyDeviceToHost));
  return BFA_OK;
}
