// bfa_group.h -- grouped ("batched, uneven sizes") launches of the dense
// LU / solve kernels.  One launch covers the same operation on many
// independent matrices: the launch's workgroups are split among problems by
// a prefix (blk0) stored in each problem descriptor, and every workgroup
// looks up its problem by binary search.  Descriptors live in device memory
// prepared once per plan, so a launch sequence can be replayed (or captured
// into a hipGraph) with no host work in between.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "bfa_internal.h"

namespace bfa {

// C(m x n) -= A(m x k) * B(k x n); tiles_m x tiles_n workgroups from blk0
struct GemmProb {
  const void *A, *B;
  void *C;
  int64_t lda, ldb, ldc;
  int m, n, k, tiles_m, blk0, pad;
};

// B(kb x n) := L^-1 B (unit lower) or U^-1 B (upper), kb <= 32; one
// workgroup per 64 columns of B from blk0
struct TrsmProb {
  const void *L;
  void *B;
  int64_t ldl, ldb;
  int kb, n, blk0, pad;
};

// rows k0..k0+kb-1 <-> ipiv[.] on columns [c0, c1) of A; 128 columns per
// workgroup from blk0
struct SwapProb {
  void *A;
  const int *ipiv;
  int64_t lda;
  int c0, c1, k0, kb, blk0, pad;
};

// Panel workspace of one problem (cross-workgroup slots and barrier)
struct PanelSlots {
  void *rowdata;    // [2][maxwg][32] complex: each workgroup's pivot candidate row
  void *rowj;       // [2][32] complex
  unsigned *bar;    // barrier counters, zero at launch (panel_sync_words() u32 with key)
  unsigned long long *key;  // per-column packed argmax keys, zero at launch
  int *flags;       // [0] spin timeout, [1] first zero pivot (1-based)
};

// m x w panel (w <= 32) factored by nwg workgroups from blk0, rpw rows each
struct PanelProb {
  void *A;
  int *ipiv;
  int64_t lda;
  int m, w, rpw, blk0, nwg, piv_off;
  PanelSlots slots;
};


// Streaming triangular solve of one problem (see k_trsv_stream): the
// diagonal-block inverses inv [2][n/64][64*64] are built once per
// factorisation (blk0 also indexes them); each refinement round reads Bin,
// writes the tag-filled Y (forward sweep) and X (backward sweep).
struct SvProb {
  const void *LU;
  void *inv;
  const void *Bin;
  void *Y, *X;
  unsigned *tick;        // [2] tickets of the two sweeps
  int *flag;             // set when a wait gives up
  const int *skip;
  int64_t ld, ldx;
  int n, nrhs, blk0;
};

// launchers (f64 selects complex128, else complex64); nblk = total workgroups
void launch_gemm_g(bool f64, int tile, const GemmProb *d, int np, int nblk, hipStream_t st);
void launch_trsm_g(bool f64, bool lower, const TrsmProb *d, int np, int nblk, hipStream_t st);
void launch_swap_g(bool f64, const SwapProb *d, int np, int nblk, hipStream_t st);
// B(kb x n) := L^-1 B, L unit lower, kb <= 256, one launch: trsm_lf_cols()
// columns of B per workgroup from blk0
int trsm_lf_cols(bool f64);
void launch_trsm_lf_g(bool f64, const TrsmProb *d, int np, int nblk, hipStream_t st);
void launch_panel_g(bool f64, const PanelProb *d, int np, int nblk, size_t lds, hipStream_t st);
void launch_sv_inv_g(const SvProb *d, int np, int nblk, hipStream_t st);            // complex64
void launch_sv_stream_g(bool lower, const SvProb *d, int np, int nblk, hipStream_t st);
// tag value that fills Y and X before a streaming sweep (complex64 word)
unsigned long long sv_tag_c64();
// panel capacity: LDS bytes available per workgroup and the co-resident
// workgroup count of the grouped panel kernel (0 on error)
size_t panel_lds_cap(bool f64);
// u32 words of one panel's barrier counters + keys (key = bar + 8 * 32)
int panel_sync_words();
int panel_max_resident(bool f64, size_t lds);

// ---------------------------------------------------------------------------
// Grouped assembly / RHS / power over the instances of a batch.  Each
// instance carries its own geometry, material and matrices, plus the first
// workgroup of each grouped launch that it owns.
struct AsmInst {
  DevGeom g;
  bfa_material mat;
  const double *table;   // static records [nnz][6]
  void *M;               // complex128 VIE matrix, N x N used, ld = ldm
  void *B;               // complex128 right-hand sides / solutions, ld = ldb
  double *Q;             // EMT scratch, N x N doubles, ld = N
  double *part;          // EMT partial sums [2][npw][row blocks]
  int64_t ldm, ldb;
  int far0, ntiles;      // far-field tiles (upper triangle of FT2 x FT2 BF tiles) from far0
  int near0, row0;       // touching-pair blocks (256) / row blocks (256)
  int emt0, nrowblk;     // EMT Q blocks (256 entries of N x N) from emt0
  double *tpt;           // two-pass far field: the instance's tetrahedron-pair moment table (null: single pass)
  int tab0, ntb;         // its 32 x 32 tetrahedron tiles (upper triangle of ntb x ntb) from tab0
};

// Frequency of the current replay: written by a one-thread kernel before a
// graph launch, read by every frequency-dependent kernel.
struct FreqState {
  double omega_re, omega_im;
  int fidx, pad;
};

#define BFA_MAXPW_BATCH 8
struct PWBatch { bfa_planewave pw[BFA_MAXPW_BATCH]; };

// realk: every frequency the launch will see is real (the kernel then never
// compiles the complex-k moments, which would spill registers)
void launch_far_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, bool realk, hipStream_t st);
// the two-pass far field of the instances (every tpt set): the moment tables
// (ntab workgroups in all), then the BF tiles gathering them (nfar)
void launch_tp_table_g(const AsmInst *d, int ni, int ntab, const FreqState *fs, bool realk, hipStream_t st);
void launch_far_gather_g(const AsmInst *d, int ni, int nfar, const FreqState *fs, hipStream_t st);
// bytes of one instance's moment table (nt tetrahedra)
size_t tp_table_bytes(int nt);
// expr: some instance's material is a compiled SVTensor (material kind 4)
void launch_near_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, bool expr, hipStream_t st);
void launch_rhs_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, const PWBatch &pw, int npw,
                  hipStream_t st);
void launch_emt_q_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, hipStream_t st);
void launch_emt_near_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, hipStream_t st);
void launch_emt_rows_g(const AsmInst *d, int ni, int nblk, const FreqState *fs, const PWBatch &pw, int npw,
                       hipStream_t st);
// out[(fidx * ni + i) * npw + j][2] = {P_abs, P_scat}
void launch_emt_final_g(const AsmInst *d, int ni, const FreqState *fs, int npw, double *out, hipStream_t st);

// GEMM tile variants for the autotuner: complex64 on FP32 MFMA in 64 x 64
// or 128 x 128 tiles, or split FP16 (hi + lo planes, 128 x 128 tiles on the
// FP16 matrix cores; the trailing updates with k = 256 only)
enum { TILE_64 = 0, TILE_128 = 1, TILE_HS = 2, NUM_TILES = 3 };
inline int tile_rows(int t) { return t == TILE_64 ? 64 : 128; }

// Split-FP16 trailing update of a batch.  Planes: X (rows of A or columns
// of B, nv of them, k each, column-major with ldx) -> P[4][pr][kp] halves
// scaled by 2^e per row / column, fac = 2^-e; blocks of 32 rows / columns
// from blk0, covering pr (a multiple of 128 >= nv).
struct HsSplitProb {
  const void *X;
  void *P, *fac;
  int64_t ldx, pr;
  int nv, k, kp, blk0;
};
// C(m x n) -= (planes of A)(planes of B) 2^-(e_row + e_col); the workgroups
// of one problem (a multiple of 8, from blk0) in G x G super-tiles
struct HsGemmProb {
  const void *PA, *PB, *rf, *cf;
  void *C;
  int64_t mp, np, ldc;
  int m, n, kp, tiles_m, tiles_n, G, blk0, pad;
};
void launch_split_g(bool cols, const HsSplitProb *d, int np, int nblk, hipStream_t st);
void launch_gemm_hs_g(const HsGemmProb *d, int np, int nblk, hipStream_t st);
// super-tile size and workgroup count of an m x n product in 128 x 128 tiles
void hs_grid(int m, int n, int *G, int *nblk);

}  // namespace bfa
