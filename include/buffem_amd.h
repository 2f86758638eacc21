/*
 * buffem_amd.h -- C ABI of the MI355X-native dense VIE hot path.
 *
 * This is the drop-in boundary beneath the SWGGeometry / HMatrix API of
 * HomerReid/buff-em (src/libs/libbuff/libbuff.h:200-283 and SCUFF-EM's
 * HMatrix::LUFactorize/LUSolve as called from
 * src/applications/buff-scatter/buff-scatter.cc:236-314).  Every entry point
 * takes plain pointers and sizes; device buffers are raw HIP device pointers,
 * streams are hipStream_t passed as void* (NULL = default stream).
 *
 * Conventions
 *  - Dense matrices are complex128 (interleaved re,im doubles), column-major
 *    with leading dimension ld (in elements).  Entry (r,c) is at
 *    ((size_t)c*ld + r).  Complex64 matrices use the same layout with floats.
 *  - Basis-function (BF) numbering follows the reference's InitFaceList
 *    (interior faces in order of their second tetrahedron visit,
 *    src/libs/libbuff/InitFaceList.cc:141-260).
 *  - Pivot vectors are 0-based: row i was interchanged with row ipiv[i].
 *  - Every function returns a status code (BFA_OK == 0); nothing calls
 *    exit().  bfa_last_error() returns a message for the calling thread.
 *    This replaces ErrExit (src/libs/libbuff/SWGVolume.cc:107,164-174,
 *    src/libs/libbuff/TetCR.cc:104).
 */
#ifndef BUFFEM_AMD_H
#define BUFFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BFA_OK             0
#define BFA_ERR_ARG       -1   /* invalid argument / size                     */
#define BFA_ERR_MESH      -2   /* invalid mesh (file format, topology)        */
#define BFA_ERR_SINGULAR  -3   /* exactly zero pivot in LU                    */
#define BFA_ERR_OOM       -4   /* device or host allocation failed            */
#define BFA_ERR_HIP       -5   /* HIP runtime error                           */
#define BFA_ERR_STATE     -6   /* call order (e.g. static table not built)    */
#define BFA_ERR_IO        -7   /* file could not be opened / read             */

typedef struct bfa_geometry bfa_geometry;

/* Material eps(omega, x): SVTensor::Evaluate (src/libs/libbuff/SVTensor.cc:
 * 306-373).  Closed forms (expr unused):
 *   kind 0: eps = p0 + i p1
 *   kind 1: diag(p0+i p1, p2+i p3, p4+i p5)
 *   kind 2: eps(x) = p0 + p2 sin(pi x) cos(pi y) + i p1
 *   kind 3: Janus, z > 0: p0 + i p1, otherwise p2 + i p3
 * kind 4: a compiled SVTensor / MatProp (bfa_material_svtensor,
 *   bfa_material_expr): expr is the device program, released by
 *   bfa_material_release.                                                  */
typedef struct { int kind; double p[6]; const void *expr; } bfa_material;

/* Plane wave E = E0 exp(i omega dir.x); E0 complex (re,im per component).
 * Stands in for SCUFF-EM's PlaneWave (buff-scatter.cc:186).               */
typedef struct { double E0[6]; double dir[3]; } bfa_planewave;

const char *bfa_last_error(void);
const char *bfa_status_string(int status);
int bfa_version(void);

/* ---- mesh construction (host only) ------------------------------------ */
/* Kuhn-subdivided cube [-1,1]^3 with n^3 subcubes x 6 tetrahedra, radially
 * mapped onto a sphere of radius R, interior nodes jittered by
 * U(-jitter,jitter)*(2R/n) per coordinate from splitmix64(seed); tetrahedra
 * positively oriented.  Pass V/tetVI == NULL to query nv/nt.             */
int bfa_mesh_kuhn_sphere(int n, double radius, double jitter, uint64_t seed,
                         int *nv, int *nt, double *V, int *tetVI);
/* GMSH .msh/.vmsh reader with 1e-6 vertex merging
 * (src/libs/libbuff/ReadGMSHFile.cc:66-261).  V/tetVI NULL: size query.   */
int bfa_mesh_read_vmsh(const char *path, int *nv, int *nt, double *V, int *tetVI);

/* ---- geometry: SWGGeometry/SWGVolume (SWGGeometry.cc:214-383,
 *      SWGVolume.cc:90-181) --------------------------------------------- */
int  bfa_geometry_create(int nv, const double *V, int nt, const int *tetVI,
                         bfa_geometry **out);
void bfa_geometry_destroy(bfa_geometry *g);
int  bfa_geometry_info(const bfa_geometry *g, int *nv, int *nt, int *nbf,
                       int *nfaces_total);
/* face tables (any pointer may be NULL), sizes nfaces_total (x3 for fV)   */
int  bfa_geometry_faces(const bfa_geometry *g, int *fV, int *fQP, int *fQM,
                        int *fPT, int *fMT, int *fPI, int *fMI, double *fArea);
/* touching BF pairs a<=b (CompareBFs > 0, SWGVolume.cc:295-344) as CSR    */
int  bfa_geometry_touch_count(const bfa_geometry *g, int64_t *nnz);
int  bfa_geometry_touch_pairs(const bfa_geometry *g, int *rowptr, int *cols);
/* upload topology and cubature tables to the current HIP device           */
int  bfa_geometry_upload(bfa_geometry *g, void *stream);

/* ---- cold step: frequency-independent static table ------------------- */
/* One record of 6 doubles per touching pair, the FIBBI data of
 * FIBBICache::GetFIBBIData / ComputeFIBBIData (FIBBICache.cc:211-245,
 * GMatrixElements.cc:295-375); Taylor-Duffy order td_order (<=0: default). */
int  bfa_static_build(bfa_geometry *g, int td_order, void *stream);
int  bfa_static_download(const bfa_geometry *g, double *table /* nnz*6 */);
/* records supplied by the caller (nnz*6, touching-pair CSR order); marks
 * the static table built without the cold step                             */
int  bfa_static_upload(bfa_geometry *g, const double *table);

/* ---- materials and geometry files -------------------------------------- */
/* SVTensor file or MatProp name -> kind 4 material on the current device.
 * is_matprop != 0: `name` is a MatProp (MATERIAL keyword of .buffgeo):
 * CONST_EPS_<complex> or a MATERIAL...ENDMATERIAL section of the file in
 * BUFF_MATPROP_FILE; otherwise an SVTensor file found as given, then on
 * BUFF_SVTENSOR_PATH (SVTensor.cc:55-127, 162-285).                        */
int  bfa_material_svtensor(const char *name, int is_matprop, bfa_material *out);
/* isotropic eps given by one expression in w, x, y, z, r, Theta, Phi      */
int  bfa_material_expr(const char *eps_expr, bfa_material *out);
int  bfa_material_release(bfa_material *m);
/* parse and compile only (host, no device): 0 or the error of the above    */
int  bfa_material_check(const char *name, int is_matprop);
int  bfa_material_expr_check(const char *eps_expr);
/* eps tensors at npts points (host xyz[3*npts]) -> Q[npts][3][3] complex
 * (re, im interleaved, row-major), evaluated on the device                 */
int  bfa_material_eval(const bfa_material *m, double omega, int npts, const double *pts, double *Q);
/* .buffgeo file (SWGGeometry.cc:52-383): one OBJECT with MESHFILE, MATERIAL
 * or SVTENSOR, DISPLACED / ROTATED; MESHPATH.  Two calls: V == T == NULL
 * returns the sizes; then V, T (transformed mesh) and, if mat != NULL, the
 * compiled material.                                                       */
int  bfa_buffgeo_load(const char *path, int *nv, int *nt, double *V, int *tetVI, bfa_material *mat);

/* ---- FIBBI cache files (FIBBICache.cc:247-463) ------------------------- */
/* "FIBBI_GCACHE\0" + records {27-float key, 6-double record}; the key of a
 * touching pair a<=b is GetFIBBICacheKey's (FIBBICache.cc:171-203).        */
/* keys[nnz*27] of the touching pairs in CSR order                          */
int  bfa_fibbi_keys(const bfa_geometry *g, float *keys);
/* raw file access: with keys == data == NULL, *nrec receives the count;
 * otherwise *nrec is the room (records) and receives the count             */
int  bfa_fibbi_read(const char *path, int64_t *nrec, float *keys, double *data);
int  bfa_fibbi_write(const char *path, int64_t nrec, const float *keys, const double *data);
/* FIBBICache::Store: each distinct key once (first pair's record)          */
int  bfa_fibbi_store(const bfa_geometry *g, const char *path, int64_t *written);
/* FIBBICache::PreLoad then lookup of every touching pair: hits take the
 * file's record; misses are computed by bfa_static_build(td_order); no
 * misses skips the cold step                                               */
int  bfa_fibbi_preload(bfa_geometry *g, const char *path, int td_order, int64_t *hits,
                       int64_t *misses, void *stream);

/* ---- warm step: VIE matrix --------------------------------------------- */
/* SWGGeometry::AssembleVIEMatrix (VIEMatrix.cc:431-460): d_M (N x N,
 * complex128, ld >= N) receives G + chi^-1 block, both triangles.        */
int  bfa_assemble_vie(bfa_geometry *g, double omega_re, double omega_im,
                      const bfa_material *mat, void *d_M, int64_t ld, void *stream);

/* SWGGeometry::AssembleRHSVector (RHSVector.cc:85-124) for npw incident
 * plane waves, one column each: d_B is N x npw complex128, ldb >= N.      */
/* Same, with the far-field (non-touching) cubature evaluated and summed in
 * FP32 (the precision form's far field); touching pairs, static records and
 * the Gram block stay FP64.  Entries are stored as complex128.             */
/* SWGGeometry::AssembleGBlock(no, no, ...) (VIEMatrix.cc:365-426): the
 * self G block, i.e. the VIE matrix without the chi^-1 overlap terms        */
int  bfa_assemble_g(bfa_geometry *g, double omega_re, double omega_im, void *d_M, int64_t ld,
                    void *stream);
/* SWGGeometry::AssembleOverlapBlocks (VIEMatrix.cc:321-360, GetOverlaps
 * :200-316): row a has cnt[a] <= 7 entries at [7a, 7a + cnt[a]) of cols,
 * V, VInv (complex, re/im) and Rytov (real); temperature (NULL: none) is a
 * compiled material whose eps_xx real part is T(x), t_env the environment
 * temperature (buff-neq units: Theta = omega / (exp(omega / (k T)) - 1),
 * k = 4.36763e-4).  Host outputs.                                          */
int  bfa_overlap_blocks(bfa_geometry *g, double omega, const bfa_material *eps,
                        const bfa_material *temperature, double t_env, int *cnt, int *cols,
                        double *V, double *VInv, double *Rytov);
/* buff-neq (one object).  S^-1, S the SWG overlap (Gram) matrix of
 * GetOverlapElements (SWGVolume.cc:405-480), as CreateBNEQData builds and
 * inverts it once per geometry (CreateBNEQData.cc:119-142): d_Sinv is
 * N x N complex128, ld >= N.                                                */
int  bfa_gram_inverse(bfa_geometry *g, void *d_Sinv, int64_t ld, void *stream);
/* ComputeDressedRytovMatrix (buff-neq GetFlux.cc:155-262) at omega:
 * DR = W (S^-1 R S^-1)^H with W = (1 + S^-1 V S^-1 G)^-1, G the G block,
 * V and R the overlap blocks of bfa_overlap_blocks (same eps, temperature,
 * t_env).  d_DR: N x N complex128 out (ld >= N); d_work: device memory of
 * bfa_dressed_rytov_work_bytes(N, ld) bytes (the reference's WorkMatrix).   */
size_t bfa_dressed_rytov_work_bytes(int n, int64_t ld);
int  bfa_dressed_rytov(bfa_geometry *g, double omega, const bfa_material *eps,
                       const bfa_material *temperature, double t_env, const void *d_Sinv,
                       void *d_DR, int64_t ld, void *d_work, void *stream);
/* GetPFTMatrix(IF = 0, Omega, {EMT, DRMatrix}) of GetFlux (GetPFT.cc:108-115,
 * EMTPFT.cc:389-476): out[0] = P_abs, out[1] = P_scat (= -P_abs) from
 * sum over a <= b of Re(Z0 DR[b,a]) Im Q_ab.  Host output.                 */
int  bfa_emt_power_dr(bfa_geometry *g, double omega, const void *d_DR, int64_t ld, double *out,
                      void *stream);
/* bfa_assemble_vie with the far-field kernel evaluated as the precision
 * form's split (AssembleVIEMatrix, VIEMatrix.cc:431-460, with the fixed rule
 * of GMatrixElements.cc:270-288): about each tetrahedron pair's centroid
 * separation, the second-order Taylor part of e^{ikr}/r in FP64 and the
 * remainder in FP32; the matrix moves by ~1e-9 relative from the FP64 rule
 * (powers within ~1e-9).                                                    */
int  bfa_assemble_vie_far32(bfa_geometry *g, double omega_re, double omega_im,
                            const bfa_material *mat, void *d_M, int64_t ld, void *stream);
int  bfa_assemble_rhs(bfa_geometry *g, double omega, const bfa_planewave *pw,
                      int npw, void *d_B, int64_t ldb, void *stream);

/* ---- dense solve: HMatrix::LUFactorize / LUSolve ----------------------- */
/* In-place partial-pivot LU of a complex128 matrix (zgetrf semantics,
 * buff-scatter.cc:285).  d_ipiv: n ints on device.  info (host, may be
 * NULL): 0, or k+1 if U(k,k) is exactly zero.                              */
int  bfa_lu_factor_z(void *d_A, int n, int64_t lda, int *d_ipiv, int *info,
                     void *stream);
/* LUSolve(HVector*)/LUSolve(HMatrix*) (buff-scatter.cc:314,
 * buff-neq GetFlux.cc:265): in-place multi-RHS solve of d_B (n x nrhs).   */
int  bfa_lu_solve_z(const void *d_A, int n, int64_t lda, const int *d_ipiv,
                    void *d_B, int nrhs, int64_t ldb, void *stream);

/* bfa_lu_solve_z without any device allocation (no reference counterpart):
 * for callers that cannot call hipMalloc -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation.  Same kernels and
 * results as bfa_lu_solve_z; its scratch is one static device array per GPU
 * (47.6 MiB, kept until the process ends), so n <= BFA_NOALLOC_MAX_N and
 * nrhs <= BFA_NOALLOC_MAX_NRHS (larger sizes return BFA_ERR_ARG), and calls
 * that overlap on one GPU race on that array.  Synchronous like
 * bfa_lu_solve_z.                                                          */
#define BFA_NOALLOC_MAX_N    20480
#define BFA_NOALLOC_MAX_NRHS 8
int  bfa_lu_solve_z_noalloc(const void *d_A, int n, int64_t lda, const int *d_ipiv,
                            void *d_B, int nrhs, int64_t ldb, void *stream);

/* LU execution on this device (no reference counterpart): on = 1 (default)
 * runs the look-ahead LU on CU-partitioned streams (the panel on 64 CUs,
 * the trailing updates on the rest; BFA_LU_MASK sets the count, 0 disables);
 * on = 0 keeps every stream on every CU, for applications that partition
 * the GPU themselves.                                                     */
int  bfa_lu_partition(int on);

/* ---- precision form (no reference counterpart) -------------------------- */
/* complex128 -> complex64 copy of an m x n column-major matrix             */
int  bfa_convert_z2c(const void *d_A64, int m, int n, int64_t lda, void *d_A32,
                     int64_t lda32, void *stream);
/* complex64 LU / solve: the same blocked algorithm on FP32 MFMA             */
int  bfa_lu_factor_c(void *d_A, int n, int64_t lda, int *d_ipiv, int *info, void *stream);
int  bfa_lu_solve_c(const void *d_A, int n, int64_t lda, const int *d_ipiv,
                    void *d_B, int nrhs, int64_t ldb, void *stream);
/* complex64 LU with pivoting restricted to each 256-row diagonal block and
 * the trailing updates in split FP16 on the matrix cores (each FP32 operand,
 * scaled by a power of two per row / column, as hi + lo FP16 halves; three
 * FP16 products per real product, FP32 accumulation): the preconditioner of
 * the fast refinement form (bfa_solve_refine with these factors).  Not a
 * partial-pivoting LU: meant for matrices that factor stably without
 * pivoting across blocks (the VIE matrices do), with FP64 refinement
 * restoring the accuracy.  d_ipiv and info as bfa_lu_factor_c.             */
int  bfa_lu_factor_c_local(void *d_A, int n, int64_t lda, int *d_ipiv, int *info, void *stream);
/* C(m x n) -= A(m x k) * B(k x n), complex64, k <= 512, by the split-FP16
 * products of bfa_lu_factor_c_local's trailing update.                     */
int  bfa_cgemm_sub_split(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B,
                         int64_t ldb, void *d_C, int64_t ldc, void *stream);
/* Solve M64 x = b with the complex64 LU factors of M64 and FP64 iterative
 * refinement (residual b - M64 x in FP64 against the resident FP64 matrix)
 * until max_j ||r_j||/||b_j|| <= tol or maxit iterations.  d_B: RHS in,
 * solution out (complex128).  iters/resid (host, may be NULL) report the
 * refinement steps taken and the final relative residual.                  */
int  bfa_solve_refine(const void *d_M64, int64_t ldm, const void *d_LU32, int64_t ldlu,
                      const int *d_ipiv, int n, void *d_B, int nrhs, int64_t ldb,
                      double tol, int maxit, int *iters, double *resid, void *stream);

/* Dense complex128 product C(m x n) -= A(m x k) * B(k x n), column-major;
 * the MFMA kernel of the LU trailing update (HMatrix::Multiply-style use).  */
int  bfa_zgemm_sub(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B,
                   int64_t ldb, void *d_C, int64_t ldc, void *stream);
/* Same for complex64 (the FP32 MFMA kernel of the precision form).        */
int  bfa_cgemm_sub(int m, int n, int k, const void *d_A, int64_t lda, const void *d_B,
                   int64_t ldb, void *d_C, int64_t ldc, void *stream);

/* ---- power (GetPFTMatrix with EMT, GetPFT.cc:133-189, EMTPFT.cc:324-546) */
/* out[2*j] = P_abs, out[2*j+1] = P_scat for RHS column j (host array).    */
int  bfa_emt_power(bfa_geometry *g, double omega, const bfa_planewave *pw,
                   int npw, const void *d_J, int64_t ldj, double *out, void *stream);
/* SWGGeometry::GetFields (GetFields.cc:288-382): total E and H at npts
 * points X[3*npts] (host) from the incident plane waves (summed, npw may be
 * 0) and the currents d_J (N complex128 on the device, or NULL for the
 * incident field alone); F[npts][6] complex (Ex Ey Ez Hx Hy Hz, re/im
 * interleaved) on the host.                                                */
int  bfa_get_fields(bfa_geometry *g, double omega, const bfa_planewave *pw, int npw,
                    const void *d_J, int npts, const double *X, double *F, void *stream);

/* ---- batching form (no reference counterpart: the reference solves one
 * problem per process, one frequency at a time, buff-scatter.cc:236-363) --
 * A batch holds `count` independent instances (each an uploaded geometry
 * with its static table built, and a material) excited by the same `npw`
 * plane waves (npw <= 8).  Per frequency every instance is assembled
 * (complex128), converted to a complex64 copy padded to a multiple of 256,
 * LU-factored on FP32 MFMA and solved with FP64 iterative refinement
 * against its complex128 matrix, then its EMT power is computed -- all in
 * grouped launches covering every instance, replayed per frequency from
 * one captured hipGraph when use_graph != 0.  Device memory per instance:
 * N*Np*16 + Np*Np*8 bytes (Np = N rounded up to 256) plus O(N) vectors.   */
typedef struct bfa_batch bfa_batch;
int  bfa_batch_create(int count, bfa_geometry *const *geoms, const bfa_material *mats,
                      const bfa_planewave *pw, int npw, bfa_batch **out);
int  bfa_batch_destroy(bfa_batch *b);
/* count, npw, and per instance N[count], Np[count] (host, may be NULL).    */
int  bfa_batch_info(const bfa_batch *b, int *count, int *npw, int *N, int *Np);
/* Autotuner: times each GEMM tile variant on the batch's own grouped
 * trailing updates, per size bucket floor(log2 m), and keeps the fastest
 * (reps timed launches each).  Without it a fixed rule picks the tile.    */
int  bfa_batch_tune(bfa_batch *b, int reps);
/* tile[16]: chosen variant per bucket (-1 = untuned; 0 = 64x64, 1 =
 * 128x128 complex64 on FP32 MFMA, 2 = split FP16 on the FP16 matrix cores,
 * timed with its plane launches); ms[16*3]: time per bucket and variant.  */
int  bfa_batch_tuning(const bfa_batch *b, int *tile, double *ms);
/* tile[16] as bfa_batch_tuning reports it (-1 = the fixed rule): a stored
 * tuning, or one variant for every bucket (no reference counterpart).      */
int  bfa_batch_set_tiles(bfa_batch *b, const int *tile);
/* Sweep nfreq (<= 256) real frequencies.  Host outputs (may be NULL):
 * power[((f*count + i)*npw + j)*2 + {0,1}] = {P_abs, P_scat};
 * iters[f*count + i], resid[f*count + i]: refinement rounds and final
 * relative residual (refinement stops at resid <= tol or maxit rounds);
 * info[f*count + i]: 0, or the 1-based column of the first zero pivot.
 * With power == NULL the EMT pass is skipped.  Synchronous.               */
int  bfa_batch_run(bfa_batch *b, const double *omega, int nfreq, int use_graph,
                   double tol, int maxit, double *power, int *iters, double *resid, int *info);
/* ms[6] for the last run, summed over its frequencies: [0] device time of
 * assembly + convert + factor + solve; [1..4] per phase (assembly,
 * convert+factor, solve, power; eager runs only, else 0); [5] power.       */
int  bfa_batch_timing(const bfa_batch *b, double *ms);
/* Host copies of the last frequency's solution J (N x npw, complex128,
 * column-major) and complex128 VIE matrix (N x N) of instance `inst`.      */
int  bfa_batch_solution(const bfa_batch *b, int inst, void *J);
int  bfa_batch_matrix(const bfa_batch *b, int inst, void *M);

/* ---- learned near-field surrogate: MLP kernels (no reference counterpart)
 * Ensemble of E networks nin -> hidden -> hidden -> hidden -> nout, GELU,
 * FP16 weights/activations into FP32-accumulating MFMA; parameter block per
 * member (FP32, row-major, kp = nin rounded up to 16):
 *   W1[hidden][kp] b1 W2[hidden][hidden] b2 W3[hidden][hidden] b3 W4[nout][hidden] b4 */
int  bfa_mlp_param_count(int nin, int hidden, int nout, int64_t *count, int *kp);
/* One forward + MSE + backward pass on host data: params [E][count],
 * X [E][B][nin], T [E][B][nout]; outputs (may be NULL) loss[E],
 * grad [E][count], Y [E][B][nout].  hidden <= 256 (multiple of 64),
 * nout a multiple of 16.                                                   */
int  bfa_mlp_grad(int nin, int hidden, int nout, int E, const float *params, int B,
                  const float *X, const float *T, double *loss, float *grad, float *Y);

/* ---- learned near-field surrogate (no reference counterpart) -----------
 * Over the non-touching tetrahedron pairs of one geometry whose centroids
 * are closer than `factor` x the larger diameter, a table of 4x4 complex
 * blocks G_ij = int_S int_T [f_i.g_j - div f_i div g_j/k^2] e^{ikr}/(4 pi r)
 * (face i of S, face j of T) feeds the far-field assembly.  The blocks come
 * from the reference's 4-point rule, from cubature (product TetCR rule of
 * npts in {4,16,33} on tetrahedra subdivided `level` in {0,1,2} times), or
 * from an ensemble of MLPs (52 inputs: the target's 4 vertices and the
 * source's shape in the source frame over the source's mean edge, log(k h),
 * the unit vector and log10 relative size of R16 - R4 (16- and 4-point
 * product rules); 32 outputs: R16's remaining error in units of
 * ||R16 - R4||) added to R16, with cubature fallback.                       */
typedef struct bfa_surrogate bfa_surrogate;
int  bfa_surr_create(bfa_geometry *g, double factor, int ensemble, int hidden, uint64_t seed,
                     bfa_surrogate **out);
int  bfa_surr_destroy(bfa_surrogate *s);
/* npairs, and the pair list (tetrahedron indices, host, may be NULL).      */
int  bfa_surr_pairs(const bfa_surrogate *s, int *npairs, int *ps, int *pt);
/* Cubature blocks of pair[n] at omega[n] (host): blocks[n][32] (re, im of
 * G_00, G_01, ..., G_33).                                                  */
int  bfa_surr_cubature(bfa_surrogate *s, const int *pair, const double *omega, int n, int npts,
                       int level, double *blocks);
/* Geometric network inputs of pair[n] at omega[n]: feat[n][19] (host).     */
int  bfa_surr_features(bfa_surrogate *s, const int *pair, const double *omega, int n, float *feat);
/* Training set: every pair at per_pair seeded log-uniform frequencies in
 * [wmin, wmax], labelled by cubature (label_npts, label_level).             */
int  bfa_surr_dataset(bfa_surrogate *s, int per_pair, double wmin, double wmax, int label_npts,
                      int label_level);
/* Adam with cosine step decay; loss_hist[epochs][ensemble] (host, may be NULL). */
int  bfa_surr_train(bfa_surrogate *s, int epochs, int batch, double lr, double *loss_hist);
/* Block table at omega: mode 0 4-point rule, 1 cubature (npts, level),
 * 2 surrogate with cubature (npts, level) where the ensemble disagreement
 * exceeds tol or an input lies outside the training range.  Disagreement:
 * the largest member's distance from the ensemble mean, in units of
 * ||R16 - R4||, the size of the residual the network corrects.            */
int  bfa_surr_table(bfa_surrogate *s, double omega, int mode, double tol, int npts, int level,
                    int *nfallback, double *max_disagree);
int  bfa_surr_table_download(const bfa_surrogate *s, double *blocks /* npairs*32 */);
/* The inverse: a table in that layout from the host (no reference
 * counterpart: lets a caller reuse or edit a table).                        */
int  bfa_surr_table_upload(bfa_surrogate *s, const double *blocks /* npairs*32 */);
/* bfa_assemble_vie with the far field read from the block table.           */
int  bfa_assemble_vie_table(bfa_surrogate *s, double omega, const bfa_material *mat, void *d_M,
                            int64_t ld, void *stream);
/* ensemble size, parameters per member, training samples; stats[4*52+2*32]
 * = input mean, sd, lower, upper bound; target mean, sd.                   */
int  bfa_surr_info(const bfa_surrogate *s, int *ensemble, int64_t *nparam, int *nsamples, float *stats);
/* Get (out) and/or set (in) the FP32 parameters [ensemble][nparam].        */
int  bfa_surr_params(bfa_surrogate *s, float *out, const float *in);

/* ---- instrumentation (no reference counterpart) ------------------------ */
/* When enabled, every trailing-update GEMM launch is bracketed by HIP
 * events; bfa_profile_read returns the summed kernel time (ms), the summed
 * algorithmic flops (8mnk per complex product) and the launch count for
 * which = 0 (complex128) or 1 (complex64), all launches; 2 / 3 the large
 * updates (m, n >= 256) on all CUs or, in the CU-partitioned LU, on the
 * update CUs; 4 / 5 the large updates run as column slices on the panel
 * CUs; reading 0 / 1 resets that type; 6 = the split-FP16 trailing updates
 * of bfa_lu_factor_c_local (flops counted as 8mnk; reading resets); 7 = the
 * algorithmic bytes (A and B operands once, C read and written, or written
 * where the product is stored) of the launches the last read of 6
 * returned, in *flops, with their count (ms = 0).                          */
int  bfa_profile_enable(int on);
int  bfa_profile_read(int which, double *ms, double *flops, int *launches);

/* Release the library's process-lifetime device objects (LU work areas,
 * the look-ahead's streams and events, instrumentation events, library
 * handles) on every device; call before the process tears down its GPU
 * runtime (the Python binding does so at exit).  Later calls re-create what
 * they need.  No reference counterpart.                                    */
int  bfa_shutdown(void);

/* ---- device memory helpers (for hosts without a tensor library) -------- */
int  bfa_device_malloc(void **ptr, size_t bytes);
int  bfa_device_free(void *ptr);
int  bfa_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int  bfa_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int  bfa_stream_synchronize(void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BUFFEM_AMD_H */
