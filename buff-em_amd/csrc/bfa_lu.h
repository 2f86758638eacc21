// bfa_lu.h -- internal entry points of the dense complex LU / solve.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace bfa {
// C(m x n) -= A(m x k) * B(k x n), complex128 (z) or complex64 (c)
int gemm_sub_z(int m, int n, int k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc,
               hipStream_t st);
int gemm_sub_c(int m, int n, int k, const void *A, int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc,
               hipStream_t st);
int lu_factor_z(void *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb);
int lu_factor_c(void *A, int n, int64_t lda, int *ipiv, int *info, hipStream_t st, int nb);
size_t lu_solve_scratch_bytes(int n, int nrhs, size_t elem);

int lu_solve_z(const void *A, int n, int64_t lda, const int *ipiv, void *B, int nrhs, int64_t ldb, hipStream_t st,
               void *scratch);
int lu_solve_c(const void *A, int n, int64_t lda, const int *ipiv, void *B, int nrhs, int64_t ldb, hipStream_t st,
               void *scratch);
// the same solve in two steps: permutation and inverted diagonal blocks
// once per factorisation, then any number of applications (complex64)
int lu_solve_prep_c(const void *A, int n, int64_t lda, const int *ipiv, hipStream_t st, void *scratch);
int lu_solve_apply_c(const void *A, int n, int64_t lda, void *B, int nrhs, int64_t ldb, hipStream_t st, void *scratch);
// nonzero timeout flag of the streaming solves that used scratch -> error
int lu_solve_check(const void *scratch, int n, hipStream_t st);

// complex64 pieces shared with the block-local LU (bfa_lu_local.hip):
// clear the LU's device flags; read them back (timeout -> error, first zero
// pivot -> info); row interchanges k0..k0+kb-1 on columns [c0, c1);
// B(kb x n) := L^-1 B with L unit lower
int lu_local_begin(hipStream_t st);
int lu_local_end(int *info, hipStream_t st);
void laswp_c(void *A, int64_t lda, int c0, int c1, int k0, int kb, const int *ipiv, hipStream_t st);
void trsm_lower_c(int kb, int n, const void *L, int64_t ldl, void *B, int64_t ldb, hipStream_t st);
// B(kb x n) := U^-1 B, U upper with its diagonal, kb <= 256 (one launch)
void trsm_upper_c(int kb, int n, const void *U, int64_t ldu, void *B, int64_t ldb, hipStream_t st);
// the LU's device flags: [0] barrier timeout, [1] first zero pivot (1-based)
int *lu_local_flags();
// GEMM instrumentation (bfa_profile_enable): whether it is on, and the
// split-FP16 trailing updates' summed kernel time / flops / launches (reset)
bool profile_on();
// release the process-lifetime device objects of the LU (bfa_shutdown)
void lu_release();
void local_release();
int profile_read_split(double *ms, double *flops, int *launches);
// the algorithmic bytes and the launch count of the split launches returned
// by the last profile_read_split
int profile_split_bytes(double *bytes, int *launches);
}  // namespace bfa
