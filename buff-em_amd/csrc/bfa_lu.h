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
}  // namespace bfa
