"""Float64 restatement of the block-local pivoting LU (test helper).

bfa_lu_factor_c_local has no counterpart in the reference, which factors
with LAPACK zgetrf (HMatrix::LUFactorize, buff-scatter.cc:285); with one
block covering the whole matrix the algorithm below is zgetrf's partial
pivoting, which tests/test_blocklu_ref.py checks against scipy's LAPACK.
"""
import numpy as np


def block_local_lu(A, nb=256):
    """Block-local pivoting LU in float64 (the algorithm of
    bfa_lu_factor_c_local): per nb-row diagonal block, partial pivoting over
    the block's own rows with izamax's rule (largest |re| + |im|, first row
    among equals), interchanges applied to whole rows, then the block row and
    column and the trailing update.  Returns (ipiv, gap): gap[i] is the
    relative distance between the best and second-best candidate of column i
    (a near tie there may legitimately resolve the other way in FP32)."""
    A = A.copy()
    n = A.shape[0]
    ipiv = np.arange(n)
    gap = np.full(n, np.inf)
    for j in range(0, n, nb):
        jb = min(nb, n - j)
        for c in range(j, j + jb):
            mag = np.abs(A[c:j + jb, c].real) + np.abs(A[c:j + jb, c].imag)
            p = c + int(np.argmax(mag))
            if len(mag) > 1:
                top = np.sort(mag)[-2:]
                gap[c] = (top[1] - top[0]) / top[1]
            ipiv[c] = p
            if p != c:
                A[[c, p], :] = A[[p, c], :]
            A[c + 1:j + jb, c] /= A[c, c]
            A[c + 1:j + jb, c + 1:j + jb] -= np.outer(A[c + 1:j + jb, c], A[c, c + 1:j + jb])
        if j + jb < n:
            L11 = np.tril(A[j:j + jb, j:j + jb], -1) + np.eye(jb)
            U11 = np.triu(A[j:j + jb, j:j + jb])
            A[j:j + jb, j + jb:] = np.linalg.solve(L11, A[j:j + jb, j + jb:])
            A[j + jb:, j:j + jb] = np.linalg.solve(U11.T, A[j + jb:, j:j + jb].T).T
            A[j + jb:, j + jb:] -= A[j + jb:, j:j + jb] @ A[j:j + jb, j + jb:]
    return ipiv, gap, A


