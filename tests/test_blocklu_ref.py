"""Pins of the block-local LU restatement used by the fast form's pivot
test (tests/blocklu.py): with a single block it is LAPACK's partial pivoting
(zgetrf through scipy: same pivots, same factors), and with 256-row blocks
its factors reproduce P A."""
import numpy as np
import scipy.linalg

from blocklu import block_local_lu


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 10.0 ** rng.uniform(0, 2, (n, n))


def test_one_block_is_zgetrf():
    A = _matrix(300, 1)
    ipiv, _, LU = block_local_lu(A, nb=300)
    lu, piv = scipy.linalg.lu_factor(A)
    assert np.array_equal(ipiv, piv)
    assert np.abs(LU - lu).max() <= 1e-12 * np.abs(lu).max()


def test_blocks_factor_pa():
    n = 600
    A = _matrix(n, 2)
    ipiv, _, LU = block_local_lu(A)
    assert all(i // 256 == ipiv[i] // 256 and ipiv[i] >= i for i in range(n))
    PA = A.copy()
    for i in range(n):
        if ipiv[i] != i:
            PA[[i, ipiv[i]], :] = PA[[ipiv[i], i], :]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.linalg.norm(PA - L @ U) <= 1e-12 * np.linalg.norm(A)
