"""bfa_lu_solve_z_noalloc: the LU solve with the library's static scratch
array (no device allocation) against bfa_lu_solve_z on the same factors.

Two calls in a row through the C ABI, the second smaller (the array keeps
the first call's contents, so whatever must start at zero has to be cleared
by the call itself): the N = 19,872 VIE matrix of the n = 12 Kuhn sphere
with its two plane waves, then the n = 5 sphere (N = 1,350) with one.  The
result must equal the allocating entry point's to 1e-13 relative (the same
kernels in the same order give the same bits) and solve the system.  Sizes
beyond the array are rejected before any device work.
"""
import numpy as np
import pytest

from conftest import kuhn

pytestmark = pytest.mark.gpu


def _case(bfa, torch, n_sub, eps, spec):
    V, T = kuhn(bfa, n_sub)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    N = g.N
    dM = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    dB = torch.empty((len(spec), N), dtype=torch.complex128, device="cuda")
    dp = torch.empty(N, dtype=torch.int32, device="cuda")
    g.assemble_vie(1.0, bfa.material("diag", eps=eps), dM)
    g.assemble_rhs(1.0, bfa.planewaves(spec), dB)
    M0 = dM.clone()
    assert bfa.lu_factor(dM, N, dp) == 0
    return N, M0, dM, dB, dp


def test_noalloc_solve_matches_allocating_solve(bfa, torch_cuda):
    torch = torch_cuda
    eps = (4 + 0.05j, 5 + 0.05j, 6 + 0.05j)
    cases = [(12, [((1, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 0, 1))]), (5, [((1, 0, 0), (0, 0, 1))])]
    for n_sub, spec in cases:
        N, M0, LU, B, piv = _case(bfa, torch, n_sub, eps, spec)
        nrhs = len(spec)
        print(f"[noalloc] N={N} nrhs={nrhs}", flush=True)
        X_ref = B.clone()
        bfa.lu_solve(LU, N, piv, X_ref, nrhs)
        X = B.clone()
        bfa.lu_solve_noalloc(LU, N, piv, X, nrhs)
        torch.cuda.synchronize()
        d = (torch.linalg.norm(X - X_ref) / torch.linalg.norm(X_ref)).item()
        same = bool(torch.equal(X, X_ref))
        # residual of the original system: M x = b (M0[c, r] = M(r, c))
        R = M0.T.contiguous() @ X.T.contiguous() - B.T
        res = (torch.linalg.norm(R) / torch.linalg.norm(B)).item()
        print(f"[noalloc] N={N}: rel diff vs bfa_lu_solve_z {d:.1e} (bit-identical {same}), residual {res:.1e}",
              flush=True)
        assert d <= 1e-13
        assert res < 1e-10
        del M0, LU, B, piv, X, X_ref, R
        torch.cuda.empty_cache()


def test_noalloc_rejects_sizes_beyond_the_array(bfa, torch_cuda):
    torch = torch_cuda
    n = 20481
    A = torch.zeros(16, dtype=torch.complex128, device="cuda")
    piv = torch.zeros(16, dtype=torch.int32, device="cuda")
    with pytest.raises(bfa.BFAError, match="exceed the static scratch"):
        bfa.lu_solve_noalloc(A, n, piv, A, 1)
    with pytest.raises(bfa.BFAError, match="exceed the static scratch"):
        bfa.lu_solve_noalloc(A, 100, piv, A, 9)
