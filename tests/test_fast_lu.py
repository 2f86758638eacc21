"""The fast refinement form's factorisation: complex64 LU with pivoting inside
each 256-row diagonal block and split-FP16 trailing updates on the matrix
cores, as the preconditioner of FP64 iterative refinement.

- bfa_cgemm_sub_split (C -= A B from hi + lo FP16 halves of row- and
  column-scaled operands) against a complex128 product on the host:
  normwise error at the split's 22-bit level, on ragged shapes, k = 1..512,
  and rows/columns spanning 1e-8..1e8 in magnitude (the per-row / per-column
  power-of-two scaling keeps every row's relative accuracy).
- bfa_lu_factor_c_local on random matrices (n = 3,000 and 8,100, not
  multiples of 256; the larger one through the paired k = 512 trailing
  updates): the solve's backward error at complex64 level; on a matrix that
  pivots in most columns (n = 600): the first block's pivots equal the
  float64 restatement's (tests/blocklu.py), every pivot dominates its
  column's candidates, and P A = L U to the split products' accuracy.
- On VIE matrices (n = 5 and n = 12 Kuhn spheres, the second the N = 19,872
  headline workload): FP64 refinement with these factors reaches a 1e-10
  residual and its powers equal the FP64 path's to 1e-9.
"""
import numpy as np
import pytest

from blocklu import block_local_lu
from conftest import kuhn

pytestmark = pytest.mark.gpu


def _c64(torch, X):
    return torch.from_numpy(np.ascontiguousarray(X.T).astype(np.complex64)).cuda()


def test_cgemm_split(bfa, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    shapes = [(16, 16, 32), (70, 33, 21), (200, 130, 256), (300, 260, 100), (129, 255, 1), (1000, 700, 256),
              (2048, 384, 256), (500, 300, 300), (1100, 640, 512)]
    for (m, n, k) in shapes:
        A = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(np.complex64)
        B = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex64)
        C = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
        for scaled in (False, True):
            if scaled:
                A = A * (10.0 ** rng.uniform(-8, 8, m))[:, None].astype(np.float32)
                B = B * (10.0 ** rng.uniform(-8, 8, n))[None, :].astype(np.float32)
                C = (A.astype(np.complex128) @ B.astype(np.complex128)).astype(np.complex64)
            print(f"[split gemm] {m}x{n}x{k} scaled={scaled}", flush=True)
            ref = C.astype(np.complex128) - A.astype(np.complex128) @ B.astype(np.complex128)
            # error per entry relative to the size of the products that form it
            scale = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64) + np.abs(C)
            errs = []
            for f in (bfa.cgemm_sub_split, bfa.cgemm_sub):
                dA, dB, dC = _c64(torch, A), _c64(torch, B), _c64(torch, C)
                f(m, n, k, dA, m, dB, k, dC, m)
                torch.cuda.synchronize()
                got = dC.cpu().numpy().T.astype(np.complex128)
                errs.append((np.abs(got - ref) / scale).max())
            print(f"  max entrywise error / (|A||B| + |C|): split FP16 {errs[0]:.2e}, FP32 MFMA {errs[1]:.2e}",
                  flush=True)
            # FP32 accumulation sets both; the 22-bit operands add little
            assert errs[0] < max(4 * errs[1], 2e-6)


@pytest.mark.parametrize("n", [3000, 8100])
def test_lu_local_random(bfa, torch_cuda, n):
    """n = 8,100 takes the paired trailing updates (k = 512) in its first
    steps, n = 3,000 only single ones."""
    torch = torch_cuda
    nrhs = 3
    gen = torch.Generator(device="cuda:0").manual_seed(n)
    A = torch.randn((n, n), dtype=torch.complex128, device="cuda:0", generator=gen)
    A += n ** 0.5 * torch.eye(n, dtype=torch.complex128, device="cuda:0")
    B = torch.randn((nrhs, n), dtype=torch.complex128, device="cuda:0", generator=gen)
    LU = A.to(torch.complex64).clone()
    piv = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    assert bfa.lu_factor_c_local(LU, n, piv) == 0
    p = piv.cpu().numpy()
    # pivots stay inside their 256-row block
    assert all(i // 256 == p[i] // 256 and p[i] >= i for i in range(n))
    X = B.to(torch.complex64).clone()
    bfa.lu_solve_c(LU, n, piv, X, nrhs)
    Am = A.T.contiguous()
    Xm = X.to(torch.complex128).T.contiguous()
    eta = (torch.linalg.norm(Am @ Xm - B.T) / (torch.linalg.norm(A) * torch.linalg.norm(Xm))).item()
    print(f"[lu local] n={n} backward error {eta:.2e}", flush=True)
    assert eta < 1e-6


@pytest.mark.gpu
def test_lu_local_pivots(bfa, torch_cuda):
    """Pivot choices of the diagonal-block kernel on a matrix that pivots in
    most columns (entry magnitudes spread over two decades), n = 600: blocks
    of 256, 256 and 88 rows (a last panel of 24 columns).
    - First block (FP32 arithmetic only): the pivots equal the float64
      restatement's up to the first column whose two best candidates lie
      within 1e-5 of each other.
    - Every column: the chosen pivot dominates the column's other candidates
      (izamax's rule, |re| + |im|) as the kernel's own factors show them
      (candidate at step d = L(i, d) U(d, d)); later blocks see operands
      rounded by the split products, so their pivots are not compared with
      the float64 ones.
    - P A = L U to the split products' accuracy."""
    torch = torch_cuda
    n = 600
    rng = np.random.default_rng(600)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 10.0 ** rng.uniform(0, 2, (n, n))
    ipiv_ref, gap, _ = block_local_lu(A)
    dA = torch.from_numpy(np.asfortranarray(A).T.copy()).to(torch.complex64).to("cuda:0")   # column-major
    piv = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    assert bfa.lu_factor_c_local(dA, n, piv) == 0
    p = piv.cpu().numpy()
    LU = dA.cpu().numpy().T.astype(np.complex128)
    tie = np.flatnonzero(gap[:256] < 1e-5)
    upto = int(tie[0]) if len(tie) else 256
    moved = int(np.sum(ipiv_ref[:upto] != np.arange(upto)))
    print(f"[lu pivots] n={n}: first block's pivots compared over {upto} columns ({moved} interchanges)", flush=True)
    assert moved > upto // 2
    assert np.array_equal(p[:upto], ipiv_ref[:upto])
    assert all(i // 256 == p[i] // 256 and p[i] >= i for i in range(n))
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    abs1 = lambda z: np.abs(z.real) + np.abs(z.imag)
    worst = 0.0
    for d in range(n):
        e = (d // 256) * 256 + 256
        cand = abs1(L[d + 1:e, d] * U[d, d])
        if len(cand):
            worst = max(worst, cand.max() / abs1(U[d, d]))
    print(f"[lu pivots] largest candidate / pivot over all columns: {worst:.6f}", flush=True)
    assert worst <= 1.0 + 1e-5
    PA = A.copy()
    for i in range(n):
        if p[i] != i:
            PA[[i, p[i]], :] = PA[[p[i], i], :]
    err = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
    # backward error of an LU with unit roundoff u: O(u) || |L||U| || / ||A||,
    # u = 2^-22 for the split-FP16 products (|L||U| ~ 700 ||A|| here: the
    # rows below a diagonal block are never pivots)
    bound = 2.0 ** -22 * np.linalg.norm(np.abs(L) @ np.abs(U)) / np.linalg.norm(A)
    print(f"[lu pivots] ||PA - LU|| / ||A|| = {err:.2e} (bound {bound:.2e})", flush=True)
    assert err < bound


@pytest.mark.parametrize("n_sub", [5, 12])
def test_lu_local_refine_vie(bfa, torch_cuda, n_sub):
    torch = torch_cuda
    V, T = kuhn(bfa, n_sub)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    N = g.N
    spec = [((1, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 0, 1))]
    pw = bfa.planewaves(spec)
    mat = bfa.material("diag", eps=(4 + 0.05j, 5 + 0.05j, 6 + 0.05j))
    dM = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    dB = torch.empty((2, N), dtype=torch.complex128, device="cuda")
    dp = torch.empty(N, dtype=torch.int32, device="cuda")
    g.assemble_vie(1.0, mat, dM)
    g.assemble_rhs(1.0, pw, dB)
    B0 = dB.clone()
    dM32 = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    bfa.convert_z2c(dM, N, N, dM32)
    assert bfa.lu_factor_c_local(dM32, N, dp) == 0
    it, res = bfa.solve_refine(dM, dM32, dp, N, dB, 2, tol=1e-10, maxit=30)
    P = g.emt_power(1.0, pw, dB)
    print(f"[lu local] N={N}: refinement {it} iterations, residual {res:.1e}", flush=True)
    assert res <= 1e-10
    del dM32
    # the FP64 path
    X = B0.clone()
    assert bfa.lu_factor(dM, N, dp) == 0
    bfa.lu_solve(dM, N, dp, X, 2)
    P64 = g.emt_power(1.0, pw, X)
    rel = np.abs(P - P64) / np.abs(P64)
    print(f"[lu local] N={N}: power vs FP64 path {rel.max():.1e}", flush=True)
    assert rel.max() < 1e-9
