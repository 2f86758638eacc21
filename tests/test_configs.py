"""The reference workloads' small configurations end to end against the oracle.

- n = 5 Kuhn sphere (750 tetrahedra, N = 1,350), eps = 4+0.1i, omega = 1,
  x-polarised plane wave along +z: FP64 matrix to 1e-10 (relative Frobenius),
  powers to 1e-9.
- The precision form on the same mesh with eps(x) = 4 + 2 sin(pi x) cos(pi y)
  + 0.1i and 32 right-hand sides (16 seeded directions uniform on the sphere
  x 2 orthonormal polarisations): complex64 LU + FP64 refinement to 1e-10
  residual; every power within 1e-8 of the FP64 oracle.  With the far field
  integrated in FP32 the error is set by the FP32 cubature, not by the
  solver: matrix 7.6e-8, worst power 4e-7 (median 1.5e-8), bounded at
  about twice that.
- An exactly singular matrix reports the LAPACK-style first zero pivot.
- A larger matrix (n = 14,000) factors through the CU-partitioned
  look-ahead, column slices of the bulk update included, with a backward
  error at the precision's level; the shared-CU look-ahead gives the same
  factors to rounding.
- With BFA_GEMM=own (a child process) both LU precisions run every trailing
  update on the hand-written GEMM tile, with the same backward-error bounds.
"""
import numpy as np
import pytest

from conftest import kuhn

pytestmark = pytest.mark.gpu


def colmajor(t):
    return t.cpu().numpy().T


@pytest.fixture(scope="module")
def mesh5(bfa, orc, torch_cuda):
    V, T = kuhn(bfa, 5)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    o = orc.Geometry(V, T)
    tab = o.static_table(10)
    return g, o, tab


def directions(nd, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nd):
        v = rng.standard_normal(3)
        n = v / np.linalg.norm(v)
        a = np.array([0.0, 0.0, 1.0]) if abs(n[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
        e1 = np.cross(n, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        out += [(tuple(e1), tuple(n)), (tuple(e2), tuple(n))]
    return out


def test_config1_fp64(bfa, orc, torch_cuda, mesh5):
    torch = torch_cuda
    g, o, tab = mesh5
    N = g.N
    print(f"[config1] N={N}", flush=True)
    assert N == 1350
    spec = [((1, 0, 0), (0, 0, 1))]
    pw, opw = bfa.planewaves(spec), orc.planewaves(spec)
    mat, omat = bfa.material("iso", eps=4 + 0.1j), orc.material("iso", eps=4 + 0.1j)
    dM = torch.zeros((N, N), dtype=torch.complex128, device="cuda")
    dB = torch.zeros((1, N), dtype=torch.complex128, device="cuda")
    dp = torch.zeros(N, dtype=torch.int32, device="cuda")
    g.assemble_vie(1.0, mat, dM)
    g.assemble_rhs(1.0, pw, dB)
    Mo = o.assemble(1.0, omat, tab)
    err = np.linalg.norm(colmajor(dM) - Mo) / np.linalg.norm(Mo)
    print("matrix rel err", err)
    assert err < 1e-10
    assert bfa.lu_factor(dM, N, dp) == 0
    bfa.lu_solve(dM, N, dp, dB, 1)
    Jo = orc.lu_solve(Mo, o.rhs(1.0, opw))
    Po = o.emt_power(1.0, opw, Jo)
    Pg = g.emt_power(1.0, pw, dB)
    print("power gpu", Pg, "oracle", Po)
    assert np.all(np.abs(Pg - Po) / np.abs(Po) < 1e-9)
    # absorbed and scattered power are positive for a lossy dielectric
    assert (Po > 0).all()


def test_config3_precision_form(bfa, orc, torch_cuda, mesh5):
    torch = torch_cuda
    g, o, tab = mesh5
    N = g.N
    spec = directions(16, 0xC0F3)
    assert len(spec) == 32
    pw, opw = bfa.planewaves(spec), orc.planewaves(spec)
    kw = dict(base=4.0, amp=2.0, imag=0.1)
    mat, omat = bfa.material("sincos", **kw), orc.material("sincos", **kw)
    Mo = o.assemble(1.0, omat, tab)
    Jo = orc.lu_solve(Mo, o.rhs(1.0, opw))
    Po = o.emt_power(1.0, opw, Jo)
    for far32 in (False, True):
        print(f"[config3] far32={far32}", flush=True)
        dM = torch.zeros((N, N), dtype=torch.complex128, device="cuda")
        dM32 = torch.zeros((N, N), dtype=torch.complex64, device="cuda")
        dB = torch.zeros((32, N), dtype=torch.complex128, device="cuda")
        dp = torch.zeros(N, dtype=torch.int32, device="cuda")
        g.assemble_vie(1.0, mat, dM, far32=far32)
        g.assemble_rhs(1.0, pw, dB)
        merr = np.linalg.norm(colmajor(dM) - Mo) / np.linalg.norm(Mo)
        bfa.convert_z2c(dM, N, N, dM32)
        assert bfa.lu_factor_c(dM32, N, dp) == 0
        it, res = bfa.solve_refine(dM, dM32, dp, N, dB, 32, tol=1e-10, maxit=12)
        assert res <= 1e-10
        Pg = g.emt_power(1.0, pw, dB)
        perr = np.abs(Pg - Po) / np.abs(Po)
        print(f"  matrix rel err {merr:.2e}; refinement {it} iters, residual {res:.2e}; "
              f"power rel err max {perr.max():.2e} median {np.median(perr):.2e}", flush=True)
        if far32:
            # the far field's second-order Taylor part in FP64, the rest of
            # the kernel in FP32: the power bar of the precision form
            assert merr < 1e-8 and perr.max() < 1e-8
        else:
            assert merr < 1e-10 and perr.max() < 1e-8


def test_lu_singular_reports_first_zero_pivot(bfa, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(3)
    for n, zc in ((40, 3), (600, 333)):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        A[:, zc] = 0.0
        for f64 in (True, False):
            dt = torch.complex128 if f64 else torch.complex64
            dA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dt).cuda()
            dp = torch.zeros(n, dtype=torch.int32, device="cuda")
            info = bfa.lu_factor(dA, n, dp) if f64 else bfa.lu_factor_c(dA, n, dp)
            assert info == zc + 1, (n, f64, info)


@pytest.mark.parametrize("f64", [False, True])
def test_lu_tall_matrix_backward_error(bfa, torch_cuda, f64):
    """n = 14,000 through the CU-partitioned look-ahead: panel workgroups of
    ~219 rows on each of the 64 panel CUs, and in the early steps the panel
    stream also updates a slice of the trailing columns; the solve's normwise
    backward error ||A x - b|| / (||A|| ||x||) is computed in complex128 on
    the device."""
    torch = torch_cuda
    n, nrhs = 14000, 4
    gen = torch.Generator(device="cuda:0").manual_seed(14000)
    A = torch.randn((n, n), dtype=torch.complex128, device="cuda:0", generator=gen)   # A[c, r]: column-major
    A += n ** 0.5 * torch.eye(n, dtype=torch.complex128, device="cuda:0")
    B = torch.randn((nrhs, n), dtype=torch.complex128, device="cuda:0", generator=gen)
    dt = torch.complex128 if f64 else torch.complex64
    LU = A.to(dt).clone()
    X = B.to(dt).clone()
    piv = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    if f64:
        assert bfa.lu_factor(LU, n, piv) == 0
        bfa.lu_solve(LU, n, piv, X, nrhs)
    else:
        assert bfa.lu_factor_c(LU, n, piv) == 0
        bfa.lu_solve_c(LU, n, piv, X, nrhs)
    Am = A.T.contiguous()                             # the matrix, row-major
    Xm = X.to(torch.complex128).T.contiguous()
    R = Am @ Xm - B.T
    eta = (torch.linalg.norm(R) / (torch.linalg.norm(A) * torch.linalg.norm(Xm))).item()
    print(f"[lu] n={n} {'z' if f64 else 'c'} backward error {eta:.2e}", flush=True)
    assert eta < (1e-15 if f64 else 1e-6)


def test_lu_shared_cu_path(bfa, torch_cuda):
    """bfa_lu_partition(0): the look-ahead LU with every stream on every CU
    (the path for matrices whose panel does not fit the panel CUs) gives the
    partitioned path's factors to rounding, and the partition comes back on."""
    torch = torch_cuda
    n = 3000
    gen = torch.Generator(device="cuda:0").manual_seed(3000)
    A = torch.randn((n, n), dtype=torch.complex128, device="cuda:0", generator=gen)
    A += n ** 0.5 * torch.eye(n, dtype=torch.complex128, device="cuda:0")
    out = {}
    try:
        for on in (False, True):
            bfa.lu_partition(on)
            for f64 in (True, False):
                LU = A.to(torch.complex128 if f64 else torch.complex64).clone()
                piv = torch.zeros(n, dtype=torch.int32, device="cuda:0")
                assert (bfa.lu_factor if f64 else bfa.lu_factor_c)(LU, n, piv) == 0
                out[(on, f64)] = (LU.to(torch.complex128), piv.clone())
    finally:
        bfa.lu_partition(True)
    for f64, tol in ((True, 1e-12), (False, 1e-4)):
        (a, pa), (b, pb) = out[(False, f64)], out[(True, f64)]
        assert torch.equal(pa, pb)
        err = (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()
        print(f"[lu] shared-CU vs partitioned {'z' if f64 else 'c'}: {err:.1e}", flush=True)
        assert err < tol


_OWN_GEMM_CHILD = r"""
import os, sys
import torch
sys.path.insert(0, os.path.join(sys.argv[1], "tests"))
from conftest import load_bfa
bfa = load_bfa()
n, nrhs = 3000, 4
gen = torch.Generator(device="cuda:0").manual_seed(3001)
A = torch.randn((n, n), dtype=torch.complex128, device="cuda:0", generator=gen)
A += n ** 0.5 * torch.eye(n, dtype=torch.complex128, device="cuda:0")
B = torch.randn((nrhs, n), dtype=torch.complex128, device="cuda:0", generator=gen)
for f64 in (True, False):
    dt = torch.complex128 if f64 else torch.complex64
    LU, X = A.to(dt).clone(), B.to(dt).clone()
    piv = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    if f64:
        assert bfa.lu_factor(LU, n, piv) == 0
        bfa.lu_solve(LU, n, piv, X, nrhs)
    else:
        assert bfa.lu_factor_c(LU, n, piv) == 0
        bfa.lu_solve_c(LU, n, piv, X, nrhs)
    Xm = X.to(torch.complex128).T.contiguous()
    R = A.T.contiguous() @ Xm - B.T
    eta = (torch.linalg.norm(R) / (torch.linalg.norm(A) * torch.linalg.norm(Xm))).item()
    print(f"eta {'z' if f64 else 'c'} {eta:.3e}", flush=True)
"""


def test_lu_own_gemm_backward_error(torch_cuda):
    """BFA_GEMM=own (read once per process, so in a child process): every
    trailing update of both LU precisions runs on the hand-written
    k_gemm_sub2 tile instead of rocBLAS; n = 3,000 gives updates of up to
    2,744 x 2,744 x 256, and the solve's normwise backward error must stay at
    the precision's level."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, BFA_GEMM="own")
    r = subprocess.run([sys.executable, "-c", _OWN_GEMM_CHILD, root], capture_output=True, text=True, env=env,
                       timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    eta = dict((w[1], float(w[2])) for w in (l.split() for l in r.stdout.splitlines()) if w and w[0] == "eta")
    assert eta["z"] < 1e-15 and eta["c"] < 1e-6
