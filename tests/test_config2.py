"""The headline workload against the CPU oracle at full size.

Kuhn sphere n = 12 (10,368 tetrahedra, N = 19,872), R = 1 um, 5 % jitter,
eps = diag(4,5,6) + 0.05i, omega = 1, x- and y-polarised plane waves along
+z (the reference driver's loop, buff-scatter.cc:236-314, at one frequency).

- Static (FIBBI) records: a seeded sample of 3,000 touching pairs against
  the oracle's ComputeFIBBIData restatement (GMatrixElements.cc:370-375).
- VIE matrix: the whole N x N device matrix against the oracle's
  AssembleVIEMatrix (VIEMatrix.cc:431-460), given the same static records,
  to 1e-10 relative Frobenius error; a sample of rows also elementwise.
- Solve and power: the device LU/solve (partial pivoting) and EMT power
  against LAPACK zgetrf/zgetrs on the oracle's matrix and the oracle's
  GetEMTPFT (EMTPFT.cc:324-546), to 1e-9 relative.

About a minute on 16 host cores (oracle assembly, LAPACK, oracle EMT).
"""
import time

import numpy as np
import pytest

from conftest import kuhn

pytestmark = pytest.mark.gpu

EPS = (4 + 0.05j, 5 + 0.05j, 6 + 0.05j)
SPEC = [((1, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 0, 1))]


@pytest.mark.timeout(900)
def test_config2_fp64_against_oracle(bfa, orc, torch_cuda):
    torch = torch_cuda
    t_start = time.perf_counter()
    V, T = kuhn(bfa, 12)
    g = bfa.Geometry(V, T)
    N = g.N
    print(f"[config2] N={N}", flush=True)
    assert N == 19872
    g.upload()
    g.build_static(10)
    o = orc.Geometry(V, T)
    tab = g.static_table()

    # static records: a seeded sample against the oracle
    rowptr, cols = o.touch_pairs()
    rng = np.random.default_rng(0xC0F2)
    idx = rng.choice(len(cols), 3000, replace=False)
    rows = np.searchsorted(rowptr, idx, side="right") - 1
    ref = np.array([o.fibbi(int(a), int(cols[i]), 10) for a, i in zip(rows, idx)])
    serr = np.abs(tab[idx] - ref).max() / np.abs(ref).max()
    print(f"[config2] static records (3000 sampled) rel err {serr:.2e}", flush=True)
    assert serr < 1e-12

    mat, omat = bfa.material("diag", eps=EPS), orc.material("diag", eps=EPS)
    pw, opw = bfa.planewaves(SPEC), orc.planewaves(SPEC)
    dM = torch.empty((N, N), dtype=torch.complex128, device="cuda")   # dM[c, r] = M(r, c)
    dB = torch.empty((2, N), dtype=torch.complex128, device="cuda")
    dp = torch.empty(N, dtype=torch.int32, device="cuda")
    g.assemble_vie(1.0, mat, dM)
    g.assemble_rhs(1.0, pw, dB)
    Mg = dM.cpu().numpy()                                            # Mg[c, r]

    Mo = o.assemble_colmajor(1.0, omat, tab)                          # Mo[r, c], Fortran order
    print(f"[config2] oracle assembly done at {time.perf_counter() - t_start:.0f}s", flush=True)
    num = den = 0.0
    for c0 in range(0, N, 2048):
        c1 = min(N, c0 + 2048)
        d = Mg[c0:c1] - Mo[:, c0:c1].T
        num += float(np.vdot(d, d).real)
        den += float(np.vdot(Mo[:, c0:c1], Mo[:, c0:c1]).real)
    err = (num / den) ** 0.5
    print(f"[config2] matrix rel Frobenius err {err:.2e}", flush=True)
    assert np.isfinite(err) and err < 1e-10
    # a few rows elementwise (relative to the row's largest entry)
    for r in (0, 7, N // 3, N - 1):
        rel = np.abs(Mg[:, r] - Mo[r, :]).max() / np.abs(Mo[r, :]).max()
        assert rel < 1e-10, (r, rel)
    del Mg

    B = o.rhs(1.0, opw)
    Bg = dB.cpu().numpy().T
    berr = np.linalg.norm(Bg - B) / np.linalg.norm(B)
    print(f"[config2] RHS rel err {berr:.2e}", flush=True)
    assert berr < 1e-12

    assert bfa.lu_factor(dM, N, dp) == 0
    bfa.lu_solve(dM, N, dp, dB, 2)
    Pg = g.emt_power(1.0, pw, dB)
    Jg = dB.cpu().numpy().T
    del dM
    torch.cuda.empty_cache()

    Jo = orc.lu_solve(Mo, B, overwrite=True)
    del Mo
    jerr = np.linalg.norm(Jg - Jo) / np.linalg.norm(Jo)
    print(f"[config2] solution rel err {jerr:.2e} at {time.perf_counter() - t_start:.0f}s", flush=True)
    Po = o.emt_power(1.0, opw, Jo)
    perr = np.abs(Pg - Po) / np.abs(Po)
    print(f"[config2] P_abs/P_scat gpu {Pg.tolist()} oracle {Po.tolist()} rel err max {perr.max():.2e} "
          f"({time.perf_counter() - t_start:.0f}s)", flush=True)
    assert jerr < 1e-9
    assert perr.max() < 1e-9
    assert (Po > 0).all()
