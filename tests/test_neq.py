"""The buff-neq consumers of the hot path (GetFlux.cc:150-399, VIEMatrix.cc:321-360).

GPU:
- AssembleGBlock plus the V^-1 overlap entries is the VIE matrix.
- AssembleOverlapBlocks: V^-1 rows equal the oracle's GetOverlaps restatement
  (indices and order); for a constant isotropic eps, V = Omega^4 (eps-1)^2 V^-1
  entry by entry and, at a uniform temperature, Rytov = -4 Omega^3/(pi Z0)
  Im(eps-1) (eps-1) V^-1 (RelDeltaTheta = 1, also without a temperature
  field, GetOverlapIntegrand's default); a temperature gradient matches
  a numpy restatement of the Rytov integrand.
- S^-1 equals the inverse of the oracle's Gram matrix; the dressed Rytov
  matrix W (S^-1 R S^-1)^H equals the oracle's step-by-step restatement
  (constant eps with every input from the oracle; sincos eps with a
  temperature gradient); its EMT power equals the oracle's, and for
  DR = J J^H the J-vector EMT power.
- The C++ drop-in's buff-neq calls (buff_scatter_amd --rytov) reproduce the
  ctypes path's EMT power of the dressed Rytov matrix.
- LUSolve(HMatrix*) with many right-hand sides (the blocked TRSM + GEMM
  form) against numpy (complex128) and LAPACK cgetrf/cgetrs (complex64).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import kuhn

pytestmark = pytest.mark.gpu
Z0 = 376.73031346177


@pytest.fixture(scope="module")
def geo(bfa, orc, torch_cuda):
    V, T = kuhn(bfa, 3)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    return V, T, g, orc.Geometry(V, T)


def test_g_block_plus_overlaps_is_the_vie_matrix(bfa, torch_cuda, geo):
    torch = torch_cuda
    V, T, g, o = geo
    N = g.N
    w = 1.2
    mat = bfa.material("sincos", base=4.0, amp=2.0, imag=0.1)
    M = torch.empty((N, N), dtype=torch.complex128, device="cuda:0")
    G = torch.empty_like(M)
    g.assemble_vie(w, mat, M)
    g.assemble_g(w, G)
    Mh, Gh = M.cpu().numpy().T, G.cpu().numpy().T
    cnt, cols, Vb, VI, R = g.overlap_blocks(w, mat)
    for a in range(N):
        for k in range(cnt[a]):
            Gh[a, cols[a, k]] += VI[a, k]
    err = np.abs(Gh - Mh).max() / np.abs(Mh).max()
    print(f"[neq] G + VInv vs VIE matrix {err:.2e}", flush=True)
    assert err < 1e-13


def test_overlap_blocks(bfa, orc, torch_cuda, geo):
    V, T, g, o = geo
    w = 0.9
    eps = 4.0 + 0.3j
    mat = bfa.material("iso", eps=eps)
    cnt, cols, Vb, VI, R = g.overlap_blocks(w, mat)
    assert cnt.max() <= 7 and np.all(cols[np.arange(g.N), 0] == np.arange(g.N))
    omat = orc.material("iso", eps=eps)
    for a in range(0, g.N, 7):
        idx, val = o.overlaps(a, w, omat)
        assert cnt[a] == len(idx) and np.array_equal(cols[a, :cnt[a]], idx)
        np.testing.assert_allclose(VI[a, :cnt[a]], val, rtol=1e-13, atol=1e-13 * np.abs(val).max())
    m = np.arange(7)[None, :] < cnt[:, None]
    np.testing.assert_allclose(Vb[m], w ** 4 * (eps - 1) ** 2 * VI[m], rtol=1e-12)
    # no temperature field: RelDeltaTheta = 1 (GetOverlapIntegrand's default);
    # a uniform one: RelDeltaTheta = (Theta(T) - Theta_env) / (Theta(T_avg) - Theta_env) = 1
    want = (-4 * w ** 3 / (np.pi * Z0) * ((eps - 1).imag * (eps - 1) * VI)).real
    np.testing.assert_allclose(R[m], want[m], rtol=1e-11, atol=1e-14 * np.abs(want).max())
    Tm = bfa.material_expr("300")
    cnt2, cols2, Vb2, VI2, R2 = g.overlap_blocks(w, mat, temperature=Tm, t_env=0.0)
    np.testing.assert_allclose(R2[m], want[m], rtol=1e-11, atol=1e-14 * np.abs(want).max())


def _rytov_row_numpy(orc, o, V, T, a, w, eps, Tfun, t_env, t_avg, nint):
    """GetOverlaps' Rytov entries of row a restated in numpy (VIEMatrix.cc:122-316)."""
    F = o.faces()
    k = 4.36763e-4
    th = lambda Tx: 0.0 if Tx == 0 else w / (np.exp(w / (k * Tx)) - 1)
    th_env = th(t_env)
    dhat = th(t_avg) - th_env
    r33 = orc.tetcr(33)
    out = {a: 0.0}
    order = [a]
    for side, (tet, qv, sg) in enumerate(((F["PT"][a], F["QP"][a], 1.0), (F["MT"][a], F["QM"][a], -1.0))):
        P0 = V[T[tet][0]]
        L = V[T[tet][1:]] - P0
        vol = np.linalg.det(L) / 6.0
        pfa = sg * F["Area"][a] / (3 * vol)
        for b in F["tetFI"][tet]:
            if b >= nint:                              # boundary faces carry no BF
                continue
            qb, sb = (F["QP"][b], 1.0) if F["PT"][b] == tet else (F["QM"][b], -1.0)
            pfb = sb * F["Area"][b] / (3 * vol)
            acc = 0.0
            for u1, u2, u3, wt in r33:
                x = P0 + u1 * L[0] + u2 * L[1] + u3 * L[2]
                rel = (th(Tfun(x)) - th_env) / dhat
                acc += 6 * vol * wt * rel * (pfa * (x - V[qv])) @ (pfb * (x - V[qb])) * (eps - 1).imag
            val = acc * 4 * w / (np.pi * Z0)
            if b == a:
                out[a] += val
            else:
                out[b] = val
                order.append(b)
    return order, np.array([out[b] for b in order])


def test_rytov_with_temperature_gradient(bfa, orc, torch_cuda, geo):
    V, T, g, o = geo
    w, eps, t_env = 0.7, 3.0 + 0.8j, 280.0
    Tfun = lambda x: 300.0 + 40.0 * x[0] - 15.0 * x[2]
    Tm = bfa.material_expr("300 + 40*x - 15*z")
    cnt, cols, _, _, R = g.overlap_blocks(w, bfa.material("iso", eps=eps), temperature=Tm, t_env=t_env)
    # T_avg as GetAverageTemperature: volume-weighted at tetrahedron centroids
    cen = V[T].mean(1)
    vols = np.array([np.linalg.det(V[t[1:]] - V[t[0]]) / 6 for t in T])
    t_avg = (vols * np.array([Tfun(c) for c in cen])).sum() / vols.sum()
    nint = g.N
    for a in range(0, nint, 11):
        order, want = _rytov_row_numpy(orc, o, V, T, a, w, eps, Tfun, t_env, t_avg, nint)
        want_sorted = dict(zip(order, want))
        got = dict(zip(cols[a, :cnt[a]], R[a, :cnt[a]]))
        assert set(got) == set(want_sorted)
        for b in got:
            assert abs(got[b] - want_sorted[b]) <= 1e-11 * np.abs(want).max(), (a, b, got[b], want_sorted[b])


@pytest.mark.parametrize("n,nrhs", [(700, 100), (1100, 300)])
def test_many_rhs_solve(bfa, torch_cuda, n, nrhs):
    torch = torch_cuda
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    X = np.linalg.solve(A, B)
    for f64 in (True, False):
        dt = torch.complex128 if f64 else torch.complex64
        dA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dt).cuda()
        dB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dt).cuda()
        piv = torch.zeros(n, dtype=torch.int32, device="cuda")
        if f64:
            assert bfa.lu_factor(dA, n, piv) == 0
            bfa.lu_solve(dA, n, piv, dB, nrhs)
        else:
            assert bfa.lu_factor_c(dA, n, piv) == 0
            bfa.lu_solve_c(dA, n, piv, dB, nrhs)
        Y = dB.cpu().numpy().T.astype(np.complex128)
        err = np.linalg.norm(Y - X) / np.linalg.norm(X)
        # normwise backward error ||A Y - B|| / (||A|| ||Y||): the solver's own
        # quality (the forward error above also carries numpy's f64 rounding)
        eta = np.linalg.norm(A @ Y - B) / (np.linalg.norm(A) * np.linalg.norm(Y))
        if f64:
            print(f"[neq] many-RHS solve n={n} nrhs={nrhs} z: forward {err:.2e} backward {eta:.2e}", flush=True)
            assert eta < 1e-15 and err < 5e-12
            continue
        # complex64: against LAPACK's single-precision cgetrf + cgetrs of the
        # same system (numpy's own complex64 solve lands 30x closer on this
        # matrix by another route; cgetrf's factors solved exactly give
        # cgetrs' error)
        lu32 = sl.lu_factor(A.astype(np.complex64))
        Y32 = sl.lu_solve(lu32, B.astype(np.complex64)).astype(np.complex128)
        ref32 = np.linalg.norm(Y32 - X) / np.linalg.norm(X)
        eta32 = np.linalg.norm(A @ Y32 - B) / (np.linalg.norm(A) * np.linalg.norm(Y32))
        print(f"[neq] many-RHS solve n={n} nrhs={nrhs} c: forward {err:.2e} backward {eta:.2e}"
              f" (LAPACK complex64 {ref32:.2e}, {eta32:.2e})", flush=True)
        assert err < 3 * ref32 and eta < 3 * eta32


# ---- dressed Rytov matrix (ComputeDressedRytovMatrix) and its EMT power ----

def _neq_bufs(torch, g):
    N = g.N
    Sinv = torch.empty((N, N), dtype=torch.complex128, device="cuda:0")
    g.gram_inverse(Sinv)
    DR = torch.empty_like(Sinv)
    work = torch.empty(g.dressed_rytov_work_bytes(), dtype=torch.uint8, device="cuda:0")
    return Sinv, DR, work


def _oracle_G(orc, o, w, omat):
    """The G block from the oracle: its VIE matrix minus GetOverlaps' V^-1 entries."""
    G = o.assemble(w, omat, o.static_table(10))
    for a in range(o.N):
        idx, val = o.overlaps(a, w, omat)
        G[a, idx] -= val
    return G


def _dense_rows(N, cnt, cols, vals):
    M = np.zeros((N, N), dtype=vals.dtype)
    for a in range(N):
        M[a, cols[a, :cnt[a]]] = vals[a, :cnt[a]]
    return M


def test_gram_inverse(bfa, orc, torch_cuda, geo):
    V, T, g, o = geo
    Sinv, _, _ = _neq_bufs(torch_cuda, g)
    want = np.linalg.inv(o.gram())
    err = np.abs(Sinv.cpu().numpy().T - want).max() / np.abs(want).max()
    print(f"[neq] S^-1 vs oracle {err:.2e}", flush=True)
    assert err < 1e-12


def test_dressed_rytov_constant_eps(bfa, orc, torch_cuda, geo):
    """Every input from the oracle: G, S, and V, R by their closed forms in
    V^-1 for a constant isotropic eps (no temperature field: RelDeltaTheta = 1)."""
    V, T, g, o = geo
    N = g.N
    w, eps = 0.8, 4.0 + 0.6j
    Sinv, DR, work = _neq_bufs(torch_cuda, g)
    g.dressed_rytov(w, bfa.material("iso", eps=eps), Sinv, DR, work)
    omat = orc.material("iso", eps=eps)
    G = _oracle_G(orc, o, w, omat)
    VI = np.zeros((N, N), dtype=complex)
    for a in range(N):
        idx, val = o.overlaps(a, w, omat)
        VI[a, idx] = val
    Vb = w ** 4 * (eps - 1) ** 2 * VI
    R = (-4 * w ** 3 / (np.pi * Z0) * (eps - 1).imag * (eps - 1) * VI).real
    want = orc.dressed_rytov(G, np.linalg.inv(o.gram()), Vb, R)
    got = DR.cpu().numpy().T
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[neq] dressed Rytov (constant eps) vs oracle {err:.2e}", flush=True)
    assert err < 1e-10
    # its EMT power against the oracle's, and P_abs = -P_scat
    p = g.emt_power_dr(w, DR)
    po = o.emt_power_dr(w, want)
    print(f"[neq] EMT(DR) {p[1]:.6e} oracle {po[1]:.6e}", flush=True)
    assert abs(p[1] - po[1]) <= 1e-9 * abs(po[1]) and p[0] == -p[1]


def test_dressed_rytov_temperature_gradient(bfa, orc, torch_cuda, geo):
    """sincos eps and a temperature gradient: V, R from AssembleOverlapBlocks
    (pinned in test_overlap_blocks / test_rytov_with_temperature_gradient), G
    and S from the oracle."""
    V, T, g, o = geo
    N = g.N
    w, t_env = 1.1, 280.0
    mat = bfa.material("sincos", base=4.0, amp=2.0, imag=0.1)
    Tm = bfa.material_expr("300 + 40*x - 15*z")
    Sinv, DR, work = _neq_bufs(torch_cuda, g)
    g.dressed_rytov(w, mat, Sinv, DR, work, temperature=Tm, t_env=t_env)
    cnt, cols, Vr, _, Rr = g.overlap_blocks(w, mat, temperature=Tm, t_env=t_env)
    G = _oracle_G(orc, o, w, orc.material("sincos", base=4.0, amp=2.0, imag=0.1))
    want = orc.dressed_rytov(G, np.linalg.inv(o.gram()), _dense_rows(N, cnt, cols, Vr), _dense_rows(N, cnt, cols, Rr))
    err = np.abs(DR.cpu().numpy().T - want).max() / np.abs(want).max()
    print(f"[neq] dressed Rytov (sincos, T gradient) vs oracle {err:.2e}", flush=True)
    assert err < 1e-10


def test_emt_power_dr_is_the_j_path_for_a_rank_one_dr(bfa, torch_cuda, geo):
    """DR = J J^H gives the J-vector EMT's scattered power (GetJJ's two forms)."""
    torch = torch_cuda
    V, T, g, o = geo
    N = g.N
    rng = np.random.default_rng(21)
    J = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    w = 0.9
    pw = bfa.planewaves([((1, 0, 0), (0, 0, 1))])
    dJ = torch.from_numpy(J[None, :].copy()).to("cuda:0")
    pj = g.emt_power(w, pw, dJ)
    DRm = np.outer(J, J.conj())
    U = np.triu(rng.standard_normal((N, N)), 1)          # never read
    dDR = torch.from_numpy(np.ascontiguousarray((DRm + U).T)).to("cuda:0")
    p = g.emt_power_dr(w, dDR)
    print(f"[neq] EMT(J J^H) {p[1]:.6e} J path {pj[0][1]:.6e}", flush=True)
    assert abs(p[1] - pj[0][1]) <= 1e-12 * abs(pj[0][1])


def test_cpp_driver_rytov_flux(bfa, torch_cuda, geo):
    """buff_scatter_amd --rytov (the C++ drop-in: ComputeDressedRytovMatrix,
    GetPFTMatrix(DR)) gives the ctypes path's EMT power of the dressed Rytov
    matrix on the same mesh and material."""
    import os
    import subprocess
    from conftest import ROOT
    V, T, g, o = geo
    w, eps, texpr, t_env = 0.8, 4.0 + 0.6j, "300 + 40*x - 15*z", 280.0
    drv = os.path.join(ROOT, "buff-em_amd", "lib", "buff_scatter_amd")
    print("[neq] driver --rytov n=3", flush=True)
    r = subprocess.run([drv, "--kuhn", "3", "1.0", "--eps", str(eps.real), str(eps.imag), "--omega", str(w),
                        "--rytov", texpr, str(t_env)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = [l.split() for l in r.stdout.strip().splitlines() if l.split()[1] == "rytov"]
    assert len(line) == 1
    got = np.array([float(line[0][2]), float(line[0][3])])
    # the drop-in builds its static table at the default Taylor-Duffy order (12)
    g12 = bfa.Geometry(V, T)
    g12.upload()
    g12.build_static(12)
    Sinv, DR, work = _neq_bufs(torch_cuda, g12)
    g12.dressed_rytov(w, bfa.material("iso", eps=eps), Sinv, DR, work, temperature=bfa.material_expr(texpr),
                      t_env=t_env)
    want = g12.emt_power_dr(w, DR)
    print(f"[neq] driver {got} ctypes {want}", flush=True)
    assert np.all(np.abs(got - want) <= 1e-9 * np.abs(want))
