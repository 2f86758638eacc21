"""Learned near-field surrogate.

CPU: the oracle's tetrahedron-pair block is pinned against the restated
reference far-field matrix element (GetGME_BFBFInt with the 4-point rule is
the signed sum of four block entries), against its own symmetry, and its
subdivided cubature converges.

GPU: cubature blocks, canonical features and the table-driven assembly match
the oracle / the standard assembly; then an ensemble is trained on the
n=5 sphere's pairs (labels from subdivided cubature over omega in [0.1, 10])
and, at a frequency it never saw, the test measures what the network gains
over the reference's 4-point rule against a finer cubature: block error,
cubature fallbacks, and the solved powers.
"""
import numpy as np
import pytest

from conftest import kuhn


def _far_pairs(o, N, count, rng):
    out = []
    while len(out) < count:
        a, b = (int(x) for x in rng.integers(0, N, 2))
        if a <= b and o.compare_bfs(a, b) == 0:
            out.append((a, b))
    return out


def test_block_rule_reproduces_far_gme(bfa, orc):
    """sum_{s,t in {P,M}} sigma_s sigma_t G[i_s][j_t](T_s, T_t) with the 4-point
    rule equals the oracle's far-field element for BF pairs sharing no vertex."""
    V, T = kuhn(bfa, 2)
    o = orc.Geometry(V, T)
    F = o.faces()
    rng = np.random.default_rng(5)
    for a, b in _far_pairs(o, o.N, 12, rng):
        for k in (0.3, 1.0, 4.0 + 0.5j):
            ref = o.gme(a, b, k)
            acc = 0.0
            for sa, (ta, ia) in enumerate([(F["PT"][a], F["PI"][a]), (F["MT"][a], F["MI"][a])]):
                for sb, (tb, jb) in enumerate([(F["PT"][b], F["PI"][b]), (F["MT"][b], F["MI"][b])]):
                    G = orc.tp_block(V[T[ta]], V[T[tb]], k, 4, 0)
                    acc += (1 if sa == sb else -1) * G[ia, jb]
            assert abs(acc - ref) <= 1e-12 * abs(ref), (a, b, k, acc, ref)


def test_block_symmetry_and_convergence(bfa, orc):
    V, T = kuhn(bfa, 3)
    cen = V[T].mean(1)
    rng = np.random.default_rng(9)
    checked = 0
    while checked < 4:
        s, t = (int(x) for x in rng.integers(0, len(T), 2))
        if s == t or set(T[s]) & set(T[t]):
            continue
        for k in (0.5, 5.0):
            G = orc.tp_block(V[T[s]], V[T[t]], k, 33, 1)
            Gt = orc.tp_block(V[T[t]], V[T[s]], k, 33, 1)
            assert np.abs(G - Gt.T).max() <= 1e-12 * np.abs(G).max()
            fine = orc.tp_block(V[T[s]], V[T[t]], k, 33, 2)
            coarse = orc.tp_block(V[T[s]], V[T[t]], k, 16, 1)
            e1 = np.linalg.norm(G - fine) / np.linalg.norm(fine)
            e0 = np.linalg.norm(coarse - fine) / np.linalg.norm(fine)
            d = np.linalg.norm(cen[s] - cen[t])
            assert e1 < 1e-6 and e1 < e0, (s, t, k, d, e0, e1)
        checked += 1


def test_features_oracle_invariance(orc):
    """Canonical inputs are invariant under rigid motion and scaling (with k
    scaled inversely), and a dropped frame axis changes them."""
    rng = np.random.default_rng(2)
    VS = rng.standard_normal((4, 3))
    VT = rng.standard_normal((4, 3)) + 3.0
    f0 = orc.tp_features(VS, VT, 1.3)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    sh, sc = rng.standard_normal(3), 2.5
    f1 = orc.tp_features(sc * VS @ Q.T + sh, sc * VT @ Q.T + sh, 1.3 / sc)
    np.testing.assert_allclose(f1, f0, atol=1e-12)
    # source shape block: v1 - v0 on e1, v2 - v0 in the e1 e2 plane
    h = np.mean([np.linalg.norm(VS[j] - VS[i]) for i in range(4) for j in range(i + 1, 4)])
    assert abs(f0[12] - np.linalg.norm(VS[1] - VS[0]) / h) < 1e-12 and f0[17] > 0
    assert abs(np.hypot(f0[13], f0[14]) - np.linalg.norm(VS[2] - VS[0]) / h) < 1e-12
    assert abs(np.linalg.norm(f0[15:18]) - np.linalg.norm(VS[3] - VS[0]) / h) < 1e-12
    assert abs(np.linalg.norm(f0[:12].reshape(4, 3), axis=1) -
               np.linalg.norm(VT - VS.mean(0), axis=1) /
               np.mean([np.linalg.norm(VS[j] - VS[i]) for i in range(4) for j in range(i + 1, 4)])).max() < 1e-12


@pytest.fixture(scope="module")
def n5(bfa, torch_cuda):
    V, T = kuhn(bfa, 5)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    S = bfa.Surrogate(g, factor=3.0, ensemble=4, hidden=256, seed=7)
    return V, T, g, S


@pytest.mark.gpu
def test_gpu_cubature_and_features_match_oracle(bfa, orc, n5):
    V, T, g, S = n5
    ps, pt = S.pairs()
    print(f"[surrogate] pairs={S.npairs}", flush=True)
    assert S.npairs > 1000
    rng = np.random.default_rng(4)
    idx = rng.choice(S.npairs, 12, replace=False)
    om = np.exp(rng.uniform(np.log(0.1), np.log(10), 12))
    for npts, level in ((4, 0), (16, 1), (33, 1)):
        G = S.cubature(idx, om, npts, level)
        for m, (p, w) in enumerate(zip(idx, om)):
            R = orc.tp_block(V[T[ps[p]]], V[T[pt[p]]], w, npts, level)
            assert np.abs(G[m] - R).max() <= 1e-12 * np.abs(R).max(), (npts, level, p, w)
    F = S.features(idx, om)
    for m, (p, w) in enumerate(zip(idx, om)):
        np.testing.assert_allclose(F[m], orc.tp_features(V[T[ps[p]]], V[T[pt[p]]], w), rtol=2e-6, atol=2e-6)


@pytest.mark.gpu
def test_gpu_table_assembly_equals_standard(bfa, n5, torch_cuda):
    """With the 4-point-rule table the table-driven assembly is the standard one."""
    torch = torch_cuda
    V, T, g, S = n5
    N = g.N
    mat = bfa.material("iso", eps=4 + 0.1j)
    A = torch.empty((N, N), dtype=torch.complex128, device="cuda:0")
    Bm = torch.empty_like(A)
    for w in (0.5, 3.0):
        g.assemble_vie(w, mat, A)
        S.table(w, mode=S.RULE)
        S.assemble(w, mat, Bm)
        err = (torch.linalg.norm(A - Bm) / torch.linalg.norm(A)).item()
        assert err < 1e-13, (w, err)


def _power(bfa, torch, g, mat, w, pw, assemble):
    N = g.N
    M = torch.empty((N, N), dtype=torch.complex128, device="cuda:0")
    B = torch.empty((len(pw), N), dtype=torch.complex128, device="cuda:0")
    piv = torch.empty(N, dtype=torch.int32, device="cuda:0")
    assemble(M)
    g.assemble_rhs(w, pw, B)
    assert bfa.lu_factor(M, N, piv) == 0
    bfa.lu_solve(M, N, piv, B, len(pw))
    return g.emt_power(w, pw, B)


@pytest.mark.gpu
def test_surrogate_gain(bfa, n5, torch_cuda):
    """Trained surrogate (16-point rule plus the network's correction) versus
    the 16-point rule and the reference's 4-point rule, measured against finer
    cubature (33-point rule on tetrahedra refined twice): block errors over
    every pair at omega = 1, and absorbed and scattered powers for 32 plane
    waves at omega = 0.5, 1, 2, 4 (inside the training range, none of them a
    training frequency)."""
    torch = torch_cuda
    V, T, g, S = n5
    import time
    from test_configs import directions
    t0 = time.time()
    S.dataset(per_pair=8, wmin=0.1, wmax=10.0, label_npts=33, label_level=1)   # ~2M samples
    t1 = time.time()
    hist = S.train(epochs=60, batch=4096, lr=2e-3)
    t2 = time.time()
    print(f"[surrogate] dataset {t1 - t0:.1f}s train {t2 - t1:.1f}s loss {hist[0].mean():.3e} -> {hist[-1].mean():.3e}",
          flush=True)
    assert hist[-1].mean() < hist[0].mean()
    n = S.npairs
    nrm = lambda X: np.linalg.norm(X.reshape(len(X), -1), axis=1)
    w = 1.0
    truth = S.cubature(np.arange(n), w, 33, 2)
    S.table(w, mode=S.RULE)
    rule = S.table_download()
    S.table(w, mode=S.CUBATURE, npts=16, level=0)
    r16 = S.table_download()
    nf, dmax = S.table(w, mode=S.SURROGATE, tol=np.inf)
    surr = S.table_download()
    e_rule, e_16, e_surr = (nrm(X - truth) / nrm(truth) for X in (rule, r16, surr))
    print(f"[surrogate] block error vs cubature(33,L2) at omega=1 over {n} pairs (median / p90): 4-point rule "
          f"{np.median(e_rule):.2e} / {np.quantile(e_rule, .9):.2e}; 16-point rule {np.median(e_16):.2e} / "
          f"{np.quantile(e_16, .9):.2e}; surrogate {np.median(e_surr):.2e} / {np.quantile(e_surr, .9):.2e}; "
          f"surrogate closer on {np.mean(e_surr < e_16):.2f} of the pairs", flush=True)
    # what the network adds over the rule it corrects, and over the reference's rule
    assert np.median(e_surr) < 0.5 * np.median(e_16)
    assert np.quantile(e_surr, 0.9) < 0.6 * np.quantile(e_16, 0.9)
    assert np.median(e_surr) < 0.01 * np.median(e_rule)
    # cubature fallbacks by ensemble disagreement (in units of ||R16 - R4||)
    counts = {}
    for tol in (1.0, 0.3, 0.1, 0.03):
        counts[tol] = S.table(w, mode=S.SURROGATE, tol=tol)[0]
        if tol == S.DEFAULT_TOL:
            fb = S.table_download()
            e_fb = nrm(fb - truth) / nrm(truth)
    print("[surrogate] cubature fallbacks by disagreement tolerance: " +
          ", ".join(f"{t}: {c}/{n}" for t, c in counts.items()) +
          f"; block error with fallback at {S.DEFAULT_TOL} (median / p90) {np.median(e_fb):.2e} / "
          f"{np.quantile(e_fb, .9):.2e}", flush=True)
    assert np.quantile(e_fb, 0.9) <= np.quantile(e_surr, 0.9)
    # wall time of each way to fill the table of all pairs (second call of
    # each, after a warm one), on the device
    def timed(**kw):
        S.table(w, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = S.table(w, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t, r
    tt = {"4-point rule": timed(mode=S.RULE)[0], "16-point rule": timed(mode=S.CUBATURE, npts=16, level=0)[0],
          "cubature 33 x refined once (labels)": timed(mode=S.CUBATURE, npts=33, level=1)[0],
          "surrogate, no fallback": timed(mode=S.SURROGATE, tol=np.inf)[0],
          f"surrogate, fallback at {S.DEFAULT_TOL}": timed(mode=S.SURROGATE, tol=S.DEFAULT_TOL)[0]}
    print("[surrogate] table time for " + str(n) + " pairs: " +
          ", ".join(f"{k} {v * 1e3:.1f} ms" for k, v in tt.items()), flush=True)
    # solved powers against those of the all-cubature far field
    mat = bfa.material("iso", eps=4 + 0.1j)
    pw = bfa.planewaves(directions(16, 0xC0F3))
    errs = {k: [] for k in ("4-point rule", "16-point rule", "surrogate", "surrogate + fallback")}
    for w in (0.5, 1.0, 2.0, 4.0):
        def power(tab):
            S.table_upload(tab)
            return _power(bfa, torch, g, mat, w, pw, lambda M: S.assemble(w, mat, M))
        P_t = power(S.cubature(np.arange(n), w, 33, 2))
        tabs = {}
        for k, kw in (("4-point rule", dict(mode=S.RULE)), ("16-point rule", dict(mode=S.CUBATURE, npts=16, level=0)),
                      ("surrogate", dict(mode=S.SURROGATE, tol=np.inf)),
                      ("surrogate + fallback", dict(mode=S.SURROGATE, tol=S.DEFAULT_TOL))):
            S.table(w, **kw)
            tabs[k] = S.table_download()
        for k, tab in tabs.items():
            errs[k].append((np.abs(power(tab) - P_t) / np.abs(P_t)).ravel())
    errs = {k: np.concatenate(v) for k, v in errs.items()}
    print("[surrogate] power vs all-cubature far field, 32 waves x 4 frequencies x (abs, scat), "
          "median / p90 / max: " + "; ".join(f"{k} {np.median(v):.2e} / {np.quantile(v, .9):.2e} / {v.max():.2e}"
                                              for k, v in errs.items()) +
          f"; surrogate closer than the 16-point rule on {np.mean(errs['surrogate'] < errs['16-point rule']):.2f}",
          flush=True)
    assert np.quantile(errs["surrogate"], 0.9) < 0.7 * np.quantile(errs["16-point rule"], 0.9)
    assert errs["surrogate"].max() < errs["16-point rule"].max()
    assert np.quantile(errs["surrogate"], 0.9) < 0.1 * np.quantile(errs["4-point rule"], 0.9)
