"""Pins of the oracle's finite-frequency parts by physics and exact integrals.

The reference's own tests hold no numeric golden vectors for this path, so
the oracle's Helmholtz far field (GMatrixElements.cc:105-145, 270-288), its
plane-wave right-hand side (RHSVector.cc:62-124) and its EMT powers
(EMTPFT.cc:85-546) are checked here against what physics and exact
integration fix, each check chosen so that a plausible slip (a sign of ik,
the -div div/k^2 term, the RHS prefactor -1/(i Omega Z0), a dropped factor
of the extinction or scattering integral) fails it:

- extinction identity: for M J = v, P_ext = (Omega Z0 / 2) J^H (Im M) J
  exactly (M complex symmetric, v = -<b, E_inc>/(i Omega Z0)); the EMT
  extinction term and the RHS are separate code paths from the matrix;
- optical theorem: for real eps no power is absorbed, so P_abs = P_ext -
  P_scat vanishes to discretisation error and shrinks under refinement;
- Mie series (Bohren & Huffman; the reference's own comparison,
  doc/docs/examples/MieScattering): the absorbed power of refined Kuhn
  spheres extrapolates to the series, the scattered power's error shrinks;
- the low-frequency dielectric sphere of unit-test-LFField.cc:50-69: the
  interior field 3 E0/(eps+2) (from the current, E = i Z0 J / (Omega
  (eps-1)) inside the body) on refined spheres, and on the reference's own
  E10Sphere_533 mesh at the EPFile.XAxis points;
- Galerkin identity: a far matrix element equals <b_a, E_b>/(i Omega Z0)
  with E_b the field of a unit current in BF b from the fields restatement,
  which Maxwell's equations pin (tests/test_fields.py);
- right-hand side: equals the exact integral of b_a . E0 e^{i Omega n.x}
  over each tetrahedron (a series in barycentric monomials, independent of
  the cubature rule).
"""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, kuhn
import mie

GOLDEN = os.path.join(ROOT, "tests", "golden")
PW_X = [((1, 0, 0), (0, 0, 1))]


def solve(orc, V, T, omega, eps, spec=PW_X, td_order=6):
    o = orc.Geometry(V, T)
    tab = o.static_table(td_order)
    pw = orc.planewaves(spec)
    M = o.assemble_colmajor(omega, orc.material("iso", eps=eps), tab)
    B = o.rhs(omega, pw)
    J = orc.lu_solve(M.copy(order="F"), B, overwrite=True)
    return o, M, B, J, o.emt_power(omega, pw, J)


@pytest.mark.parametrize("eps", [4 + 0.1j, 4.0 + 0j, 12 + 1j])
def test_extinction_identity(bfa, orc, eps):
    V, T = kuhn(bfa, 4)
    for omega in (0.7, 2.0):
        o, M, B, J, P = solve(orc, V, T, omega, eps, spec=[((1, 0, 0), (0, 0, 1)), ((0, 0.6, 0.8j), (1, 0, 0))])
        for j in range(2):
            x = J[:, j]
            p_ext = 0.5 * omega * orc.ZVAC * float(np.real(np.conj(x) @ (M.imag @ x)))
            got = P[j, 0] + P[j, 1]
            assert abs(got - p_ext) <= 1e-10 * abs(p_ext), (eps, omega, j, got, p_ext)


def test_optical_theorem_lossless(bfa, orc):
    ratios = {}
    for n in (4, 5):
        V, T = kuhn(bfa, n)
        for omega in (1.0, 2.0):
            P = solve(orc, V, T, omega, 4.0 + 0j)[4][0]
            ratios[(n, omega)] = P[0] / P[1]
            assert P[1] > 0
    print("P_abs / P_scat for real eps:", ratios)
    # measured: -1.4e-2, -5.5e-2 (n = 4); -7.1e-3, -3.7e-2 (n = 5)
    for omega, bound in ((1.0, 1.5e-2), (2.0, 7e-2)):
        assert abs(ratios[(5, omega)]) < abs(ratios[(4, omega)])
        assert abs(ratios[(4, omega)]) < bound


def test_mie_series(bfa, orc):
    """omega = 1, R = 1 um (x = kR = 1), eps = 4 + 0.1i, n = 4, 6, 8."""
    eps = 4 + 0.1j
    pa, ps = mie.powers(1.0, 1.0, eps)
    ns = (4, 6, 8)
    Pa, Ps = [], []
    for n in ns:
        V, T = bfa.kuhn_sphere(n, 1.0, 0.0, 0x5EED)
        P = solve(orc, V, T, 1.0, eps)[4][0]
        Pa.append(P[0])
        Ps.append(P[1])
    ea = [p / pa - 1 for p in Pa]
    es = [p / ps - 1 for p in Ps]
    print("Mie P_abs", pa, "errors", ea, "; P_scat", ps, "errors", es)
    # both approach the series from below, monotonically
    assert all(e < 0 for e in ea + es)
    assert abs(ea[0]) > abs(ea[1]) > abs(ea[2])
    assert abs(es[0]) > abs(es[1]) > abs(es[2])
    # the absorbed power follows a power law in h = 2R/n; its limit is the series
    lim, p = mie.richardson([2.0 / n for n in ns], Pa)
    print(f"P_abs extrapolated {lim:.6e} (order {p:.2f}) vs Mie {pa:.6e}")
    assert 1.0 < p < 2.5
    assert abs(lim / pa - 1) < 1e-2      # measured -4.9e-3 (order 1.48)
    # the scattered power converges at first order (BFs on interior faces only:
    # the surface charge lives in the outer layer of tetrahedra, of thickness ~h)
    orders = [math.log(es[i] / es[i + 1]) / math.log(ns[i + 1] / ns[i]) for i in range(2)]
    lim1 = Ps[2] + (Ps[2] - Ps[1]) * ns[1] / (ns[2] - ns[1])
    print(f"P_scat observed orders {orders}; first-order extrapolation {lim1:.6e} vs Mie {ps:.6e}")
    assert all(0.8 < q < 1.3 for q in orders)
    assert abs(lim1 / ps - 1) < 2e-2     # measured 1.0e-2


def _interior_field(o, J, omega, eps, pts):
    """E at points inside the body from the current: E = i Z0 J(x) / (Omega (eps - 1))
    (the VIE's constitutive relation), J(x) = sum_a J_a b_a(x) over the BFs of
    the tetrahedron holding x."""
    F = o.faces()
    V, T = o.V, o.T
    out = []
    for x in pts:
        for t in range(len(T)):
            P0 = V[T[t][0]]
            L = (V[T[t][1:]] - P0).T
            lam = np.linalg.solve(L, x - P0)
            if lam.min() >= -1e-12 and lam.sum() <= 1 + 1e-12:
                break
        else:
            raise AssertionError(f"point {x} in no tetrahedron")
        vol = np.linalg.det(L) / 6.0
        Jx = np.zeros(3, complex)
        for a in F["tetFI"][t]:
            if a < 0 or a >= o.N:
                continue
            s, q = (1.0, F["QP"][a]) if F["PT"][a] == t else (-1.0, F["QM"][a])
            Jx += J[a] * s * F["Area"][a] / (3.0 * vol) * (x - V[q])
        out.append(1j * orc_Z0 * Jx / (omega * (eps - 1)))
    return np.array(out)


orc_Z0 = mie.ZVAC


def _mean_interior_field(o, J, omega, eps):
    """Volume average of the interior field: i Z0 / (Omega (eps-1) Vol) sum_a J_a int b_a,
    int b_a = (A/3)(c_P - Q_P) - (A/3)(c_M - Q_M) over its two tetrahedra."""
    F = o.faces()
    V, T = o.V, o.T
    cen = V[T].mean(1)
    vols = np.linalg.det(V[T[:, 1:]] - V[T[:, :1]]) / 6.0
    ib = (F["Area"][:o.N, None] / 3.0) * ((cen[F["PT"][:o.N]] - V[F["QP"][:o.N]]) -
                                           (cen[F["MT"][:o.N]] - V[F["QM"][:o.N]]))
    return 1j * orc_Z0 * (J @ ib) / (omega * (eps - 1) * vols.sum())


def test_low_frequency_sphere_interior_field(bfa, orc):
    """unit-test-LFField.cc:50-69 at Omega = 0.1, eps = 10: the quasi-static
    interior field 3 E0 / (eps + 2) = 0.25 x.  On Kuhn spheres n = 4, 6, 8 the
    volume-averaged interior field converges to it; on the reference's own
    mesh (E10Sphere_533.buffgeo: Sphere_533.vmsh, CONST_EPS_10) the field at
    the EPFile.XAxis points well inside the sphere is within the coarse
    mesh's discretisation error of it."""
    omega, eps = 0.1, 10.0 + 0j
    target = 3.0 / (eps.real + 2.0)
    ns = (4, 6, 8)
    avg = []
    for n in ns:
        V, T = bfa.kuhn_sphere(n, 1.0, 0.0, 0x5EED)
        o, M, B, J, P = solve(orc, V, T, omega, eps)
        E = _mean_interior_field(o, J[:, 0], omega, eps)
        # the Kuhn split is not cubically symmetric: a small cross-polarised part
        assert abs(E[1]) < 3e-2 * abs(E[0]) and abs(E[2]) < 3e-2 * abs(E[0])
        avg.append(E[0].real)
    err = [e / target - 1 for e in avg]
    # first-order convergence (the outer layer of tetrahedra carries the surface charge)
    orders = [math.log(err[i] / err[i + 1]) / math.log(ns[i + 1] / ns[i]) for i in range(2)]
    lim1 = avg[2] + (avg[2] - avg[1]) * ns[1] / (ns[2] - ns[1])
    print(f"mean interior E_x {avg} (errors {err}, orders {orders}); first-order extrapolation "
          f"{lim1:.5f} vs {target}")
    assert abs(err[0]) > abs(err[1]) > abs(err[2])
    assert all(0.7 < q < 1.3 for q in orders)
    # measured -0.5%; (kR)^2 = 0.01 corrections to the quasi-static value are ~1e-3
    assert abs(lim1 / target - 1) < 2e-2

    V, T = bfa.read_vmsh(os.path.join(GOLDEN, "Sphere_533.vmsh"))
    assert "CONST_EPS_10" in open(os.path.join(GOLDEN, "E10Sphere_533.buffgeo")).read()
    X = np.loadtxt(os.path.join(GOLDEN, "EPFile.XAxis"))
    assert X.shape == (82, 3)
    o, M, B, J, P = solve(orc, V, T, omega, eps)
    inner = X[np.linalg.norm(X, axis=1) <= 0.5]
    E = _interior_field(o, J[:, 0], omega, eps, inner)
    ex = E[:, 0].real
    print(f"Sphere_533 interior E_x at {len(inner)} points: mean {ex.mean():.4f} min {ex.min():.4f} "
          f"max {ex.max():.4f} (3/(eps+2) = {target})")
    assert abs(ex.mean() / target - 1) < 0.2
    assert np.abs(E[:, 1:]).max() < 0.2 * target


def test_far_element_galerkin_identity(bfa, orc):
    """G_ab (non-touching, 4 x 4-point rule) = <b_a, E_b> / (i Omega Z0), E_b the
    field of unit current in BF b (4-point outer rule over b_a's tetrahedra)."""
    V, T = kuhn(bfa, 6)
    o = orc.Geometry(V, T)
    F = o.faces()
    N = o.N
    cen = np.array([V[F["V"][a]].mean(0) for a in range(N)])
    rad = np.array([max(np.linalg.norm(V[i] - cen[a]) for i in (F["QP"][a], *F["V"][a], F["QM"][a]))
                    for a in range(N)])
    r4 = orc.tetcr(4)
    rng = np.random.default_rng(7)
    for omega in (1.0, 3.0):
        g, gg = [], []
        while len(g) < 12:
            a, b = (int(x) for x in rng.integers(0, N, 2))
            if o.compare_bfs(a, b) > 0 or np.linalg.norm(cen[a] - cen[b]) < 3 * max(rad[a], rad[b]):
                continue
            J = np.zeros(N, complex)
            J[b] = 1.0
            tot = 0j
            for tet, q, s in ((F["PT"][a], F["QP"][a], 1.0), (F["MT"][a], F["QM"][a], -1.0)):
                P0 = V[T[tet][0]]
                L = V[T[tet][1:]] - P0
                vol = np.linalg.det(L) / 6.0
                pts = P0 + r4[:, :3] @ L
                E = orc.fields(o, V, T, J, omega, pts)[:, :3]
                tot += np.sum(6 * vol * r4[:, 3] * (s * F["Area"][a] / (3 * vol)) * np.einsum("ij,ij->i", pts - V[q], E))
            g.append(o.gme(a, b, omega))
            gg.append(tot / (1j * omega * orc.ZVAC))
        g, gg = np.array(g), np.array(gg)
        rel = np.linalg.norm(g - gg) / np.linalg.norm(g)
        print(f"omega={omega}: far elements vs field Galerkin, aggregate rel diff {rel:.2e}")
        assert rel < 1e-2      # measured ~2e-3..4e-3: the two cubatures of one integral


def _simplex_exp_moments(z, deg=16):
    """m[t, i] = (1/6V) int_T lambda_i exp(sum_j lambda_j z_j) dV
    = sum over multi-indices beta of z^beta (beta_i + 1) / (|beta| + 4)!
    (Dirichlet's integral int_T lambda^alpha = 6V alpha! / (|alpha| + 3)!),
    for z[t, 4] of modest size (terms of degree > deg are below 1e-14 here)."""
    m = np.zeros(z.shape, complex)
    for d in range(deg + 1):
        B = np.array([(b0, b1, b2, d - b0 - b1 - b2) for b0 in range(d + 1)
                      for b1 in range(d + 1 - b0) for b2 in range(d + 1 - b0 - b1)])
        zp = np.prod(z[:, None, :] ** B[None, :, :], axis=2)          # [t, nbeta]
        m += zp @ ((B + 1) / math.factorial(d + 4))                   # [t, 4]
    return m


def test_rhs_exact_integral(bfa, orc):
    """v_a = -1/(i Omega Z0) int b_a . E0 e^{i Omega n.x} (RHSVector.cc:62-124) against
    the exact integral: x = sum lambda_j v_j, so b_a . E0 is linear in lambda and
    the plane wave is exp(sum lambda_j z_j), z_j = i Omega n.v_j (shifted by z_0)."""
    V, T = kuhn(bfa, 6)
    o = orc.Geometry(V, T)
    F = o.faces()
    N = o.N
    E0 = np.array([0.3 + 0.2j, -0.7, 0.5j])
    nh = np.array([1.0, 2.0, 2.0]) / 3.0
    for omega in (1.0, 3.0):
        got = o.rhs(omega, orc.planewaves([(tuple(E0), tuple(nh))]))[:, 0]
        vol = np.linalg.det(V[T[:, 1:]] - V[T[:, :1]]) / 6.0
        z = 1j * omega * (V[T] @ nh)                                   # [t, 4]
        m = _simplex_exp_moments(z - z[:, :1]) * (6 * vol * np.exp(z[:, 0]))[:, None]
        want = np.zeros(N, complex)
        for side, (PT, QV, s) in enumerate(((F["PT"], F["QP"], 1.0), (F["MT"], F["QM"], -1.0))):
            t = PT[:N]
            d = (V[T[t]] - V[QV[:N]][:, None, :]) @ E0                 # [a, 4]: (v_j - Q).E0
            want += s * F["Area"][:N] / (3 * vol[t]) * np.sum(d * m[t], axis=1)
        want *= -1.0 / (1j * omega * orc.ZVAC)
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"omega={omega}: RHS vs exact integral rel err {err:.2e}")
        assert err < 1e-8      # the 33-point rule is exact to degree 7
