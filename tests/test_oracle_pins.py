"""Pins of the CPU oracle against closed forms and independent mathematics.

The reference's own tests hold no numeric golden vectors for this path
(its make-check programs assert nothing), so each oracle function is pinned
here by a check a plausible mistake would fail:

  TetCR rules          exact monomial integrals (degree 2 / 4 / 7), orbit symmetry
  Taylor-Duffy         unit cube = 6 Kuhn tetrahedra (every pair shares 2-4
                       vertices): int int |x-x'|^-1 and |x-x'|^+1 (Robbins'
                       constant) and the bilinear moment int int (x-c).(x'-c)|x-x'|^p
                       against an independent covariogram reduction; p=0
                       moments against closed forms on jittered pairs
  ExpRel series        mpmath e^z - 1 - z - z^2/2
  Gram / chi^-1 block  textbook tetrahedron second moments
  static p=0 terms     (int b_a).(int b_b) / 4 pi and int div = 0
"""
import itertools
import math

import mpmath
import numpy as np
import pytest

from conftest import kuhn


# ---------------------------------------------------------------- TetCR rules
def _unit_tet_monomial(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("npts,degree", [(4, 2), (16, 4), (33, 7)])
def test_tetcr_exactness(orc, npts, degree):
    cr = orc.tetcr(npts)
    assert abs(cr[:, 3].sum() - 1.0 / 6.0) < 1e-15
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                q = np.sum(cr[:, 3] * cr[:, 0] ** a * cr[:, 1] ** b * cr[:, 2] ** c)
                assert abs(q - _unit_tet_monomial(a, b, c)) < 2e-15, (a, b, c)
    # not exact one degree higher (the stated degree is tight)
    errs = [abs(np.sum(cr[:, 3] * cr[:, 0] ** (degree + 1)) - _unit_tet_monomial(degree + 1, 0, 0))]
    errs += [abs(np.sum(cr[:, 3] * (cr[:, 0] * cr[:, 1]) ** ((degree + 1) // 2) * cr[:, 2] ** ((degree + 1) % 2))
                 - _unit_tet_monomial((degree + 1) // 2, (degree + 1) // 2, (degree + 1) % 2))]
    assert max(errs) > 1e-9


@pytest.mark.parametrize("npts", [4, 16, 33])
def test_tetcr_permutation_invariant(orc, npts):
    cr = orc.tetcr(npts)
    bary = np.column_stack([1 - cr[:, :3].sum(1), cr[:, :3]])
    pts = {tuple(np.round(b, 12)) for b in bary}
    for perm in itertools.permutations(range(4)):
        assert {tuple(np.round(b[list(perm)], 12)) for b in bary} == pts


# ---------------------------------------------------------- Taylor-Duffy KATs
def _kuhn_cube_tets():
    tets = []
    for perm in itertools.permutations(range(3)):
        v, cur = [np.zeros(3)], np.zeros(3)
        for a in perm:
            cur = cur.copy()
            cur[a] = 1.0
            v.append(cur)
        tets.append(np.array(v))
    return tets


def _shared_first(TA, TB):
    ka = [tuple(p) for p in TA]
    kb = [tuple(p) for p in TB]
    shared = [p for p in ka if p in kb]
    A = shared + [p for p in ka if p not in shared]
    B = shared + [p for p in kb if p not in shared]
    return len(shared), np.array(A, float), np.array(B, float)


def _cube_reference(p):
    """int_cube int_cube |x-x'|^p and int int (x-c).(x'-c) |x-x'|^p, c = cube centre.

    Covariogram reduction: for z = x'-x the overlap of the cube with its shift
    is a box, giving int |z|^p prod(1-|z_j|) dz and, for the bilinear moment,
    sum_i g(z_i) prod_{j!=i}(1-|z_j|) with g(t) = int_0^{1-t} (x-1/2)(x+t-1/2) dx.
    The positive octant is split into 3 pyramids z = s (1, a, b), which makes
    the integrand polynomial in s; (a, b) use 60-point Gauss-Legendre.
    """
    xg, wg = np.polynomial.legendre.leggauss(60)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    sg, swt = np.polynomial.legendre.leggauss(12)
    sg, swt = 0.5 * (sg + 1), 0.5 * swt
    A, B = np.meshgrid(xg, xg, indexing="ij")
    W = np.outer(wg, wg)
    norm = np.sqrt(1 + A * A + B * B)

    def g(t):
        # int_0^{1-t} (x-1/2)(x+t-1/2) dx
        # (x - 1/2)(x + t - 1/2) = x^2 - (1 - t) x + (1/2)(1/2 - t)
        u = 1 - t
        return u ** 3 / 3 - (1.0 - t) * u ** 2 / 2 + 0.5 * (0.5 - t) * u

    m0 = mx = 0.0
    for s, ws in zip(sg, swt):
        z0, z1, z2 = s * np.ones_like(A), s * A, s * B
        jac = s ** 2 * (s * norm) ** p
        cov = (1 - z0) * (1 - z1) * (1 - z2)
        bil = g(z0) * (1 - z1) * (1 - z2) + g(z1) * (1 - z0) * (1 - z2) + g(z2) * (1 - z0) * (1 - z1)
        m0 += ws * np.sum(W * jac * cov)
        mx += ws * np.sum(W * jac * bil)
    return 8 * 3 * m0, 8 * 3 * mx


def _cube_td(orc, order):
    tets = _kuhn_cube_tets()
    c = np.full(3, 0.5)
    tot = np.zeros((3, 8))
    for TA in tets:
        for TB in tets:
            ncv, A, B = _shared_first(TA, TB)
            assert ncv >= 2
            tot += orc.td_moments(ncv, A, B, c, c, order)
    return tot


def test_td_cube_kat(orc):
    tot = _cube_td(orc, 16)
    robbins = (4 + 17 * math.sqrt(2) - 6 * math.sqrt(3) + 21 * math.log(1 + math.sqrt(2))
               + 42 * math.log(2 + math.sqrt(3)) - 7 * math.pi) / 105
    m0m1, mxm1 = _cube_reference(-1)
    m0p1, mxp1 = _cube_reference(1)
    assert abs(m0p1 - robbins) < 1e-14          # the reduction itself reproduces Robbins' constant
    assert abs(tot[1, 0] - 1.0) < 1e-13         # p = 0: volume^2
    assert abs(tot[0, 0] / m0m1 - 1) < 1e-10    # p = -1 (Coulomb self-energy of the cube)
    assert abs(tot[2, 0] / robbins - 1) < 1e-11  # p = +1 (mean distance)
    assert abs(tot[0, 7] / mxm1 - 1) < 1e-9     # bilinear moment, p = -1
    assert abs(tot[2, 7] / mxp1 - 1) < 1e-10    # bilinear moment, p = +1
    # by the cube's reflection symmetry the linear moments vanish
    assert np.abs(tot[:, 1:7]).max() < 1e-12


def test_td_converges_with_order(orc):
    e = [abs(_cube_td(orc, n)[0, 0] / _cube_reference(-1)[0] - 1) for n in (6, 10, 14)]
    assert e[0] > e[1] > e[2]


def _tet_volume(T):
    return abs(np.linalg.det(np.array([T[1] - T[0], T[2] - T[0], T[3] - T[0]]))) / 6


@pytest.mark.parametrize("ncv", [2, 3, 4])
def test_td_p0_closed_forms(orc, ncv):
    """p = 0: M0 = V_A V_B, Mx = V_A V_B (c_A - refA), Mxx' = V_A V_B (c_A-refA).(c_B-refB)."""
    rng = np.random.default_rng(ncv)
    # jittered Kuhn tetrahedra that share ncv vertices (non-overlapping)
    tets = _kuhn_cube_tets()
    for TA, TB in itertools.product(tets, tets):
        k, A, B = _shared_first(TA, TB)
        if k != ncv:
            continue
        jit = {}
        for p in list(map(tuple, A)) + list(map(tuple, B)):
            jit.setdefault(p, rng.uniform(-0.06, 0.06, 3))
        A = np.array([a + jit[tuple(a)] for a in A])
        B = np.array([b + jit[tuple(b)] for b in B])
        refA, refB = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        m = orc.td_moments(ncv, A, B, refA, refB, 8)[1]
        VA, VB = _tet_volume(A), _tet_volume(B)
        dA, dB = A.mean(0) - refA, B.mean(0) - refB
        assert abs(m[0] - VA * VB) < 1e-14
        np.testing.assert_allclose(m[1:4], VA * VB * dA, rtol=0, atol=1e-13)
        np.testing.assert_allclose(m[4:7], VA * VB * dB, rtol=0, atol=1e-13)
        assert abs(m[7] - VA * VB * dA @ dB) < 1e-13


# ------------------------------------------------------------------- ExpRel
def test_exprel_series(orc):
    mpmath.mp.dps = 40
    for mag in np.logspace(-3, 1, 17):
        for ph in (0.5 * np.pi, 0.3, 2.0, np.pi):   # ikr is purely imaginary on the real axis
            z = mag * complex(math.cos(ph), math.sin(ph))
            ref = mpmath.exp(mpmath.mpc(z)) - 1 - mpmath.mpc(z) - mpmath.mpc(z) ** 2 / 2
            ref = complex(ref)
            got = orc.exprel3(z)
            assert abs(got - ref) <= 1e-7 * abs(ref), (z, got, ref)


# ------------------------------------------------------------ Gram / chi^-1
def _tet_second_moment(P):
    s = P.sum(0)
    return (np.einsum("ia,ib->ab", P, P) + np.outer(s, s)) / 20.0


def test_gram_closed_form(bfa, orc):
    V, T = kuhn(bfa, 2)
    o = orc.Geometry(V, T)
    f = o.faces()
    eps = 4 + 0.1j
    omega = 0.7
    mat = orc.material("iso", eps=eps)
    chi_inv = 1.0 / (eps - 1.0)
    for a in range(o.N):
        idx, val = o.overlaps(a, omega, mat)
        for b, v in zip(idx, val):
            gram = 0.0
            for ta, qa, sa in ((f["PT"][a], f["QP"][a], 1), (f["MT"][a], f["QM"][a], -1)):
                for tb, qb, sb in ((f["PT"][b], f["QP"][b], 1), (f["MT"][b], f["QM"][b], -1)):
                    if ta != tb or (a == b and sa != sb):
                        continue
                    Pt = V[T[ta]]
                    vol = np.linalg.det(np.array([Pt[1] - Pt[0], Pt[2] - Pt[0], Pt[3] - Pt[0]])) / 6
                    QA, QB = V[qa], V[qb]
                    S = _tet_second_moment(Pt) * abs(vol)
                    c = Pt.mean(0)
                    integral = np.trace(S) - (QA + QB) @ c * abs(vol) + QA @ QB * abs(vol)
                    gram += sa * sb * f["Area"][a] * f["Area"][b] / (9 * vol * vol) * integral * np.sign(vol)
            expect = -chi_inv / omega ** 2 * gram
            assert abs(v - expect) < 1e-13 * max(1.0, abs(expect)), (a, b, v, expect)


# ------------------------------------------------------- static p = 0 terms
def test_static_p0_terms(bfa, orc):
    V, T = kuhn(bfa, 2)
    o = orc.Geometry(V, T)
    f = o.faces()
    tab = o.static_table(8)
    rowptr, cols = o.touch_pairs()

    def int_b(a):
        s = np.zeros(3)
        for t, q, sg in ((f["PT"][a], f["QP"][a], 1), (f["MT"][a], f["QM"][a], -1)):
            Pt = V[T[t]]
            vol = abs(np.linalg.det(np.array([Pt[1] - Pt[0], Pt[2] - Pt[0], Pt[3] - Pt[0]]))) / 6
            s += sg * f["Area"][a] / (3 * vol) * vol * (Pt.mean(0) - V[q])
        return s

    ib = np.array([int_b(a) for a in range(o.N)])
    for a in range(o.N):
        for j in range(rowptr[a], rowptr[a + 1]):
            b = cols[j]
            assert abs(tab[j, 2] - ib[a] @ ib[b] / (4 * np.pi)) < 1e-14
            assert abs(tab[j, 3]) < 1e-13
