"""Oracle pins for buff-neq's dressed Rytov path (CPU).

- Gram matrix (GetOverlapElements, SWGVolume.cc:405-480): the closed forms
  equal the 33-point cubature of int b_a . b_b over the common tetrahedra
  (exact: the integrand has degree 2); the matrix is symmetric although each
  row is built from face a's own geometry, and positive definite.
- dressed_rytov (ComputeDressedRytovMatrix, GetFlux.cc:155-262): for real
  symmetric S and R its step order equals the closed form
  (S + V S^-1 G)^-1 R S^-1, reached by another route, and reduces to
  S^-1 R S^-1 when V = 0.
- orc_emt_power_dr (GetEMTPFT with a DRMatrix): DR = J J^H reproduces the
  J-vector EMT's scattered power, and entries above the diagonal (b < a) are
  never read (GetJJ returns DR(b, a) for a <= b).
"""
import numpy as np
import pytest

from conftest import kuhn


def _cubature_gram(orc, o, V, T):
    """int b_a . b_b over every tetrahedron both BFs live on, 33-point rule."""
    F = o.faces()
    N = o.N
    r33 = orc.tetcr(33)
    S = np.zeros((N, N))
    for t, tet in enumerate(T):
        P0 = V[tet[0]]
        L = V[tet[1:]] - P0
        vol = np.linalg.det(L) / 6.0
        x = P0 + r33[:, :3] @ L
        w = 6.0 * vol * r33[:, 3]
        bfs = []
        for b in F["tetFI"][t]:
            if b >= N:
                continue
            s, q = (1.0, V[F["QP"][b]]) if F["PT"][b] == t else (-1.0, V[F["QM"][b]])
            bfs.append((b, s * F["Area"][b] / (3 * vol) * (x - q)))
        for a, fa in bfs:
            for b, fb in bfs:
                S[a, b] += w @ (fa * fb).sum(1)
    return S


def test_gram_closed_form_is_the_overlap_integral(bfa, orc):
    V, T = kuhn(bfa, 2)
    o = orc.Geometry(V, T)
    S = o.gram()
    C = _cubature_gram(orc, o, V, T)
    err = np.abs(S - C).max() / np.abs(C).max()
    print(f"[neq] Gram closed form vs cubature {err:.1e}")
    assert err < 1e-13
    assert np.array_equal(S != 0, C != 0)
    assert np.abs(S - S.T).max() < 1e-14 * np.abs(S).max()
    assert np.linalg.eigvalsh(S).min() > 0


def test_dressed_rytov_step_order(orc):
    rng = np.random.default_rng(3)
    n = 40
    A = rng.standard_normal((n, n))
    S = A @ A.T + n * np.eye(n)
    R = rng.standard_normal((n, n))
    R = R + R.T
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Vb = np.diag(rng.standard_normal(n) + 1j * rng.standard_normal(n)) * 0.3
    Sinv = np.linalg.inv(S)
    DR = orc.dressed_rytov(G, Sinv, Vb, R)
    closed = np.linalg.solve(S + Vb @ Sinv @ G, R @ Sinv)
    assert np.abs(DR - closed).max() < 1e-12 * np.abs(closed).max()
    DR0 = orc.dressed_rytov(G, Sinv, np.zeros_like(Vb), R)
    assert np.abs(DR0 - Sinv @ R @ Sinv).max() < 1e-13 * np.abs(DR0).max()


def test_emt_power_dr_oracle(bfa, orc):
    V, T = kuhn(bfa, 2)
    o = orc.Geometry(V, T)
    N = o.N
    rng = np.random.default_rng(8)
    J = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    w = 1.1
    pw = orc.planewaves([((1, 0, 0), (0, 0, 1))])
    ref = o.emt_power(w, pw, J[:, None])[0]
    DR = np.outer(J, J.conj())                     # DR[b, a] = J_b conj(J_a)
    p = o.emt_power_dr(w, DR)
    assert abs(p[1] - ref[1]) < 1e-12 * abs(ref[1]) and p[0] == -p[1]
    U = np.triu(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)), 1)
    p2 = o.emt_power_dr(w, DR + U)
    assert abs(p2[1] - p[1]) < 1e-12 * abs(p[1])
    Lw = np.tril(rng.standard_normal((N, N)), -1)
    p3 = o.emt_power_dr(w, DR + Lw)
    assert abs(p3[1] - p[1]) > 1e-6 * abs(p[1])     # the lower triangle is read
