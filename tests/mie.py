"""Mie series for a homogeneous sphere in a plane wave (test infrastructure).

Bohren & Huffman, "Absorption and Scattering of Light by Small Particles",
eqs. 4.53 (a_n, b_n from Riccati-Bessel functions) and 4.61-4.62 (cross
sections); the reference compares buff-scatter with the same series
(doc/docs/examples/MieScattering/Mie.math, index.md:169-177), with power =
cross section x the incident flux 1/(2 Z0) of a unit plane wave.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn

ZVAC = 376.73031346177


def _psi(n, z):
    return z * spherical_jn(n, z)


def _dpsi(n, z):
    return spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)


def _h1(n, z, d=False):
    return spherical_jn(n, z, derivative=d) + 1j * spherical_yn(n, z, derivative=d)


def _xi(n, z):
    return z * _h1(n, z)


def _dxi(n, z):
    return _h1(n, z) + z * _h1(n, z, True)


def coefficients(x, m, nmax=None):
    """a_n, b_n for size parameter x = kR and relative index m = sqrt(eps)."""
    nmax = nmax or int(x + 4 * x ** (1.0 / 3.0) + 12)
    n = np.arange(1, nmax + 1)
    mx = m * x
    a = (m * _psi(n, mx) * _dpsi(n, x) - _psi(n, x) * _dpsi(n, mx)) / \
        (m * _psi(n, mx) * _dxi(n, x) - _xi(n, x) * _dpsi(n, mx))
    b = (_psi(n, mx) * _dpsi(n, x) - m * _psi(n, x) * _dpsi(n, mx)) / \
        (_psi(n, mx) * _dxi(n, x) - m * _xi(n, x) * _dpsi(n, mx))
    return n, a, b


def powers(k, R, eps):
    """(P_abs, P_scat) of a sphere of radius R (um) at wavenumber k (1/um,
    = omega in units of 3e14 rad/s) under a unit-amplitude plane wave."""
    x = k * R
    n, a, b = coefficients(x, np.sqrt(complex(eps)))
    s_ext = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * (a + b).real)
    s_sca = 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return (s_ext - s_sca) / (2 * ZVAC), s_sca / (2 * ZVAC)


def richardson(hs, vals):
    """Limit h -> 0 of three values assumed to follow v(h) = v0 + C h^p with
    unknown p (solved from the three points); returns (v0, p)."""
    h1, h2, h3 = hs
    v1, v2, v3 = vals
    r = (v1 - v2) / (v2 - v3)
    lo, hi = 0.05, 8.0
    f = lambda p: (h1 ** p - h2 ** p) / (h2 ** p - h3 ** p) - r
    if f(lo) * f(hi) > 0:
        raise ValueError("no power law through the three values")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    p = 0.5 * (lo + hi)
    return v3 + (v3 - v2) * h3 ** p / (h2 ** p - h3 ** p), p
