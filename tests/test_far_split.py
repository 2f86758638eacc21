"""The precision form's far-field kernel split, checked on the CPU.

The split kernel (bfa_assembly.hip, tp_moments_split) evaluates
g(|d + w|) = e^{ik|d+w|} / |d+w| as g(D) [T2 + E3]: T2 the second-order
Taylor part about the centroid separation d (length D, direction e), E3 the
remainder, computed in FP32 from formulas in which no two large terms
cancel.  Restated here in float64, the formulas must reproduce g itself (an
identity: a dropped term or a wrong sign breaks it), E3 must be of third
order in |w| / D (halving w divides it by ~8), and the series branch of
X3 = e^x - 1 - x - x^2/2 must meet the direct one up to the series' tail.
Evaluated in float32, E3 stays within a few units of FP32 rounding of itself
for separations |w| / D up to 0.5 -- what keeps the far-field elements
within ~1e-9 of the FP64 rule.
"""
import numpy as np

FACT = [1.0]
for n in range(1, 30):
    FACT.append(FACT[-1] * n)
KERNEL_TERMS = 12   # the kernel's series ends at x^12 / 12!


def x3(x, dtype, last=KERNEL_TERMS):
    """e^x - 1 - x - x^2/2: series to x^last / last! (|x| <= 1) or direct"""
    x = np.asarray(x, dtype=np.complex64 if dtype == np.float32 else np.complex128)
    out = np.empty_like(x)
    small = np.abs(x) <= 1.0
    xs = x[small]
    acc = np.full_like(xs, 1.0 / FACT[last])
    for n in range(last - 1, 2, -1):
        acc = acc * xs + 1.0 / FACT[n]
    out[small] = acc * xs ** 3
    xb = x[~small]
    out[~small] = np.exp(xb) - 1 - xb - xb * xb / 2
    return out


def e3(dvec, w, k, dtype=np.float64, last=KERNEL_TERMS):
    """E3 of the kernel split, in the given precision (the kernel's formulas)"""
    D = np.linalg.norm(dvec, axis=-1)
    e = (dvec / D[..., None]).astype(dtype)
    w = w.astype(dtype)
    Df = D.astype(dtype)
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    k = cdt(k)
    p = (e * w).sum(-1)
    q = w - p[..., None] * e
    s = (q * q).sum(-1)
    dp = Df + p
    r = np.sqrt(dp * dp + s)
    den = dp + r
    e1 = s / den
    Dl = p + e1
    u = Dl / Df
    x = 1j * k * Dl
    X = x3(x, dtype, last)
    i1u = 1 / (1 + u)
    U3 = u * u * u * i1u
    R3 = X * i1u - (1 + x + x * x / 2) * U3 + x * u * u + x * x * u * (u - 1) / 2
    c1 = 1j * k - 1 / Df
    c2 = c1 * c1 + 1 / (Df * Df)
    m = -s * (p + Dl) / (2 * Df * den)
    h = e1 * (p + Dl) / 2
    return R3 + c1 * m + c2 * h


def t2(dvec, w, k):
    D = np.linalg.norm(dvec, axis=-1)
    e = dvec / D[..., None]
    p = (e * w).sum(-1)
    s = (w * w).sum(-1) - p * p
    c1 = 1j * k - 1 / D
    c2 = c1 * c1 + 1 / D ** 2
    return 1 + c1 * (p + s / (2 * D)) + c2 * p * p / 2


def g(r, k):
    return np.exp(1j * k * r) / r


def geometry(rng, n, ratio):
    """separations d of length 1..3 and offsets w with |w| / D up to ratio"""
    d = rng.standard_normal((n, 3))
    d *= (1 + 2 * rng.random(n))[:, None] / np.linalg.norm(d, axis=1)[:, None]
    w = rng.standard_normal((n, 3))
    w *= (ratio * rng.random(n) * np.linalg.norm(d, axis=1))[:, None] / np.linalg.norm(w, axis=1)[:, None]
    return d, w


def test_split_reproduces_the_kernel():
    rng = np.random.default_rng(7)
    d, w = geometry(rng, 20000, 0.7)
    for k in (1.0, 1.7, 4.0, 1.7 + 0.2j, 3 + 1j):
        D = np.linalg.norm(d, axis=1)
        exact = g(np.linalg.norm(d + w, axis=1), k)
        # the algebra, with the series carried to x^25 / 25!
        split = g(D, k) * (t2(d, w, k) + e3(d, w, k, last=25))
        err = np.abs(split - exact).max() / np.abs(exact).max()
        assert err < 1e-13, (k, err)
        # the kernel's series: its tail beyond x^12 / 12! is < 1.7e-10 for |x| <= 1
        split = g(D, k) * (t2(d, w, k) + e3(d, w, k))
        err = np.abs(split - exact).max() / np.abs(exact).max()
        assert err < 2e-10, (k, err)


def test_remainder_is_third_order():
    rng = np.random.default_rng(8)
    d, w = geometry(rng, 2000, 0.05)
    for k in (1.0, 2.5 + 0.3j):
        a = np.abs(e3(d, w, k))
        b = np.abs(e3(d, w / 2, k))
        ratio = np.median(a / b)
        assert 7.6 < ratio < 8.4, (k, ratio)


def test_x3_branches_meet():
    """the kernel's series (to x^12 / 12!) against e^x - 1 - x - x^2/2 on
    circles up to the switch at |x| = 1: the difference is the series' tail,
    at most 1.1 |x|^13 / 13!"""
    for rad in (0.3, 0.6, 0.9, 1.0):
        x = rad * np.exp(1j * np.linspace(0, 2 * np.pi, 50))
        diff = np.abs(x3(x, np.float64) - (np.exp(x) - 1 - x - x * x / 2)).max()
        assert diff <= 1.1 * rad ** 13 / FACT[13] + 1e-15, (rad, diff)


def test_float32_remainder_keeps_its_relative_accuracy():
    rng = np.random.default_rng(9)
    d, w = geometry(rng, 20000, 0.5)
    for k in (1.0, 1.7 + 0.2j):
        ref = e3(d, w, k)
        f32 = e3(d.astype(np.float32).astype(np.float64), w.astype(np.float32).astype(np.float64), k,
                 dtype=np.float32).astype(np.complex128)
        # ~2.5 units of FP32 rounding of the remainder at the median
        rel = np.abs(f32 - ref) / np.abs(ref)
        assert np.median(rel) < 3e-7, (k, np.median(rel))
        # everywhere within a few units of rounding of the size of its terms,
        # (|w| / D)^3 (1 + |k D|^3) (where its terms cancel, the remainder
        # is smaller than they are and its relative error larger)
        D = np.linalg.norm(d, axis=1)
        size = (np.linalg.norm(w, axis=1) / D) ** 3 * (1 + np.abs(k * D) ** 3)
        assert (np.abs(f32 - ref) / size).max() < 1.5e-6, k
