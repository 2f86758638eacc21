"""Python driver of the CPU oracle.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as what results are compared against, never by
the product path.

Wraps liboracle.so (the C restatement of the reference algorithm, see
buffem_oracle.h) and adds the dense solve through LAPACK (scipy's zgetrf /
zgetrs), which is what the reference's HMatrix::LUFactorize / LUSolve call
(src/applications/buff-scatter/buff-scatter.cc:285,314).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


class orc_material(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("p", ctypes.c_double * 6)]


class orc_planewave(ctypes.Structure):
    _fields_ = [("E0", ctypes.c_double * 6), ("dir", ctypes.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "orc_geom_build": (I, [I, P, I, P, P]),
            "orc_geom_free": (None, [P]),
            "orc_geom_info": (I, [P, P, P, P, P]),
            "orc_geom_faces": (None, [P, P, P, P, P, P, P, P, P, P]),
            "orc_compare_bfs": (I, [P, I, I]),
            "orc_compare_tets": (I, [P, I, I, P, P]),
            "orc_tetcr": (P, [I]),
            "orc_exprel": (None, [D, D, I, I, P]),
            "orc_td_moments": (None, [I, P, P, P, P, I, P]),
            "orc_fibbi": (None, [P, I, I, I, P]),
            "orc_gme": (None, [P, I, I, D, D, P, P]),
            "orc_touch_pairs": (I, [P, P, P]),
            "orc_static_table": (None, [P, P, P, I, P]),
            "orc_assemble": (None, [P, D, D, P, P, P, P, P]),
            "orc_assemble_rows": (None, [P, D, D, I, P, P, P, P, P]),
            "orc_overlaps": (I, [P, I, D, P, P, P]),
            "orc_rhs": (None, [P, D, P, I, P]),
            "orc_emt_power": (None, [P, D, P, I, P, P]),
            "orc_emt_power_dr": (None, [P, D, P, P]),
            "orc_tp_block": (None, [P, P, D, D, I, I, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def material(kind="iso", **kw):
    m = orc_material()
    p = [0.0] * 6
    if kind == "iso":
        e = complex(kw["eps"])
        m.kind, p[0], p[1] = 0, e.real, e.imag
    elif kind == "diag":
        es = [complex(x) for x in kw["eps"]]
        m.kind = 1
        p = [es[0].real, es[0].imag, es[1].real, es[1].imag, es[2].real, es[2].imag]
    elif kind == "sincos":
        m.kind = 2
        p = [kw.get("base", 4.0), kw.get("imag", 0.1), kw.get("amp", 2.0), 0, 0, 0]
    elif kind == "janus":
        up, dn = complex(kw["up"]), complex(kw["down"])
        m.kind = 3
        p = [up.real, up.imag, dn.real, dn.imag, 0, 0]
    for i in range(6):
        m.p[i] = p[i]
    return m


def planewaves(specs):
    arr = (orc_planewave * len(specs))()
    for j, (e0, d) in enumerate(specs):
        e0 = [complex(x) for x in e0]
        for c in range(3):
            arr[j].E0[2 * c] = e0[c].real
            arr[j].E0[2 * c + 1] = e0[c].imag
            arr[j].dir[c] = float(d[c])
    return arr


def tetcr(npts):
    p = lib().orc_tetcr(npts)
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(npts, 4)).copy()


def exprel3(z):
    out = np.zeros(2)
    lib().orc_exprel(z.real, z.imag, 3, 3, out.ctypes.data)
    return complex(out[0], out[1])


def td_moments(ncv, VA, VB, cA, cB, order):
    out = np.zeros((3, 8))
    VA = np.ascontiguousarray(VA, float)
    VB = np.ascontiguousarray(VB, float)
    cA = np.ascontiguousarray(cA, float)
    cB = np.ascontiguousarray(cB, float)
    lib().orc_td_moments(ncv, VA.ctypes.data, VB.ctypes.data, cA.ctypes.data, cB.ctypes.data, order,
                         out.ctypes.data)
    return out


def tp_block(VS, VT, k, npts=4, level=0):
    """4x4 complex block of a tetrahedron pair (face i of VS, face j of VT)."""
    out = np.zeros(32)
    VS = np.ascontiguousarray(VS, float).reshape(12)
    VT = np.ascontiguousarray(VT, float).reshape(12)
    k = complex(k)
    lib().orc_tp_block(VS.ctypes.data, VT.ctypes.data, k.real, k.imag, int(npts), int(level), out.ctypes.data)
    return (out[0::2] + 1j * out[1::2]).reshape(4, 4)


def tp_features(VS, VT, omega):
    """Canonical surrogate inputs of a tetrahedron pair in the source frame
    (origin = source centroid, e1 along v1 - v0, e2 in the v0 v1 v2 plane,
    e3 = e1 x e2), lengths over the source's mean edge h: the 4 target
    vertices (12), the source's shape v1-v0 = (a,0,0), v2-v0 = (b,c,0),
    v3-v0 = (d,e,f) (6), then log(omega h)."""
    VS = np.asarray(VS, float).reshape(4, 3)
    VT = np.asarray(VT, float).reshape(4, 3)
    c = VS.mean(0)
    h = np.mean([np.linalg.norm(VS[j] - VS[i]) for i in range(4) for j in range(i + 1, 4)])
    e1 = (VS[1] - VS[0]) / np.linalg.norm(VS[1] - VS[0])
    u = VS[2] - VS[0]
    e2 = u - (u @ e1) * e1
    e2 /= np.linalg.norm(e2)
    e3 = np.cross(e1, e2)
    R = np.stack([e1, e2, e3])
    f = ((VT - c) @ R.T / h).ravel()
    sh = (VS[1:] - VS[0]) @ R.T / h
    return np.concatenate([f, [sh[0, 0], sh[1, 0], sh[1, 1], sh[2, 0], sh[2, 1], sh[2, 2]], [np.log(omega * h)]])


class Geometry:
    def __init__(self, V, T):
        self.V = np.ascontiguousarray(V, float)
        self.T = np.ascontiguousarray(T, np.int32)
        h = P()
        rc = lib().orc_geom_build(len(self.V), self.V.ctypes.data, len(self.T), self.T.ctypes.data,
                                  ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"orc_geom_build failed ({rc})")
        self.h = h
        nv, nt, nf, nft = I(), I(), I(), I()
        lib().orc_geom_info(h, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(nf), ctypes.byref(nft))
        self.nv, self.nt, self.N, self.nfaces = nv.value, nt.value, nf.value, nft.value
        self._touch = None

    def __del__(self):
        try:
            lib().orc_geom_free(self.h)
        except Exception:
            pass

    def faces(self):
        n = self.nfaces
        out = {k: np.zeros(n, dtype=np.int32) for k in ("QP", "QM", "PT", "MT", "PI", "MI")}
        fV = np.zeros((n, 3), dtype=np.int32)
        area = np.zeros(n)
        tetFI = np.zeros((self.nt, 4), dtype=np.int32)
        lib().orc_geom_faces(self.h, fV.ctypes.data, out["QP"].ctypes.data, out["QM"].ctypes.data,
                             out["PT"].ctypes.data, out["MT"].ctypes.data, out["PI"].ctypes.data,
                             out["MI"].ctypes.data, area.ctypes.data, tetFI.ctypes.data)
        out["V"] = fV
        out["Area"] = area
        out["tetFI"] = tetFI
        return out

    def touch_pairs(self):
        if self._touch is None:
            rowptr = np.zeros(self.N + 1, dtype=np.int32)
            nnz = lib().orc_touch_pairs(self.h, rowptr.ctypes.data, None)
            cols = np.zeros(nnz, dtype=np.int32)
            lib().orc_touch_pairs(self.h, rowptr.ctypes.data, cols.ctypes.data)
            self._touch = (rowptr, cols)
        return self._touch

    def compare_bfs(self, a, b):
        return lib().orc_compare_bfs(self.h, a, b)

    def static_table(self, td_order=12):
        rowptr, cols = self.touch_pairs()
        t = np.zeros((len(cols), 6))
        lib().orc_static_table(self.h, rowptr.ctypes.data, cols.ctypes.data, td_order, t.ctypes.data)
        return t

    def fibbi(self, a, b, td_order=12):
        G = np.zeros(6)
        lib().orc_fibbi(self.h, a, b, td_order, G.ctypes.data)
        return G

    def gme(self, a, b, omega, fibbi=None):
        out = np.zeros(2)
        omega = complex(omega)
        fp = None if fibbi is None else np.ascontiguousarray(fibbi, float)
        lib().orc_gme(self.h, a, b, omega.real, omega.imag, None if fp is None else fp.ctypes.data,
                      out.ctypes.data)
        return complex(out[0], out[1])

    def overlaps(self, a, omega, mat):
        idx = np.zeros(7, dtype=np.int32)
        val = np.zeros((7, 2))
        n = lib().orc_overlaps(self.h, a, float(omega), ctypes.byref(mat), idx.ctypes.data, val.ctypes.data)
        return idx[:n], val[:n, 0] + 1j * val[:n, 1]

    def assemble(self, omega, mat, table=None, td_order=12):
        """Dense VIE matrix as a numpy array M[r, c] (complex128)."""
        rowptr, cols = self.touch_pairs()
        if table is None:
            table = self.static_table(td_order)
        table = np.ascontiguousarray(table, float)
        N = self.N
        Mcol = np.zeros((N, N), dtype=np.complex128)       # Mcol[c, r] = M(r, c)
        omega = complex(omega)
        lib().orc_assemble(self.h, omega.real, omega.imag, ctypes.byref(mat), rowptr.ctypes.data,
                           cols.ctypes.data, table.ctypes.data, Mcol.ctypes.data)
        return Mcol.T.copy()

    def assemble_colmajor(self, omega, mat, table):
        """Dense VIE matrix M[r, c] as a Fortran-ordered array (no transposed copy:
        LAPACK can factor it in place, which matters at N ~ 20k)."""
        rowptr, cols = self.touch_pairs()
        table = np.ascontiguousarray(table, float)
        N = self.N
        Mcol = np.zeros((N, N), dtype=np.complex128)       # Mcol[c, r] = M(r, c)
        omega = complex(omega)
        lib().orc_assemble(self.h, omega.real, omega.imag, ctypes.byref(mat), rowptr.ctypes.data,
                           cols.ctypes.data, table.ctypes.data, Mcol.ctypes.data)
        return Mcol.T

    def assemble_rows(self, omega, rows, table=None):
        """Upper-triangle entries (b >= a) of the given rows, zero static records if table is None."""
        rowptr, cols = self.touch_pairs()
        rows = np.ascontiguousarray(rows, np.int32)
        out = np.zeros((len(rows), self.N), dtype=np.complex128)
        omega = complex(omega)
        tp = None if table is None else np.ascontiguousarray(table, float)
        lib().orc_assemble_rows(self.h, omega.real, omega.imag, len(rows), rows.ctypes.data, rowptr.ctypes.data,
                                cols.ctypes.data, None if tp is None else tp.ctypes.data, out.ctypes.data)
        return out

    def rhs(self, omega, pws):
        N = self.N
        B = np.zeros((len(pws), N), dtype=np.complex128)
        lib().orc_rhs(self.h, float(omega), pws, len(pws), B.ctypes.data)
        return B.T.copy()

    def emt_power(self, omega, pws, J):
        Jc = np.ascontiguousarray(np.asarray(J).T, dtype=np.complex128)   # columns contiguous
        out = np.zeros(2 * len(pws))
        lib().orc_emt_power(self.h, float(omega), pws, len(pws), Jc.ctypes.data, out.ctypes.data)
        return out.reshape(-1, 2)

    def emt_power_dr(self, omega, DR):
        """GetEMTPFT with a DRMatrix, no incident field: (P_abs, P_scat)."""
        D = np.asfortranarray(DR, dtype=np.complex128)
        out = np.zeros(2)
        lib().orc_emt_power_dr(self.h, float(omega), D.ctypes.data, out.ctypes.data)
        return out

    def gram(self):
        """Dense SWG overlap (Gram) matrix, GetOverlapElements row by row
        (SWGVolume.cc:405-480): for each of BF a's tetrahedra (P then M), the
        self term Area_a^2 (|X-Q|^2/20 + sum|L|^2/180)/Vol and, per other
        interior face b of the tetrahedron,
        2 s_a s_b Area_a Area_b/(3 Vol) (3|X-Q|^2/40 + sum|L|^2/120 + (X-Q).(Q_a-Q_b)/8),
        with L_i = V_i - Q_a over face a's vertices and X its centroid;
        Vol is the signed volume (SWGVolume.cc:70-80)."""
        F = self.faces()
        V, T, N = self.V, self.T, self.N
        S = np.zeros((N, N))
        for a in range(N):
            fv = V[F["V"][a]]
            for sA, tet, qa in ((1.0, F["PT"][a], F["QP"][a]), (-1.0, F["MT"][a], F["QM"][a])):
                QA = V[qa]
                vol = np.linalg.det(V[T[tet][1:]] - V[T[tet][0]]) / 6.0
                L = fv - QA
                l2 = (L * L).sum()
                xmq = fv.mean(0) - QA
                x2 = xmq @ xmq
                S[a, a] += F["Area"][a] ** 2 * (x2 / 20.0 + l2 / 180.0) / vol
                for b in F["tetFI"][tet]:
                    if b == a or b >= N:
                        continue
                    sB, QB = (1.0, V[F["QP"][b]]) if F["PT"][b] == tet else (-1.0, V[F["QM"][b]])
                    pre = 2.0 * sA * sB * F["Area"][a] * F["Area"][b] / (3.0 * vol)
                    S[a, b] = pre * (3.0 * x2 / 40.0 + l2 / 120.0 + xmq @ (QA - QB) / 8.0)
        return S


def dressed_rytov(G, Sinv, Vb, R):
    """ComputeDressedRytovMatrix (buff-neq GetFlux.cc:155-262), step by step
    as the reference orders it: M3 = 1 + S^-1 (V (S^-1 G)); M1 = (S^-1 R) S^-1;
    DR = M3 \\ M1^H (LAPACK zgetrf/zgetrs)."""
    M2 = Sinv @ G
    M1 = Vb @ M2
    M3 = Sinv @ M1 + np.eye(len(G))
    M1 = (Sinv @ R) @ Sinv
    return lu_solve(M3, M1.conj().T)


def lu_solve(M, B, overwrite=False):
    """LAPACK zgetrf + zgetrs (the reference's HMatrix::LUFactorize/LUSolve).
    overwrite=True factors a Fortran-ordered M in place."""
    import scipy.linalg as sl
    lu, piv = sl.lu_factor(M, check_finite=False, overwrite_a=overwrite)
    return sl.lu_solve((lu, piv), B, check_finite=False)


ZVAC = 376.73031346177


def fields(geom, V, T, J, omega, X, pws=()):
    """SWGGeometry::GetFields restated in numpy (GetFields.cc:181-224, 285-382).

    Per point X and BF a: the volume integral of Get1BFFields_VI over the
    BF's two tetrahedra with TetCR4 when |X - face centroid| > 10 x the
    face radius (InitFaceList.cc:194-200: max distance from the centroid to
    QP, V1, V2, V3, QM), else TetCR16 (BFInt, Cubature.cc:234-258);
    E = i w Z0 G.b, H = -i w C.b with the vacuum dyadic pair of CalcGC
    (restated from the Helmholtz Green's function: G = (I + grad grad/k^2) g,
    C.b = -(1 + i/kr) g Rhat x b, g = e^{ikr}/4 pi r, k = w), times J_a; plus
    the incident plane waves (E0 e^{ik n.x}, H = n x E / Z0).
    """
    F = geom.faces()
    Vt = np.asarray(V, dtype=np.float64)
    Tt = np.asarray(T)
    rules = {4: tetcr(4), 16: tetcr(16)}
    N = len(J)
    k = float(omega)
    out = np.zeros((len(X), 6), dtype=complex)
    for p, x in enumerate(np.asarray(X, dtype=np.float64)):
        Et = np.zeros(3, complex)
        Ht = np.zeros(3, complex)
        for a in range(N):
            if J[a] == 0:
                continue                    # no current in this BF: no field
            fv = Vt[F["V"][a]]
            cen = fv.mean(0)
            rad = max(np.linalg.norm(Vt[i] - cen) for i in (F["QP"][a], *F["V"][a], F["QM"][a]))
            npts = 4 if np.linalg.norm(x - cen) > 10.0 * rad else 16
            r4 = rules[npts]
            E = np.zeros(3, complex)
            H = np.zeros(3, complex)
            for tet, qv, sign in ((F["PT"][a], F["QP"][a], 1.0), (F["MT"][a], F["QM"][a], -1.0)):
                P0 = Vt[Tt[tet][0]]
                L = Vt[Tt[tet][1:]] - P0
                vol = np.linalg.det(L) / 6.0
                pf = sign * F["Area"][a] / (3.0 * vol)
                for u1, u2, u3, w in r4:
                    xs = P0 + u1 * L[0] + u2 * L[1] + u3 * L[2]
                    b = pf * (xs - Vt[qv])
                    R = x - xs
                    r = np.linalg.norm(R)
                    u = R / r
                    g = np.exp(1j * k * r) / (4 * np.pi * r)
                    kr = k * r
                    G = g * ((1 + 1j / kr - 1 / kr ** 2) * np.eye(3) + (-1 - 3j / kr + 3 / kr ** 2) * np.outer(u, u))
                    E += 6 * vol * w * (G @ b)
                    H += 6 * vol * w * (-(1 + 1j / kr) * g * np.cross(u, b))
            Et += J[a] * 1j * k * ZVAC * E
            Ht += J[a] * (-1j * k) * H
        for e0, d in pws:
            e0 = np.asarray(e0, complex)
            d = np.asarray(d, float)
            Ei = e0 * np.exp(1j * k * d @ x)
            Et += Ei
            Ht += np.cross(d, Ei) / ZVAC
        out[p, :3], out[p, 3:] = Et, Ht
    return out
