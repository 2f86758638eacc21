#!/usr/bin/env python3
"""Generate the Taylor-Duffy region table ``td_regions.h``.

Singular static integrals over a pair of tetrahedra that share ncv = 2, 3 or
4 vertices are evaluated by the Taylor-Duffy strategy of the reference
(src/libs/libTTaylorDuffy/TTaylorDuffy.cc:284-343, doc/docs/tex/TaylorDuffy.tex):
both tetrahedra are parameterised from the shared vertex V1,

    x  = V1 + xi1*A  + xi2*B  + xi3*C ,   1 >= xi1 >= xi2 >= xi3 >= 0
    x' = V1 + eta1*A'+ eta2*B'+ eta3*C',  1 >= eta1>= eta2>= eta3>= 0

with A = V2-V1, B = V3-V2, C = V4-V3 (primed for the second tetrahedron,
A'=A, B'=B, C'=C for the shared edges).  The displacement x'-x depends only on
the D "singular" variables s (relative coordinates u_i = eta_i - xi_i along
shared directions plus the free coordinates of the unshared vertices):

    ncv=4: s=(u1,u2,u3)                rem=(xi1,xi2,xi3)
    ncv=3: s=(u1,u2,xi3,eta3)          rem=(xi1,xi2)
    ncv=2: s=(u1,xi2,xi3,eta2,eta3)    rem=(xi1,)

For fixed s the remaining variables range over a nested simplex of size
sigma(s) = 1 - l(s), where l is piecewise linear and homogeneous.  Each
linearity piece ("pattern") is a polyhedral cone; the cone truncated by
l(s) <= 1 is a pyramid with apex at s = 0, i.e. at the singularity.  This
script enumerates those pyramids and triangulates their bases; the kernel
integral then becomes an exact radial integral (polynomial in the radial
variable w) times a smooth integral over the base simplices.

The reference derives its own (18-region) decomposition symbolically and
tabulates the polynomial coefficients; this table is an independent
derivation of the same construction, generated here from the constraint
definitions.  Output: C header with, per case, the pattern offset matrices
O_P (rem = O_P s + sigma * tau) and the base simplices (pattern id, D vertices,
|det|).
"""
import itertools
import os
import sys

import numpy as np
from scipy.spatial import Delaunay


def case_definition(ncv):
    """Return (D, m, terms, pure, offset_fn).

    terms: list of (kind, f0, f1, sign) where kind is 'max' or 'min', f0/f1
    linear forms (length-D vectors) and sign the coefficient of the chosen
    branch in sigma(s) = 1 + sum sign*branch.
    pure: list of linear forms g with g.s >= 0 on the whole domain.
    offset_fn(choices) -> (m x D) matrix O such that rem = O s + sigma*tau.
    """
    if ncv == 4:
        D = 3
        u1, u2, u3 = np.eye(3)
        z = np.zeros(3)
        terms = [("max", z, u1, -1.0),         # a = 1 - max(0,u1)
                 ("min", z, u1 - u2, +1.0),    # b = min(0,u1-u2)
                 ("min", z, u2 - u3, +1.0),    # c = min(0,u2-u3)
                 ("max", z, -u3, -1.0)]        # d = max(0,-u3)
        pure = []

        def offset(br):
            b, c, d = br[1], br[2], br[3]
            return np.array([d - c - b, d - c, d])
        return D, 3, terms, pure, offset
    if ncv == 3:
        D = 4
        u1, u2, x3, e3 = np.eye(4)
        z = np.zeros(4)
        terms = [("max", z, u1, -1.0),         # a = 1 - max(0,u1)
                 ("min", z, u1 - u2, +1.0),    # b = min(0,u1-u2)
                 ("max", x3, e3 - u2, -1.0)]   # d = max(xi3, eta3-u2)
        pure = [x3, e3]

        def offset(br):
            b, d = br[1], br[2]
            return np.array([d - b, d])
        return D, 2, terms, pure, offset
    if ncv == 2:
        D = 5
        u1, x2, x3, e2, e3 = np.eye(5)
        z = np.zeros(5)
        terms = [("max", z, u1, -1.0),         # a = 1 - max(0,u1)
                 ("max", x2, e2 - u1, -1.0)]   # d = max(xi2, eta2-u1)
        pure = [x2 - x3, x3, e2 - e3, e3]

        def offset(br):
            return np.array([br[1]])
        return D, 1, terms, pure, offset
    raise ValueError(ncv)


def enumerate_regions(ncv):
    D, m, terms, pure, offset = case_definition(ncv)
    patterns = []
    simplices = []
    for choice in itertools.product((0, 1), repeat=len(terms)):
        ineq = [g for g in pure]
        branches = []
        ell = np.zeros(D)
        for (kind, f0, f1, sign), k in zip(terms, choice):
            fk, fo = (f0, f1) if k == 0 else (f1, f0)
            ineq.append(fk - fo if kind == "max" else fo - fk)
            branches.append(fk)
            ell -= sign * fk           # sigma = 1 - ell
        G = np.array(ineq)
        # vertex enumeration of {G s >= 0, ell.s = 1}
        verts = []
        for combo in itertools.combinations(range(len(G)), D - 1):
            Msys = np.vstack([ell[None, :], G[list(combo)]])
            if abs(np.linalg.det(Msys)) < 1e-12:
                continue
            rhs = np.zeros(D)
            rhs[0] = 1.0
            v = np.linalg.solve(Msys, rhs)
            if np.all(G @ v >= -1e-12):
                if not any(np.allclose(v, w, atol=1e-10) for w in verts):
                    verts.append(v)
        if len(verts) < D:
            continue
        verts = np.array(verts)
        # coordinates inside the hyperplane ell.s = 1
        _, _, vt = np.linalg.svd(ell[None, :])
        basis = vt[1:]                       # (D-1) x D, orthonormal, ell-orthogonal
        proj = (verts - verts[0]) @ basis.T
        if np.linalg.matrix_rank(proj, tol=1e-10) < D - 1:
            continue                         # measure-zero cone
        O = offset(branches)
        pid = len(patterns)
        patterns.append(O)
        if len(verts) == D:
            tri = [list(range(D))]
        else:
            tri = Delaunay(proj).simplices.tolist()
        for simp in tri:
            sv = verts[simp]
            cols = [sv[0]] + [sv[k] - sv[k - 1] for k in range(1, D)]
            det = abs(np.linalg.det(np.array(cols).T))
            if det < 1e-13:
                continue
            # l(s) must be positive on the cone: check on the base vertices
            assert np.all(np.abs(sv @ ell - 1.0) < 1e-10)
            simplices.append((pid, sv, det))
    return D, m, patterns, simplices


def fmt(x):
    r = repr(float(x))
    if r in ("-0.0",):
        r = "0.0"
    return r


def main(out_path):
    lines = []
    lines.append("// Generated by gen_td_regions.py -- Taylor-Duffy pyramid decomposition.")
    lines.append("// Do not edit by hand; rerun the generator instead.")
    lines.append("#pragma once")
    lines.append("// TD_QUAL lets device code place the tables in constant memory.")
    lines.append("#ifndef TD_QUAL")
    lines.append("#define TD_QUAL")
    lines.append("#endif")
    lines.append("")
    summary = []
    for ncv in (4, 3, 2):
        D, m, patterns, simplices = enumerate_regions(ncv)
        tot = sum(s[2] for s in simplices)
        summary.append((ncv, len(patterns), len(simplices), tot))
        lines.append(f"// ncv={ncv}: D={D} singular dims, m={m} remaining dims, "
                     f"{len(patterns)} patterns, {len(simplices)} base simplices")
        lines.append(f"#define TD{ncv}_D {D}")
        lines.append(f"#define TD{ncv}_M {m}")
        lines.append(f"#define TD{ncv}_NPAT {len(patterns)}")
        lines.append(f"#define TD{ncv}_NSIMP {len(simplices)}")
        lines.append(f"TD_QUAL static const double TD{ncv}_OFFSET[{len(patterns)}][{m}][{D}] = {{")
        for O in patterns:
            rows = ", ".join("{" + ", ".join(fmt(x) for x in row) + "}" for row in O)
            lines.append(f"  {{{rows}}},")
        lines.append("};")
        lines.append(f"TD_QUAL static const int TD{ncv}_SIMP_PAT[{len(simplices)}] = {{"
                     + ", ".join(str(s[0]) for s in simplices) + "};")
        lines.append(f"TD_QUAL static const double TD{ncv}_SIMP_DET[{len(simplices)}] = {{"
                     + ", ".join(fmt(s[2]) for s in simplices) + "};")
        lines.append(f"TD_QUAL static const double TD{ncv}_SIMP_V[{len(simplices)}][{D}][{D}] = {{")
        for _, sv, _ in simplices:
            rows = ", ".join("{" + ", ".join(fmt(x) for x in row) + "}" for row in sv)
            lines.append(f"  {{{rows}}},")
        lines.append("};")
        lines.append("")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    for ncv, npat, nsimp, tot in summary:
        print(f"ncv={ncv}: patterns={npat} simplices={nsimp} sum|det|={tot:.6f}")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "td_regions.h"))
