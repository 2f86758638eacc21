"""Material and geometry front end: SVTensor / MatProp expressions and .buffgeo.

CPU: the .buffgeo loader (golden E10Sphere_533.buffgeo, DISPLACED/ROTATED
against scipy's rotations, the reference's error cases) and expression
parsing errors, without a device.
GPU: device evaluation of compiled expressions against Python's cmath on
the same inputs; the SVTensor tensor rules (defaults, Qxx copied, symmetric
off-diagonals, MP_ references, MATERIAL sections, constants); assembly with
compiled materials equal to the closed-form kinds the oracle pins; rotation
invariance of the full-tensor Gram path; the golden .buffgeo end to end and
through the C++ driver.
"""
import cmath
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import ROOT, kuhn

GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_buffgeo_golden_file(bfa):
    V, T = bfa.read_vmsh(os.path.join(GOLDEN, "Sphere_533.vmsh"))
    L = bfa.lib()
    import ctypes
    nv, nt = ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(os.path.join(GOLDEN, "E10Sphere_533.buffgeo"))
    assert L.bfa_buffgeo_load(p, ctypes.byref(nv), ctypes.byref(nt), None, None, None) == 0
    assert (nv.value, nt.value) == V.shape[:1] + T.shape[:1]
    V2 = np.zeros_like(V)
    T2 = np.zeros_like(T)
    assert L.bfa_buffgeo_load(p, ctypes.byref(nv), ctypes.byref(nt), V2.ctypes.data, T2.ctypes.data, None) == 0
    assert np.array_equal(V2, V) and np.array_equal(T2, T)
    bfa.material_check("CONST_EPS_10", matprop=True)


def _load_xform(bfa, path):
    import ctypes
    L = bfa.lib()
    nv, nt = ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(str(path))
    rc = L.bfa_buffgeo_load(p, ctypes.byref(nv), ctypes.byref(nt), None, None, None)
    if rc:
        raise bfa.BFAError(rc, L.bfa_last_error().decode())
    V = np.zeros((nv.value, 3))
    T = np.zeros((nt.value, 4), dtype=np.int32)
    assert L.bfa_buffgeo_load(p, ctypes.byref(nv), ctypes.byref(nt), V.ctypes.data, T.ctypes.data, None) == 0
    return V


def test_buffgeo_transformations(bfa, tmp_path):
    V0, _ = bfa.read_vmsh(os.path.join(GOLDEN, "Sphere_533.vmsh"))
    mesh = os.path.join(GOLDEN, "Sphere_533.vmsh")
    g = tmp_path / "t.buffgeo"
    g.write_text(f"# comment\nOBJECT S\n MESHFILE {mesh}\n MATERIAL CONST_EPS_4+0.1i\n"
                 " DISPLACED 1 -2 0.5\n ROTATED 30 ABOUT 0 0 1\n ROTATED -45 ABOUT 1 1 0 DISPLACED 0 0 3\n"
                 "ENDOBJECT\n")
    V = _load_xform(bfa, g)
    R1 = Rotation.from_rotvec(np.radians(30) * np.array([0, 0, 1.0]))
    R2 = Rotation.from_rotvec(np.radians(-45) * np.array([1, 1, 0.0]) / math.sqrt(2))
    ref = R2.apply(R1.apply(V0 + [1, -2, 0.5])) + [0, 0, 3]
    np.testing.assert_allclose(V, ref, atol=1e-13)
    # relative MESHFILE found next to the .buffgeo, MESHPATH, errors
    d = tmp_path / "sub"
    d.mkdir()
    (d / "m.vmsh").write_bytes(open(mesh, "rb").read())
    (tmp_path / "rel.buffgeo").write_text("MESHPATH sub\nOBJECT A\nMESHFILE m.vmsh\nSVTENSOR x.SVTensor\nENDOBJECT\n")
    assert np.array_equal(_load_xform(bfa, tmp_path / "rel.buffgeo"), V0)
    bad = {
        "two.buffgeo": f"OBJECT A\nMESHFILE {mesh}\nMATERIAL CONST_EPS_2\nENDOBJECT\nOBJECT B\nENDOBJECT\n",
        "nomesh.buffgeo": "OBJECT A\nMATERIAL CONST_EPS_2\nENDOBJECT\n",
        "nomat.buffgeo": f"OBJECT A\nMESHFILE {mesh}\nENDOBJECT\n",
        "junk.buffgeo": f"OBJECT A\nMESHFILE {mesh}\nMATERIAL CONST_EPS_2\nDISPLACED 1 2 3 4\nENDOBJECT\n",
        "kw.buffgeo": f"OBJECT A\nMESHFILE {mesh}\nCOLOR red\nENDOBJECT\n",
        "missing.buffgeo": "OBJECT A\nMESHFILE nowhere.vmsh\nMATERIAL CONST_EPS_2\nENDOBJECT\n",
    }
    for name, text in bad.items():
        (tmp_path / name).write_text(text)
        with pytest.raises(bfa.BFAError):
            _load_xform(bfa, tmp_path / name)


def test_expression_and_svtensor_parse_errors(bfa, tmp_path):
    for ok in ("4+0.1*i", "-x^2", "2^-1", "sin(pi*x)*cos(pi*y) + 1e-3*I", "exp(-(r/0.3)^2)*Theta+Phi", "w/3e14"):
        bfa.material_check(expr=ok)
    deep = "1" + "+(2*(3" * 9 + ")" * 18      # 19 values live at once: beyond the 16-slot stack
    for bad in ("4+", "sin x", "foo(x)", "q+1", "(1+2", "1+2)", "3 4", deep):
        with pytest.raises(bfa.BFAError):
            bfa.material_check(expr=bad)
    f = tmp_path / "bad.SVTensor"
    f.write_text("EpsQZ = 4\n")
    with pytest.raises(bfa.BFAError):
        bfa.material_check(str(f))
    f.write_text("Eps = MP_Unobtainium\n")
    with pytest.raises(bfa.BFAError):
        bfa.material_check(str(f))
    f.write_text("A = notanumber\n")
    with pytest.raises(bfa.BFAError):
        bfa.material_check(str(f))
    with pytest.raises(bfa.BFAError):
        bfa.material_check("nonexistent.SVTensor")
    for bad in ("VACUUM", "CONST_EPS_", "CONST_EPS_4+i"):
        with pytest.raises(bfa.BFAError):
            bfa.material_check(bad, matprop=True)


# ---- device ----------------------------------------------------------------

def _py_eval(expr, v):
    """The same expression evaluated by Python's cmath (independent evaluator)."""
    env = {k: complex(val) for k, val in v.items()}
    env.update(pi=math.pi, e=math.e, i=1j, I=1j)
    fns = {n: getattr(cmath, n) for n in ("exp", "log", "sqrt", "sin", "cos", "tan", "asin", "acos", "atan",
                                          "sinh", "cosh", "tanh")}
    fns.update(cot=lambda z: 1 / cmath.tan(z), sec=lambda z: 1 / cmath.cos(z), csc=lambda z: 1 / cmath.sin(z),
               abs=lambda z: complex(abs(z)), step=lambda z: complex(1.0 if z.real >= 0 else 0.0))
    env.update(fns)
    return eval(expr.replace("^", "**"), {"__builtins__": {}}, env)


EXPRS = [
    "4 + 2*sin(pi*x)*cos(pi*y) + 0.1*i",
    "exp(-(x^2+y^2)) + sqrt(r)*i - log(2+z)",
    "tan(0.3*x) + cot(1+y) + sec(z) - csc(2+x) + atan(y)*i",
    "asin(0.4*x) + acos(0.3*y) + sinh(z) - cosh(x)*tanh(y)",
    "abs(x - 2*i) + step(z) + step(-y)*3 - Theta/Phi",
    "(w/3e14)^2 - 1/(w/3e14 + 0.2*i) + x^3 - 2^-2 + (1+i)^0.5",
    "-x^2 + 2^3^2/100 - -y",
]


@pytest.mark.gpu
def test_device_expressions_match_cmath(bfa, torch_cuda):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.2, 1.2, (64, 3))
    omega = 0.7
    for ex in EXPRS:
        m = bfa.material_expr(ex)
        Q = bfa.material_eval(m, omega, pts)
        for k, x in enumerate(pts):
            rho = math.hypot(x[0], x[1])
            v = dict(w=omega * 3e14, x=x[0], y=x[1], z=x[2], r=math.sqrt(rho ** 2 + x[2] ** 2),
                     Theta=math.atan2(rho, x[2]), Phi=math.atan2(x[1], x[0]))
            ref = _py_eval(ex, v)
            for d in range(3):
                assert abs(Q[k, d, d] - ref) <= 1e-13 * max(1.0, abs(ref)), (ex, x, Q[k, d, d], ref)
            assert np.all(Q[k][~np.eye(3, dtype=bool)] == 0)


@pytest.mark.gpu
def test_svtensor_rules(bfa, torch_cuda, tmp_path):
    f = tmp_path / "aniso.SVTensor"
    f.write_text("""# anisotropic, inhomogeneous
MATERIAL Lossy
  g = 0.25
  Eps(w) = 3 + g*i
ENDMATERIAL
A = 0.5
EpsXX = MP_Lossy + A*x;
EpsZZ = 6 + 0.05*i
EpsXY = 0.1*y
EpsYX = 0.3
Q31 = 0.2*i
EpsYZ = MP_CONST_EPS_2-0.5i * z
""")
    m = bfa.material_svtensor(str(f))
    pts = np.array([[0.1, -0.2, 0.3], [0.7, 0.4, -0.5]])
    Q = bfa.material_eval(m, 1.0, pts)
    for k, (x, y, z) in enumerate(pts):
        q = np.eye(3, dtype=complex)
        q[0, 0] = 3 + 0.25j + 0.5 * x
        q[2, 2] = 6 + 0.05j
        q[0, 1], q[1, 0] = 0.1 * y, 0.3
        q[2, 0] = 0.2j
        q[1, 2] = (2 - 0.5j) * z
        for mu in range(3):                      # SVTensor::Evaluate's symmetrisation
            nu = (mu + 1) % 3
            a = 0.5 * (q[mu, nu] + q[nu, mu])
            q[mu, nu] = q[nu, mu] = a
        q[1, 1] = q[0, 0]                        # EpsYY absent: copied from EpsXX
        np.testing.assert_allclose(Q[k], q, rtol=0, atol=1e-14)
    # homogeneous MatProp (MATERIAL keyword): eps on the diagonal only
    Qm = bfa.material_eval(bfa.material_svtensor("CONST_EPS_10", matprop=True), 1.0, pts)
    assert np.array_equal(Qm, np.broadcast_to(np.eye(3) * 10.0, Qm.shape))


def _matrix(bfa, torch, g, mat, w=1.1):
    M = torch.empty((g.N, g.N), dtype=torch.complex128, device="cuda:0")
    g.assemble_vie(w, mat, M)
    return M.cpu().numpy()


@pytest.fixture(scope="module")
def geo3(bfa, torch_cuda):
    V, T = kuhn(bfa, 3)
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    return V, T, g


@pytest.mark.gpu
def test_compiled_materials_reproduce_closed_forms(bfa, torch_cuda, geo3, tmp_path):
    torch = torch_cuda
    V, T, g = geo3
    cases = [
        (bfa.material("sincos", base=4.0, amp=2.0, imag=0.1), bfa.material_expr("4 + 2*sin(pi*x)*cos(pi*y) + 0.1*i")),
        (bfa.material("iso", eps=4 + 0.1j), bfa.material_svtensor("CONST_EPS_4+0.1i", matprop=True)),
    ]
    f = tmp_path / "diag.SVTensor"
    f.write_text("EpsXX = 4 + 0.05*i\nEpsYY = 5 + 0.05*i\nEpsZZ = 6 + 0.05*i\n")
    cases.append((bfa.material("diag", eps=(4 + 0.05j, 5 + 0.05j, 6 + 0.05j)), bfa.material_svtensor(str(f))))
    for closed, compiled in cases:
        A, B = _matrix(bfa, torch, g, closed), _matrix(bfa, torch, g, compiled)
        err = np.abs(A - B).max() / np.abs(A).max()
        print(f"[material] compiled vs closed form: {err:.2e}", flush=True)
        assert err < 1e-14


@pytest.mark.gpu
def test_full_tensor_path_is_rotation_invariant(bfa, torch_cuda, tmp_path):
    """Mesh rotated by R with eps' = R eps R^T gives the same VIE matrix: pins
    the off-diagonal tensor evaluation and the full 3x3 chi inverse."""
    torch = torch_cuda
    V, T = kuhn(bfa, 3)
    R = Rotation.from_rotvec([0.4, -0.7, 0.3]).as_matrix()
    eps = np.diag([4 + 0.05j, 5 + 0.05j, 6 + 0.05j])
    ep = R @ eps @ R.T
    lines = []
    for a in range(3):
        for b in range(3):
            lines.append(f"Eps{'XYZ'[a]}{'XYZ'[b]} = ({float(ep[a, b].real):.17g}) + ({float(ep[a, b].imag):.17g})*i")
    f = tmp_path / "rot.SVTensor"
    f.write_text("\n".join(lines) + "\n")
    g0 = bfa.Geometry(V, T)
    g0.upload()
    g0.build_static(10)
    g1 = bfa.Geometry(V @ R.T, T)
    g1.upload()
    g1.build_static(10)
    A = _matrix(bfa, torch, g0, bfa.material("diag", eps=(4 + 0.05j, 5 + 0.05j, 6 + 0.05j)))
    B = _matrix(bfa, torch, g1, bfa.material_svtensor(str(f)))
    err = np.linalg.norm(A - B) / np.linalg.norm(A)
    print(f"[material] rotated mesh + rotated tensor vs original: {err:.2e}", flush=True)
    assert err < 1e-11


@pytest.mark.gpu
def test_golden_buffgeo_end_to_end(bfa, torch_cuda):
    torch = torch_cuda
    V, T, mat = bfa.buffgeo_load(os.path.join(GOLDEN, "E10Sphere_533.buffgeo"))
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    A = _matrix(bfa, torch, g, mat, w=0.8)
    B = _matrix(bfa, torch, g, bfa.material("iso", eps=10.0), w=0.8)
    assert np.abs(A - B).max() <= 1e-15 * np.abs(B).max()
    drv = os.path.join(ROOT, "buff-em_amd", "lib", "buff_scatter_amd")
    common = ["--omega", "0.8", "--pw", "1", "0", "0", "0", "0", "1"]
    r1 = subprocess.run([drv, "--geometry", os.path.join(GOLDEN, "E10Sphere_533.buffgeo")] + common,
                        capture_output=True, text=True, timeout=120, cwd=os.environ.get("TMPDIR", "/tmp"))
    assert r1.returncode == 0, r1.stderr
    r2 = subprocess.run([drv, "--vmsh", os.path.join(GOLDEN, "Sphere_533.vmsh"), "--eps", "10", "0"] + common,
                        capture_output=True, text=True, timeout=120, cwd=os.environ.get("TMPDIR", "/tmp"))
    assert r2.returncode == 0, r2.stderr
    p1 = np.array(r1.stdout.split()[2:4], dtype=float)
    p2 = np.array(r2.stdout.split()[2:4], dtype=float)
    print(f"[material] driver .buffgeo vs .vmsh + eps: {p1} {p2}", flush=True)
    assert np.all(np.abs(p1 - p2) <= 1e-12 * np.abs(p2))
