"""The source-compatible C++ drop-in (buff-em_amd/host/libbuff.h).

Two programs written against it with the reference's API -- SWGGeometry
(GeoFile), AllocateVIEMatrix / AssembleVIEMatrix, AllocateRHSVector /
AssembleRHSVector(Omega, IncField*, HVector*), HMatrix::LUFactorize /
LUSolve, GetFields(IncField*, HVector*, Omega, HMatrix*) returning an
HMatrix, GetPFTMatrix(HVector*, Omega, PFTOptions*) returning NumObjects x
NUMPFT, PlaneWave(E0, nHat) chained through IncField::Next, HMatrix text
files, GetEntries / GetEntriesD / Copy, ErrExit / Log -- are built by
buff-em_amd/build.py (host/examples/*.cc):

- lffield_dropin: the calls of src/unitTests/unit-test-LFField.cc:50-69 on
  the reference's own inputs (E10Sphere_533.buffgeo, EPFile.XAxis,
  Omega = 0.1).  The incident field equals E0 e^{i Omega z}; the scattered
  field equals the oracle's GetFields restatement applied to the oracle's
  solution.
- scatter_dropin: buff-scatter's loop (buff-scatter.cc:182-314,
  OutputModules.cc:149-211) with two plane waves chained into one incident
  field at two frequencies, on the Sphere_533 mesh with eps = 10+1i; the
  PFT file's absorbed and scattered power equal the oracle's for the summed
  field.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
LIBDIR = os.path.join(ROOT, "buff-em_amd", "lib")


def _exe(name):
    p = os.path.join(LIBDIR, name)
    assert os.path.exists(p), f"{p} missing: run python buff-em_amd/build.py"
    return p


def test_dropin_programs_built():
    for name in ("lffield_dropin", "scatter_dropin"):
        r = subprocess.run([_exe(name)], capture_output=True, text=True)
        assert r.returncode == 2 and "usage" in r.stderr


def _golden_dir(tmp_path):
    for f in ("E10Sphere_533.buffgeo", "Sphere_533.vmsh", "EPFile.XAxis"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    return tmp_path


def _oracle_solution(bfa, orc, omega, eps, pws):
    V, T = bfa.read_vmsh(os.path.join(GOLDEN, "Sphere_533.vmsh"))
    o = orc.Geometry(V, T)
    M = o.assemble_colmajor(omega, orc.material("iso", eps=eps), o.static_table(12))
    B = o.rhs(omega, orc.planewaves(pws)).sum(axis=1, keepdims=True)
    J = orc.lu_solve(M, B, overwrite=True)
    return o, V, T, J[:, 0]


@pytest.mark.gpu
def test_lffield_program(bfa, orc, torch_cuda, tmp_path):
    d = _golden_dir(tmp_path)
    out = d / "lffield.out"
    r = subprocess.run([_exe("lffield_dropin"), "E10Sphere_533.buffgeo", "EPFile.XAxis", "0.1", str(out)],
                       cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stderr[-2000:])
    assert r.returncode == 0
    D = np.loadtxt(out)
    assert D.shape == (82, 3 + 12 + 12)
    X = D[:, :3]
    FS = D[:, 3:15:2] + 1j * D[:, 4:15:2]
    FI = D[:, 15::2] + 1j * D[:, 16::2]
    assert np.allclose(X, np.loadtxt(os.path.join(GOLDEN, "EPFile.XAxis")))
    # incident: E = x e^{i 0.1 z}, H = z x E / Z0
    Ei = np.exp(1j * 0.1 * X[:, 2])
    want = np.zeros((82, 6), complex)
    want[:, 0] = Ei
    want[:, 4] = Ei / orc.ZVAC
    assert np.abs(FI - want).max() < 1e-14
    # scattered: the oracle's fields of the oracle's solution
    o, V, T, J = _oracle_solution(bfa, orc, 0.1, 10.0 + 0j, [((1, 0, 0), (0, 0, 1))])
    ref = orc.fields(o, V, T, J, 0.1, X)
    err = np.linalg.norm(FS - ref) / np.linalg.norm(ref)
    print(f"[dropin] LFField scattered field vs oracle: {err:.2e}; E_x at the centre "
          f"{(FS[20, 0] + FI[20, 0]).real:.4f}")
    assert err < 1e-9


@pytest.mark.gpu
def test_scatter_program(bfa, orc, torch_cuda, tmp_path):
    d = _golden_dir(tmp_path)
    # the reference's sphere with a lossy material, so that both powers are O(1)
    with open(d / "LossySphere.buffgeo", "w") as f:
        f.write("OBJECT TheSphere\n\tMESHFILE Sphere_533.vmsh\n\tMATERIAL CONST_EPS_10+1i\nENDOBJECT\n")
    pws = [((1, 0, 0), (0, 0, 1)), ((0, 0.6, 0.8), (1, 0, 0))]
    omegas = [0.1, 1.0]
    args = [_exe("scatter_dropin"), "LossySphere.buffgeo", "out.PFT"]
    for w in omegas:
        args += ["--omega", str(w)]
    for e0, n in pws:
        args += ["--pw", *map(str, e0), *map(str, n)]
    r = subprocess.run(args, cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stderr[-2000:])
    assert r.returncode == 0
    rows = [l.split() for l in open(d / "out.PFT") if l.strip()]
    assert len(rows) == len(omegas)
    for row, w in zip(rows, omegas):
        assert float(row[0]) == pytest.approx(w) and row[1] == "TheSphere"
        vals = [float(x) for x in row[2:]]
        assert len(vals) == 8 and all(np.isnan(vals[2:]))
        o, V, T, J = _oracle_solution(bfa, orc, w, 10.0 + 1j, pws)
        # the chain's power: P_scat from J alone, P_ext summed over the waves
        pext, pscat = 0.0, None
        for pw in pws:
            p = o.emt_power(w, orc.planewaves([pw]), J[:, None])[0]
            pext += p[0] + p[1]
            pscat = p[1]
        want = np.array([pext - pscat, pscat])
        rel = np.abs(np.array(vals[:2]) - want) / np.abs(want)
        print(f"[dropin] omega={w}: PFT {vals[:2]} oracle {want.tolist()} rel {rel.max():.1e}")
        assert rel.max() < 1e-9
