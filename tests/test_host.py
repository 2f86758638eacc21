"""Host-side tests of the product library (no GPU compute calls).

The library loads, exports every entry point include/buffem_amd.h declares,
and its host code (mesh construction, GMSH reader, SWG topology, touching
pairs, error handling) agrees bit-exactly with the oracle's restatement of
InitFaceList / CompareBFs (src/libs/libbuff/InitFaceList.cc:69-279,
src/libs/libbuff/SWGVolume.cc:295-344).
"""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, kuhn

GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "buffem_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bfa_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header(bfa):
    L = bfa.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(bfa.declared_symbols()) == set(syms)
    assert L.bfa_version() >= 1


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_kuhn_counts_and_orientation(bfa, n):
    V, T = kuhn(bfa, n)
    assert V.shape == ((n + 1) ** 3, 3) and T.shape == (6 * n ** 3, 4)
    P = V[T]
    vol = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6
    assert (vol > 0).all()
    g = bfa.Geometry(V, T)
    assert g.N == 12 * n ** 3 - 6 * n ** 2
    # boundary lattice nodes lie on the sphere, interior ones inside
    r = np.linalg.norm(V, axis=1)
    ijk = np.array(np.unravel_index(np.arange(len(V)), (n + 1,) * 3, order="F")).T
    bnd = ((ijk == 0) | (ijk == n)).any(1)
    np.testing.assert_allclose(r[bnd], 1.0, atol=1e-14)
    assert (r[~bnd] < 1.0).all()


def test_kuhn_deterministic_and_jitter_bounded(bfa):
    V1, T1 = bfa.kuhn_sphere(4, 1.5, 0.05, 7)
    V2, T2 = bfa.kuhn_sphere(4, 1.5, 0.05, 7)
    V3, _ = bfa.kuhn_sphere(4, 1.5, 0.05, 8)
    V0, _ = bfa.kuhn_sphere(4, 1.5, 0.0, 7)
    assert np.array_equal(V1, V2) and np.array_equal(T1, T2)
    assert not np.array_equal(V1, V3)
    h = 2 * 1.5 / 4
    assert np.abs(V1 - V0).max() <= 0.05 * h + 1e-15
    assert np.abs(V1 - V0).max() > 0.02 * h


def _topology_equal(g, o):
    fg, fo = g.faces(), o.faces()
    for k in ("QP", "QM", "PT", "MT", "PI", "MI", "V"):
        assert np.array_equal(fg[k], fo[k]), k
    assert np.array_equal(fg["Area"], fo["Area"])
    rg, cg = g.touch_pairs()
    ro, co = o.touch_pairs()
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)


@pytest.mark.parametrize("n", [1, 2, 4])
def test_topology_matches_oracle(bfa, orc, n):
    V, T = kuhn(bfa, n)
    _topology_equal(bfa.Geometry(V, T), orc.Geometry(V, T))


def test_touch_pairs_definition(bfa, orc):
    V, T = kuhn(bfa, 2)
    g = bfa.Geometry(V, T)
    o = orc.Geometry(V, T)
    rowptr, cols = g.touch_pairs()
    touch = set()
    for a in range(g.N):
        for j in range(rowptr[a], rowptr[a + 1]):
            touch.add((a, int(cols[j])))
    for a in range(g.N):
        for b in range(a, g.N):
            assert ((a, b) in touch) == (o.compare_bfs(a, b) > 0)


@pytest.mark.parametrize("name,nt", [("Sphere_48.vmsh", None), ("Sphere_533.vmsh", 323), ("Sphere_677.vmsh", 677)])
def test_vmsh_reader(bfa, orc, name, nt):
    V, T = bfa.read_vmsh(os.path.join(GOLDEN, name))
    if nt is not None:
        assert len(T) == nt
    # every vertex used by a tetrahedron is distinct from every other (merge at 1e-6)
    used = np.unique(T)
    d = np.linalg.norm(V[used][:, None] - V[used][None], axis=2) + np.eye(len(used))
    assert d.min() > 1e-6
    _topology_equal(bfa.Geometry(V, T), orc.Geometry(V, T))


def test_vmsh_merges_duplicate_vertex(bfa):
    # Sphere_48.vmsh holds node 20 at (-2.2e-16, 0, 0), a duplicate of node 1
    V, T = bfa.read_vmsh(os.path.join(GOLDEN, "Sphere_48.vmsh"))
    assert len(V) == 20
    assert 19 not in set(T.ravel().tolist())
    assert 0 in set(T.ravel().tolist())


def test_error_codes(bfa, tmp_path):
    L = bfa.lib()
    nv, nt = ctypes.c_int(), ctypes.c_int()
    assert L.bfa_mesh_kuhn_sphere(0, 1.0, 0.05, 1, ctypes.byref(nv), ctypes.byref(nt), None, None) == -1
    assert b"bad arguments" in L.bfa_last_error()
    assert L.bfa_mesh_read_vmsh(str(tmp_path / "nope.vmsh").encode(), ctypes.byref(nv), ctypes.byref(nt),
                                None, None) == -7
    bad = tmp_path / "bad.vmsh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\nx\n")
    assert L.bfa_mesh_read_vmsh(str(bad).encode(), ctypes.byref(nv), ctypes.byref(nt), None, None) == -2
    # three tetrahedra on one face: invalid topology
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 1]], float)
    T = np.array([[0, 1, 2, 3], [0, 2, 1, 4], [0, 1, 2, 5]], np.int32)
    with pytest.raises(bfa.BFAError, match="topology"):
        bfa.Geometry(V, T)
    with pytest.raises(bfa.BFAError):
        bfa.Geometry(V, np.array([[0, 1, 2, 9]], np.int32))
    assert L.bfa_status_string(-3) == b"singular matrix"


def test_cpp_dropin_builds():
    """The C++ drop-in header and the buff-scatter driver compile against the C ABI."""
    import subprocess
    drv = os.path.join(ROOT, "buff-em_amd", "lib", "buff_scatter_amd")
    assert os.path.exists(drv)
    r = subprocess.run([drv], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
