"""Batching form: many instances of uneven size swept over frequencies in
grouped launches (optionally one captured hipGraph per frequency), complex64
LU + FP64 refinement, EMT power per instance and frequency.

Every per-instance result is checked against the single-instance FP64 GPU
path (assemble_vie + lu_factor + lu_solve + emt_power, itself pinned to the
oracle by test_gpu_parity.py) and, for the smallest instance, against the
oracle directly.  Eager and graph runs must agree bit for bit, on repeated
replays with other work in between.
"""
import numpy as np
import pytest

from conftest import kuhn

pytestmark = pytest.mark.gpu

OMEGAS = [0.5, 1.0, 3.0]
SPEC = [((1, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 0, 1))]


def _instances(bfa):
    # (n, radius, seed, material)
    cases = [
        (2, 1.0, 11, bfa.material("iso", eps=4 + 0.1j)),
        (3, 0.8, 12, bfa.material("janus", up=4 + 0.1j, down=12 + 1j)),
        (4, 1.2, 13, bfa.material("iso", eps=4 + 0.1j)),
        (2, 0.6, 14, bfa.material("janus", up=4 + 0.1j, down=12 + 1j)),
        (5, 1.0, 15, bfa.material("diag", eps=(4 + 0.05j, 5 + 0.05j, 6 + 0.05j))),
    ]
    geoms, mats = [], []
    for n, R, seed, m in cases:
        V, T = bfa.kuhn_sphere(n, R, 0.05, seed)
        g = bfa.Geometry(V, T)
        g.upload()
        g.build_static(10)
        geoms.append(g)
        mats.append(m)
    return geoms, mats


def _single_fp64(bfa, torch, g, mat, omega, pws):
    dev = torch.device("cuda:0")
    N = g.N
    M = torch.empty((N, N), dtype=torch.complex128, device=dev)
    B = torch.empty((len(pws), N), dtype=torch.complex128, device=dev)
    piv = torch.empty(N, dtype=torch.int32, device=dev)
    g.assemble_vie(omega, mat, M)
    g.assemble_rhs(omega, pws, B)
    Mh = M.T.cpu().numpy().copy()
    assert bfa.lu_factor(M, N, piv) == 0
    bfa.lu_solve(M, N, piv, B, len(pws))
    P = g.emt_power(omega, pws, B)
    return Mh, B.cpu().numpy().T, P


@pytest.fixture(scope="module")
def batch_setup(bfa, torch_cuda):
    geoms, mats = _instances(bfa)
    pws = bfa.planewaves(SPEC)
    b = bfa.Batch(geoms, mats, pws)
    return geoms, mats, pws, b


def test_batch_matches_single_fp64(bfa, torch_cuda, batch_setup):
    torch = torch_cuda
    geoms, mats, pws, b = batch_setup
    print(f"[batch] N={b.N.tolist()} Np={b.Np.tolist()}", flush=True)
    r = b.run(OMEGAS, use_graph=False)
    assert (r["info"] == 0).all()
    assert (r["resid"] <= 1e-10).all(), r["resid"]
    assert (r["iters"] >= 1).all()
    for f, w in enumerate(OMEGAS):
        for i, (g, m) in enumerate(zip(geoms, mats)):
            Mh, J, P = _single_fp64(bfa, torch, g, m, w, pws)
            rel = np.abs(r["power"][f, i] - P) / np.abs(P)
            assert rel.max() < 1e-9, (f, i, rel, r["power"][f, i], P)
            if f == len(OMEGAS) - 1:
                # the batch's complex128 matrix is the single path's: same rule and
                # summation order; the two kernels' code may contract multiply-adds
                # differently, so entries agree to the last bits
                Mb = b.matrix(i)
                d = np.abs(Mb - Mh).max()
                assert d <= 1e-14 * np.abs(Mh).max(), d
                Jb = b.solution(i)
                assert np.linalg.norm(Jb - J) / np.linalg.norm(J) < 1e-9


def test_batch_graph_replays_match_eager(bfa, torch_cuda, batch_setup):
    torch = torch_cuda
    geoms, mats, pws, b = batch_setup
    ref = b.run(OMEGAS, use_graph=False)
    for rep in range(3):
        print(f"[batch] graph replay {rep}", flush=True)
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda:0")
        (x * 2).sum().item()
        r = b.run(OMEGAS, use_graph=True)
        assert np.array_equal(r["power"], ref["power"]), np.abs(r["power"] - ref["power"]).max()
        assert np.array_equal(r["iters"], ref["iters"])


def test_batch_smallest_against_oracle(bfa, orc, torch_cuda, batch_setup):
    geoms, mats, pws, b = batch_setup
    r = b.run([1.0], use_graph=True)
    g = geoms[0]
    o = orc.Geometry(g.V, g.T)
    tab = o.static_table(10)
    M = o.assemble(1.0, orc.material("iso", eps=4 + 0.1j), tab)
    opw = orc.planewaves(SPEC)
    J = orc.lu_solve(M, o.rhs(1.0, opw))
    P = o.emt_power(1.0, opw, J)
    rel = np.abs(r["power"][0, 0] - P) / np.abs(P)
    assert rel.max() < 1e-9, rel


def test_batch_autotuner(bfa, torch_cuda, batch_setup):
    geoms, mats, pws, b = batch_setup
    before = b.run(OMEGAS, use_graph=True)
    tile, ms = b.tune(reps=2)
    print(f"[batch] tuned tiles {tile.tolist()}", flush=True)
    used = tile >= 0
    assert used.any()
    assert set(tile[used].tolist()) <= {0, 1, 2}
    # the chosen variant is the faster one measured for its bucket
    for bk in np.nonzero(used)[0]:
        assert ms[bk, tile[bk]] == ms[bk].min()
    after = b.run(OMEGAS, use_graph=True)
    rel = np.abs(after["power"] - before["power"]) / np.abs(before["power"])
    assert rel.max() < 1e-9
    # every variant forced on every bucket: the refined solutions agree (the
    # split-FP16 trailing update changes the complex64 factors, and the FP64
    # refinement restores the same powers)
    for t in (0, 1, 2):
        b.set_tiles(t)
        r = b.run(OMEGAS, use_graph=True)
        rel = np.abs(r["power"] - before["power"]) / np.abs(before["power"])
        print(f"[batch] all buckets on variant {t}: power vs fixed rule {rel.max():.2e}, refinement iterations "
              f"max {int(r['iters'].max())}", flush=True)
        assert (r["info"] == 0).all() and (r["resid"] <= 1e-10).all()
        assert rel.max() < 1e-9
    b.set_tiles(-1)


_SINGLE_PASS_BATCH = r"""
import os, sys
import numpy as np
import torch
sys.path.insert(0, os.path.join(sys.argv[1], "tests"))
from conftest import load_bfa
from test_batch import _instances, SPEC
bfa = load_bfa()
geoms, mats = _instances(bfa)
b = bfa.Batch(geoms, mats, bfa.planewaves(SPEC))
r = b.run([1.0], use_graph=True)
np.savez(sys.argv[2], power=r["power"], **{f"m{i}": b.matrix(i) for i in range(len(geoms))})
"""


def test_batch_single_pass_far_field(bfa, torch_cuda, batch_setup, tmp_path):
    """The batch's far field without the moment tables (BFA_FAR_TABLE=0, in a
    child process: the switch is read once per process) against the two-pass
    one: the same sums in another order, matrices to 1e-14 relative, powers
    to 1e-12."""
    import os
    import subprocess
    import sys
    geoms, mats, pws, b = batch_setup
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = str(tmp_path / "single_pass.npz")
    r = subprocess.run([sys.executable, "-c", _SINGLE_PASS_BATCH, root, out], capture_output=True, text=True,
                       env=dict(os.environ, BFA_FAR_TABLE="0"), timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    single = np.load(out)
    two = b.run([1.0], use_graph=True)
    for i in range(len(geoms)):
        A, B = b.matrix(i), single[f"m{i}"]
        err = np.linalg.norm(A - B) / np.linalg.norm(B)
        print(f"[batch] instance {i}: two-pass vs single-pass far field {err:.2e}", flush=True)
        assert err < 1e-14
    rel = np.abs(two["power"] - single["power"]) / np.abs(single["power"])
    assert rel.max() < 1e-12, rel.max()
