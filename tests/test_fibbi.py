"""FIBBI cache files: the reference's on-disk format for the static records.

Format (FIBBICache.cc:247-463): b"FIBBI_GCACHE\\0", then records of a
27-float key (GetFIBBICacheKey, FIBBICache.cc:171-203) and 6 doubles.

CPU: keys against the oracle's own face tables, file round trip, the
reference's validation rules (signature, record-size multiple).
GPU: store -> preload gives back the same table and matrix without the cold
step; a cache written from the oracle's static table (an independent
producer of the same format) is consumed exactly; a partial cache mixes
file records (hits) and device-computed ones (misses); the C++ drop-in
driver writes <mesh>.cache into BUFF_CACHE_PATH and preloads it on the next
run, as buff-scatter does.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, kuhn

REC = np.dtype([("key", "<f4", 27), ("data", "<f8", 6)])
SIG = b"FIBBI_GCACHE\x00"


def _ref_keys(o, V):
    """Keys from the oracle's face tables (numpy: double difference, then float)."""
    F = o.faces()
    rowptr, cols = o.touch_pairs()
    keys = []
    for a in range(len(rowptr) - 1):
        for j in range(rowptr[a], rowptr[a + 1]):
            b = cols[j]
            qpa = V[F["QP"][a]]
            pts = [V[F["QM"][a]], *V[F["V"][a]], V[F["QP"][b]], V[F["QM"][b]], *V[F["V"][b]]]
            keys.append(np.concatenate([(p - qpa).astype(np.float32) for p in pts]))
    return np.array(keys, dtype=np.float32)


def test_keys_match_oracle_face_tables(bfa, orc):
    V, T = kuhn(bfa, 2)
    g = bfa.Geometry(V, T)
    o = orc.Geometry(V, T)
    k = g.fibbi_keys()
    ref = _ref_keys(o, V)
    assert k.shape == ref.shape == (g.touch_pairs()[1].size, 27)
    assert np.array_equal(k.view(np.uint32), ref.view(np.uint32))
    # self pair: QP_B - QP_A = 0 and the B block repeats the A block
    rowptr, cols = g.touch_pairs()
    d = rowptr[:-1]
    assert np.all(cols[d] == np.arange(g.N))
    assert np.all(k[d, 12:15] == 0) and np.array_equal(k[d, 15:18], k[d, 0:3])
    assert np.array_equal(k[d, 18:27], k[d, 3:12])


def test_write_read_round_trip_and_layout(bfa, tmp_path):
    rng = np.random.default_rng(1)
    keys = rng.standard_normal((37, 27)).astype(np.float32)
    data = rng.standard_normal((37, 6))
    p = tmp_path / "x.cache"
    bfa.fibbi_write(p, keys, data)
    raw = p.read_bytes()
    assert raw[:13] == SIG and len(raw) == 13 + 37 * (27 * 4 + 6 * 8)
    recs = np.frombuffer(raw[13:], dtype=REC)
    assert np.array_equal(recs["key"], keys) and np.array_equal(recs["data"], data)
    k2, d2 = bfa.fibbi_read(p)
    assert np.array_equal(k2, keys) and np.array_equal(d2, data)
    bfa.fibbi_write(p, keys[:0], data[:0])
    k3, d3 = bfa.fibbi_read(p)
    assert k3.shape == (0, 27) and d3.shape == (0, 6)


def test_reader_rejects_what_the_reference_rejects(bfa, tmp_path):
    rng = np.random.default_rng(2)
    good = SIG + np.zeros(2, dtype=REC).tobytes()
    cases = {
        "missing.cache": None,
        "short.cache": SIG[:5],
        "badsig.cache": b"FIBBI_XCACHE\x00" + good[13:],
        "ragged.cache": good + b"\x00" * 7,
    }
    for name, content in cases.items():
        p = tmp_path / name
        if content is not None:
            p.write_bytes(content)
        with pytest.raises(bfa.BFAError):
            bfa.fibbi_read(p)
    p = tmp_path / "ok.cache"
    p.write_bytes(good)
    k, d = bfa.fibbi_read(p)
    assert len(k) == 2


@pytest.fixture(scope="module")
def mesh3(bfa, orc, torch_cuda):
    V, T = kuhn(bfa, 3)
    o = orc.Geometry(V, T)
    return V, T, o, o.static_table(12)


def _assemble(bfa, torch, g, w=1.3):
    mat = bfa.material("iso", eps=4 + 0.1j)
    M = torch.empty((g.N, g.N), dtype=torch.complex128, device="cuda:0")
    g.assemble_vie(w, mat, M)
    return M.cpu().numpy()


@pytest.mark.gpu
def test_store_then_preload_skips_the_cold_step(bfa, torch_cuda, mesh3, tmp_path):
    torch = torch_cuda
    V, T, o, _ = mesh3
    g = bfa.Geometry(V, T)
    g.upload()
    g.build_static(10)
    t0 = g.static_table()
    p = tmp_path / "kuhn3.cache"
    n = g.fibbi_store(p)
    keys = g.fibbi_keys()
    assert n == len(np.unique(keys.view(np.uint8).reshape(len(keys), -1), axis=0))
    print(f"[fibbi] pairs {len(keys)} records {n}", flush=True)
    g2 = bfa.Geometry(V, T)
    g2.upload()
    hits, misses = g2.fibbi_preload(p, td_order=10)
    assert (hits, misses) == (len(keys), 0)
    assert np.array_equal(g2.static_table(), t0)
    assert np.array_equal(_assemble(bfa, torch, g2), _assemble(bfa, torch, g))


@pytest.mark.gpu
def test_oracle_written_cache_is_consumed_exactly(bfa, orc, torch_cuda, mesh3, tmp_path):
    """Records from an independent producer (the oracle's TD order 12) replace the
    device's own; the assembled matrix is the oracle's to round-off."""
    torch = torch_cuda
    V, T, o, tab = mesh3
    g = bfa.Geometry(V, T)
    keys = g.fibbi_keys()
    p = tmp_path / "oracle.cache"
    bfa.fibbi_write(p, keys, tab)
    g.upload()
    hits, misses = g.fibbi_preload(p, td_order=10)
    assert (hits, misses) == (len(keys), 0)
    assert np.array_equal(g.static_table(), tab)
    M = _assemble(bfa, torch, g)
    Mo = o.assemble(1.3, orc.material("iso", eps=4 + 0.1j), tab)
    err = np.linalg.norm(M.T - Mo) / np.linalg.norm(Mo)
    print(f"[fibbi] matrix from oracle cache vs oracle: {err:.2e}", flush=True)
    assert err < 1e-12


@pytest.mark.gpu
def test_partial_cache_mixes_hits_and_computed_misses(bfa, orc, torch_cuda, mesh3, tmp_path):
    V, T, o, tab = mesh3
    g = bfa.Geometry(V, T)
    keys = g.fibbi_keys()
    sel = np.arange(0, len(keys), 2)
    p = tmp_path / "half.cache"
    # half of the oracle's records, plus a duplicate key with a different
    # record after the first: the first occurrence wins (unordered_map::insert)
    bad = tab[sel[:1]] * 2.0
    bfa.fibbi_write(p, np.concatenate([keys[sel], keys[sel[:1]]]), np.concatenate([tab[sel], bad]))
    g.upload()
    hits, misses = g.fibbi_preload(p, td_order=10)
    assert hits == len(sel) and misses == len(keys) - len(sel)
    t = g.static_table()
    assert np.array_equal(t[sel], tab[sel])
    ref = bfa.Geometry(V, T)
    ref.upload()
    ref.build_static(10)
    odd = np.setdiff1d(np.arange(len(keys)), sel)
    assert np.array_equal(t[odd], ref.static_table()[odd])


@pytest.mark.gpu
def test_cpp_driver_writes_and_preloads_mesh_cache(bfa, tmp_path):
    drv = os.path.join(ROOT, "buff-em_amd", "lib", "buff_scatter_amd")
    mesh = os.path.join(ROOT, "tests", "golden", "Sphere_533.vmsh")
    env = dict(os.environ, BUFF_CACHE_PATH=str(tmp_path))
    args = [drv, "--vmsh", mesh, "--eps", "4", "0.1", "--omega", "1", "--pw", "1", "0", "0", "0", "0", "1"]
    r1 = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env, cwd=tmp_path)
    assert r1.returncode == 0, r1.stderr
    cache = tmp_path / "Sphere_533.cache"
    assert cache.exists() and cache.read_bytes()[:13] == SIG
    k, d = bfa.fibbi_read(cache)
    print(f"[fibbi] driver cache records {len(k)}", flush=True)
    r2 = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env, cwd=tmp_path)
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout == r1.stdout
    assert "preloaded" in r2.stderr
