"""The batching form on its full workload: 48 Kuhn spheres of uneven size
(n ~ U{4..10}, N = 672..11,400, R ~ U(0.5, 1.5) um, every second one Janus:
z > 0 eps = 4+0.1i, z < 0 eps = 12+1i), 8 log-spaced omega in [0.1, 10], one
x-polarised plane wave along +z -- the instances bench.py --workload batch
times, from the same generator.

- Every instance at every frequency: the batch's power (grouped launches
  replayed from a hipGraph, complex64 LU + FP64 refinement) against the
  plain FP64 GPU path run on that instance alone (assemble_vie + lu_factor
  + lu_solve + emt_power), to 1e-9.
- The smallest instance and the first one with n >= 8, at the sweep's two
  end frequencies: against the CPU oracle end to end (its own static table,
  AssembleVIEMatrix, LAPACK, GetEMTPFT), to 1e-9.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_config4_batch_48_instances(bfa, orc, torch_cuda):
    import bench
    torch = torch_cuda
    t0 = time.perf_counter()
    geoms, mats, desc = bench.batch_workload(bfa, 48, 10)
    omegas = bench.BATCH_OMEGAS
    spec = [((1, 0, 0), (0, 0, 1))]
    pw = bfa.planewaves(spec)
    Ns = [d["N"] for d in desc]
    print(f"[config4] 48 instances, N {min(Ns)}..{max(Ns)}, n {sorted(set(d['n'] for d in desc))}", flush=True)
    assert len(geoms) == 48 and min(d["n"] for d in desc) >= 4 and max(d["n"] for d in desc) <= 10
    b = bfa.Batch(geoms, mats, pw)
    b.tune(reps=2)
    r = b.run(omegas, use_graph=True)
    print(f"[config4] sweep done at {time.perf_counter() - t0:.0f}s: refine iters max {int(r['iters'].max())}, "
          f"resid max {r['resid'].max():.2e}", flush=True)
    assert (r["info"] == 0).all()
    assert (r["resid"] <= 1e-10).all(), r["resid"].max()

    dev = torch.device("cuda:0")
    worst = 0.0
    for f, w in enumerate(omegas):
        for i, (g, m) in enumerate(zip(geoms, mats)):
            N = g.N
            dM = torch.empty((N, N), dtype=torch.complex128, device=dev)
            dB = torch.empty((1, N), dtype=torch.complex128, device=dev)
            dp = torch.empty(N, dtype=torch.int32, device=dev)
            g.assemble_vie(float(w), m, dM)
            g.assemble_rhs(float(w), pw, dB)
            assert bfa.lu_factor(dM, N, dp) == 0
            bfa.lu_solve(dM, N, dp, dB, 1)
            P = g.emt_power(float(w), pw, dB)
            rel = float(np.max(np.abs(r["power"][f, i] - P) / np.abs(P)))
            worst = max(worst, rel)
            assert rel < 1e-9, (f, i, desc[i], rel)
            del dM, dB, dp
        print(f"[config4] omega={w:.3g}: all 48 instances match the FP64 path (worst so far {worst:.1e})", flush=True)
    torch.cuda.empty_cache()

    # two instances end to end against the oracle
    small = int(np.argmin(Ns))
    big = next(i for i, d in enumerate(desc) if d["n"] >= 8)
    opw = orc.planewaves(spec)
    for i in (small, big):
        d = desc[i]
        o = orc.Geometry(geoms[i].V, geoms[i].T)
        tab = o.static_table(10)
        omat = (orc.material("janus", up=4 + 0.1j, down=12 + 1j) if d["janus"]
                else orc.material("iso", eps=4 + 0.1j))
        for f in (0, len(omegas) - 1):
            w = float(omegas[f])
            Mo = o.assemble_colmajor(w, omat, tab)
            Jo = orc.lu_solve(Mo, o.rhs(w, opw), overwrite=True)
            del Mo
            Po = o.emt_power(w, opw, Jo)
            rel = float(np.max(np.abs(r["power"][f, i] - Po) / np.abs(Po)))
            print(f"[config4] instance {i} (n={d['n']}, N={d['N']}, janus={d['janus']}) omega={w:.3g}: "
                  f"power rel err vs oracle {rel:.1e} ({time.perf_counter() - t0:.0f}s)", flush=True)
            assert rel < 1e-9, (i, f, rel)
