"""Surrogate MLP: the CPU oracle's backward pass pinned by finite
differences, and the GPU training kernels (forward, MSE, backward) checked
against the oracle on the same FP16-rounded weights and inputs."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_gelu_grad_finite_difference():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(z), fd, rtol=0, atol=1e-8)
    # GELU(0) = 0, GELU(z) -> z for large z, -> 0 for very negative z
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mo.gelu(np.array([8.0]))[0] - 8.0) < 1e-12
    assert abs(mo.gelu(np.array([-8.0]))[0]) < 1e-12


def test_oracle_backward_finite_difference():
    rng = np.random.default_rng(3)
    nin, h, nout, B = 13, 32, 16, 7
    kp, offs, npar = mo.shape(nin, h, nout)
    p = rng.standard_normal(npar) * 0.3
    X = rng.standard_normal((B, nin))
    T = rng.standard_normal((B, nout))
    loss, g, _ = mo.loss_and_grad(p, X, T, nin, h, nout, half=False)
    # parameters from every block, padding columns of W1 excluded
    idx = []
    for key, (o, shp) in offs.items():
        size = int(np.prod(shp))
        cand = np.arange(o, o + size)
        if key == "W0":
            cand = cand[(cand - o) % kp < nin]
        idx.extend(rng.choice(cand, 6, replace=False))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mo.loss_and_grad(pp, X, T, nin, h, nout, half=False)[0] -
              mo.loss_and_grad(pm, X, T, nin, h, nout, half=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
    # a dropped term (the GELU' factor) fails this check
    assert np.all(g[offs["W0"][0]:offs["W0"][0] + h * kp].reshape(h, kp)[:, nin:] == 0)


def test_param_count_matches_oracle(bfa):
    n, kp = bfa.mlp_param_count(13, 256, 32)
    kpo, _, no = mo.shape(13, 256, 32)
    assert (n, kp) == (no, kpo)


@pytest.mark.gpu
def test_gpu_mlp_gradients_match_oracle(bfa, torch_cuda):
    rng = np.random.default_rng(11)
    nin, h, nout, E, B = 13, 256, 32, 2, 300
    npar, kp = bfa.mlp_param_count(nin, h, nout)
    _, offs, _ = mo.shape(nin, h, nout)
    P = np.zeros((E, npar), np.float32)
    for e in range(E):
        for key, (o, shp) in offs.items():
            size = int(np.prod(shp))
            fan_in = shp[1] if len(shp) == 2 else 1
            P[e, o:o + size] = rng.standard_normal(size) * (1.0 / np.sqrt(fan_in) if len(shp) == 2 else 0.1)
    X = rng.standard_normal((E, B, nin)).astype(np.float32)
    T = rng.standard_normal((E, B, nout)).astype(np.float32)
    print(f"[mlp] E={E} B={B} nparam={npar}", flush=True)
    loss, grad, Y = bfa.mlp_grad(P, X, T, nin, h, nout)
    for e in range(E):
        lo, go, yo = mo.loss_and_grad(P[e].astype(np.float64), X[e].astype(np.float64), T[e].astype(np.float64),
                                      nin, h, nout, half=True)
        assert abs(loss[e] - lo) / lo < 1e-5, (loss[e], lo)
        assert np.linalg.norm(Y[e] - yo) / np.linalg.norm(yo) < 1e-4
        for key, (o, shp) in offs.items():
            size = int(np.prod(shp))
            a, b = grad[e, o:o + size], go[o:o + size]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert rel < 5e-4, (e, key, rel)
