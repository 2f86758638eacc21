"""CPU oracle of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

Forward and backward of an ensemble member in float64 numpy, written from
the definitions: layers x -> 256 -> 256 -> 256 -> nout, GELU (exact, erf form)
between layers, mean-squared-error loss over all outputs.  Parameter layout
matches the product's flat block per member:
    W1[h][kp] b1[h] W2[h][h] b2[h] W3[h][h] b3[h] W4[nout][h] b4[nout]
with kp = nin rounded up to 16 (padding columns multiply zero inputs).

`half=True` mirrors the product's number formats: weights and every layer
input are rounded to IEEE binary16 before the products (which the product
accumulates in FP32); everything else is float64.  The backward pass is
pinned by finite differences in tests/ (half=False).
"""
import math

import numpy as np
from scipy.special import erf


def shape(nin, hidden=256, nout=32):
    kp = (nin + 15) // 16 * 16
    offs = {}
    o = 0
    for l, (n, k) in enumerate([(hidden, kp), (hidden, hidden), (hidden, hidden), (nout, hidden)]):
        offs[f"W{l}"] = (o, (n, k))
        o += n * k
        offs[f"b{l}"] = (o, (n,))
        o += n
    return kp, offs, o


def unpack(p, nin, hidden=256, nout=32):
    kp, offs, n = shape(nin, hidden, nout)
    assert p.size == n
    out = {}
    for key, (o, shp) in offs.items():
        size = int(np.prod(shp))
        out[key] = p[o:o + size].reshape(shp)
    return kp, out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _h(x, half):
    return x.astype(np.float16).astype(np.float64) if half else np.asarray(x, np.float64)


def forward(p, X, nin, hidden=256, nout=32, half=True):
    """X [B][nin] -> (Y [B][nout], cache)."""
    kp, P = unpack(np.asarray(p, np.float64), nin, hidden, nout)
    B = X.shape[0]
    x = np.zeros((B, kp))
    x[:, :nin] = X
    inputs, zs = [], []
    a = _h(x, half)
    for l in range(4):
        W = _h(P[f"W{l}"], half)
        z = a @ W.T + P[f"b{l}"]
        inputs.append(a)
        zs.append(z)
        if l < 3:
            a = _h(gelu(z), half)
    return zs[-1], {"inputs": inputs, "zs": zs, "P": P, "kp": kp}


def loss_and_grad(p, X, T, nin, hidden=256, nout=32, half=True):
    """Mean squared error over all outputs and its gradient (flat, same layout as p)."""
    Y, c = forward(p, X, nin, hidden, nout, half)
    n = Y.size
    d = Y - T
    loss = float(np.sum(d * d) / n)
    dz = 2.0 * d / n
    kp, offs, npar = shape(nin, hidden, nout)
    g = np.zeros(npar)
    for l in range(3, -1, -1):
        a = c["inputs"][l]
        o, shp = offs[f"W{l}"]
        g[o:o + shp[0] * shp[1]] = (dz.T @ a).ravel()
        o, shp = offs[f"b{l}"]
        g[o:o + shp[0]] = dz.sum(0)
        if l == 0:
            break
        W = _h(c["P"][f"W{l}"], half)
        dz = (dz @ W) * gelu_grad(c["zs"][l - 1])
    return loss, g, Y
