"""CPU ORACLE of the learned p* guess (test infrastructure only).

numpy FP64 restatement of the network the HIP kernels train and evaluate
(asteriy_2d_amd/csrc/mlp.hpp, mlp.hip): 6 -> 32 -> 32 -> 32 -> 1, tanh hidden
units, linear output, mean-squared-error loss over masked rows, Adam.  The
parameter vector layout is the kernels': W1[32][6] b1 W2[32][32] b2 W3[32][32]
b3 w4[32] b4 (2369 values).  The analytic gradient here is pinned by central
finite differences in tests/test_oracle_mlp.py; the GPU gradient is checked
against it.  Features of a face restate mlp.hpp::features, and the target
log(p*/p_ref) comes from the restated exact solver (asteriy_oracle.c).
"""
from __future__ import annotations

import numpy as np

IN, H = 6, 32
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("w4", (H,)), ("b4", (1,))]
NP = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("b"):
            d[name] = np.zeros(shp)
        else:
            fan_in = shp[1] if len(shp) == 2 else shp[0]
            fan_out = shp[0] if len(shp) == 2 else 1
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            d[name] = rng.uniform(-lim, lim, shp)
    return pack(d)


def forward(p, X):
    """X: (n, 6) -> y (n,), plus the hidden activations."""
    d = unpack(p)
    h1 = np.tanh(X @ d["W1"].T + d["b1"])
    h2 = np.tanh(h1 @ d["W2"].T + d["b2"])
    h3 = np.tanh(h2 @ d["W3"].T + d["b3"])
    y = h3 @ d["w4"] + d["b4"][0]
    return y, (h1, h2, h3)


def loss_grad(p, X, t, mask=None):
    """Mean squared error over the masked rows and its gradient."""
    d = unpack(p)
    y, (h1, h2, h3) = forward(p, X)
    w = np.ones(len(y)) if mask is None else np.asarray(mask, dtype=np.float64)
    cnt = w.sum()
    e = (y - t) * w
    loss = (e * e).sum() / cnt
    dy = 2.0 * e / cnt
    g = {}
    g["w4"] = h3.T @ dy
    g["b4"] = np.array([dy.sum()])
    d3 = np.outer(dy, d["w4"]) * (1 - h3 * h3)
    g["W3"] = d3.T @ h2
    g["b3"] = d3.sum(0)
    d2 = (d3 @ d["W3"]) * (1 - h2 * h2)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ d["W2"]) * (1 - h1 * h1)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(faces, gamma, n1=1.0, n2=0.0):
    """faces (8, n) SoA -> (n, 6) features, log p_ref, post-swap states."""
    rL, pL = faces[0], faces[1]
    uL = faces[2] * n1 + faces[3] * n2
    rR, pR = faces[4], faces[5]
    uR = faces[6] * n1 + faces[7] * n2
    sw = pL > pR
    enI, enII = np.where(sw, -uR, uL), np.where(sw, -uL, uR)
    pI, pII = np.where(sw, pR, pL), np.where(sw, pL, pR)
    rI, rII = np.where(sw, rR, rL), np.where(sw, rL, rR)
    lr = 0.5 * (np.log(rI) + np.log(rII))
    lp = 0.5 * (np.log(pI) + np.log(pII))
    cref = np.sqrt(gamma * np.exp(lp - lr))
    um = 0.5 * (enI + enII)
    X = np.stack([np.log(rI) - lr, (enI - um) / cref, np.log(pI) - lp,
                  np.log(rII) - lr, (enII - um) / cref, np.log(pII) - lp], axis=1)
    return X, lp, (rI, pI, enI, rII, pII, enII)


def two_shock(rI, pI, enI, rII, pII, enII, gamma):
    """The pattern test of kernel.cu:200-231 for two shocks (Udf > Uud)."""
    a = np.sqrt(rI * ((gamma + 1) * pII + (gamma - 1) * pI) / 2.0)
    return (enI - enII) > (pII - pI) / a
