"""CPU oracle of the cost model (TEST INFRASTRUCTURE ONLY).

numpy float64 forward pass and backpropagation of the 16 -> 64 -> 64 -> 64 -> 2
ReLU MLP with the loss (1/n) sum_i sum_o (y[i,o] - T[i,o])^2, in the
parameter layout of
gpu-colocationmininig_amd/csrc/colo_mlp.hip (row-major W[out][in], then bias,
per layer).  Pinned by central finite differences in tests/test_mlp.py.
The model has no counterpart in the reference.
"""
import numpy as np

IN, H, OUT = 16, 64, 2
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (OUT, H)), ("b4", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def forward(p, X, mean, std):
    w = unpack(p)
    a0 = (np.asarray(X, np.float64) - mean) / std
    z1 = a0 @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ w["W3"].T + w["b3"]
    h3 = np.maximum(z3, 0)
    y = h3 @ w["W4"].T + w["b4"]
    return y, (a0, h1, h2, h3)


def loss(p, X, T, mean, std):
    y, _ = forward(p, X, mean, std)
    return float(np.sum((y - T) ** 2) / len(X))


def grad(p, X, T, mean, std):
    w = unpack(p)
    y, (a0, h1, h2, h3) = forward(p, X, mean, std)
    n = len(X)
    dy = 2.0 * (y - np.asarray(T, np.float64)) / n
    g = {}
    g["W4"] = dy.T @ h3
    g["b4"] = dy.sum(0)
    d3 = (dy @ w["W4"]) * (h3 > 0)
    g["W3"] = d3.T @ h2
    g["b3"] = d3.sum(0)
    d2 = (d3 @ w["W3"]) * (h2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    g["W1"] = d1.T @ a0
    g["b1"] = d1.sum(0)
    return np.concatenate([g[name].reshape(-1) for name, _ in SHAPES]), float(np.sum((y - T) ** 2) / n)
