"""Reference MLP 8 -> 64 -> 64 -> 2 (ReLU) in numpy, FP64 — the checker for the
GPU MLP kernels.  Weight packing as in include/fmm2d.h:
W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]."""
import numpy as np

SHAPES = [("W1", (64, 8)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (2, 64)),
          ("b3", (2,))]
NW = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = w[o:o + k].reshape(shp)
        o += k
    return out


def forward(w, x):
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(z2, 0.0)
    return h2 @ p["W3"].T + p["b3"], (x, z1, h1, z2, h2)


def loss(w, x, y):
    o, _ = forward(w, x)
    return float(((o - y) ** 2).sum() / x.shape[0])


def grad(w, x, y):
    p = unpack(w)
    o, (x, z1, h1, z2, h2) = forward(w, x)
    n = x.shape[0]
    d3 = 2.0 * (o - y) / n
    g = {"W3": d3.T @ h2, "b3": d3.sum(0)}
    d2 = (d3 @ p["W3"]) * (z2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * (z1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SHAPES])


def init(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            parts.append(rng.standard_normal(shp).ravel() * scale * np.sqrt(2.0 / shp[1]))
        else:
            parts.append(np.zeros(int(np.prod(shp))))
    return np.concatenate(parts).astype(np.float32)
