"""Seeded synthetic blob systems (positions, circulations, core radii).

All draws come from a counter-based splitmix64 stream so that a (seed, stream,
index) triple always yields the same number, independent of array sizes:
uniforms are the top 53 bits, normals are Box-Muller pairs.

Workloads (sizes as named by the benchmark configurations):
  uniform_square   N blobs uniform in [0,1]^2, circulation N(0, 1e-3^2), core 0.01
  sheet            N blobs along y = 0.01 sin(2 pi x), normal jitter N(0, 1e-4^2),
                   circulation 1/N, core 1e-3
  gaussian_mixture N blobs from 8 isotropic Gaussians (centres U[0,1]^2, std
                   log-uniform in [1e-3, 5e-2]), circulation N(0, 1/N), core 5e-4
  batch_systems    256 systems, N log-uniform in [2e3, 2e5], uniform or 2-cluster
  elliptic_patch   N blobs in an ellipse 0.3 x 0.15 with Gaussian vorticity, core 1e-3
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Counter-based splitmix64: output for counter idx of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0x8CB92BA72F3D8DD7) & 0xFFFFFFFFFFFFFFFF)
        z = base + (idx.astype(np.uint64) + np.uint64(1)) * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) doubles (53-bit)."""
    z = splitmix64(seed, stream, np.arange(offset, offset + n, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller on two uniform streams."""
    u1 = uniform(seed, 2 * stream + 1000, n)
    u2 = uniform(seed, 2 * stream + 1001, n)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def uniform_square(n: int, seed: int = 1, circ_std: float = 1e-3, core: float = 0.01):
    xy = np.stack([uniform(seed, 0, n), uniform(seed, 1, n)], axis=1)
    cir = circ_std * normal(seed, 2, n)
    cor = np.full(n, core)
    return np.ascontiguousarray(xy), cir, cor


def sheet(n: int, seed: int = 2, jitter: float = 1e-4, core: float = 1e-3):
    x = (np.arange(n, dtype=np.float64) + 0.5) / n
    y = 0.01 * np.sin(2.0 * np.pi * x)
    # unit normal of the curve (x, 0.01 sin 2 pi x)
    dy = 0.02 * np.pi * np.cos(2.0 * np.pi * x)
    inv = 1.0 / np.sqrt(1.0 + dy * dy)
    eps = jitter * normal(seed, 0, n)
    xy = np.stack([x - dy * inv * eps, y + inv * eps], axis=1)
    cir = np.full(n, 1.0 / n)
    cor = np.full(n, core)
    return np.ascontiguousarray(xy), cir, cor


def gaussian_mixture(n: int, seed: int = 3, k: int = 8, std_lo: float = 1e-3, std_hi: float = 5e-2,
                     core: float = 5e-4):
    centres = np.stack([uniform(seed, 10, k), uniform(seed, 11, k)], axis=1)
    std = np.exp(np.log(std_lo) + (np.log(std_hi) - np.log(std_lo)) * uniform(seed, 12, k))
    comp = np.minimum((uniform(seed, 13, n) * k).astype(np.int64), k - 1)
    g = np.stack([normal(seed, 14, n), normal(seed, 15, n)], axis=1)
    xy = centres[comp] + std[comp, None] * g
    cir = normal(seed, 16, n) / np.sqrt(n)
    cor = np.full(n, core)
    return np.ascontiguousarray(xy), cir, cor


def batch_systems(count: int = 256, seed: int = 4, n_lo: float = 2e3, n_hi: float = 2e5,
                  core: float = 1e-3, scale: float = 1.0):
    """List of (xy, cir, cor) systems; scale < 1 shrinks every N (for tests)."""
    u = uniform(seed, 20, count)
    ns = np.exp(np.log(n_lo) + (np.log(n_hi) - np.log(n_lo)) * u)
    ns = np.maximum((ns * scale).astype(np.int64), 1)
    layout = uniform(seed, 21, count) < 0.5
    out = []
    for s in range(count):
        n = int(ns[s])
        ss = seed * 1000 + s
        if layout[s]:
            xy, cir, _ = uniform_square(n, seed=ss)
        else:
            c = np.stack([uniform(ss, 30, 2), uniform(ss, 31, 2)], axis=1)
            sd = 0.02 + 0.08 * uniform(ss, 32, 2)
            comp = (uniform(ss, 33, n) < 0.5).astype(np.int64)
            g = np.stack([normal(ss, 34, n), normal(ss, 35, n)], axis=1)
            xy = np.ascontiguousarray(c[comp] + sd[comp, None] * g)
        cir = normal(ss, 36, n) / np.sqrt(n)
        out.append((xy, cir, np.full(n, core)))
    return out


def elliptic_patch(n: int, seed: int = 5, a: float = 0.3, b: float = 0.15, core: float = 1e-3,
                   total_circulation: float = 1.0):
    """Blobs uniform in the ellipse, circulation weighted by a Gaussian vorticity profile."""
    r = np.sqrt(uniform(seed, 40, n))
    t = 2.0 * np.pi * uniform(seed, 41, n)
    xy = np.stack([0.5 + a * r * np.cos(t), 0.5 + b * r * np.sin(t)], axis=1)
    w = np.exp(-4.0 * r * r)
    cir = total_circulation * w / w.sum()
    return np.ascontiguousarray(xy), cir, np.full(n, core)


def sample_targets(n: int, count: int, seed: int = 99) -> np.ndarray:
    """count distinct indices of 0..n-1 (seeded, without replacement)."""
    if count >= n:
        return np.arange(n)
    keys = splitmix64(seed, 50, np.arange(n, dtype=np.uint64))
    return np.sort(np.argsort(keys, kind="stable")[:count])
