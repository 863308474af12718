"""Synthetic point sets for the batch Delaunay path.

The generator is splitmix64, the reference bench's platform-independent
generator (test/qdt_bench.cpp:264-293): next() adds the golden gamma to the
state and mixes it; uni() = (next() >> 11) * 2^-53.  It is vectorised here
with numpy so that tens of millions of points take well under a second.
"""
import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """Outputs number start .. start+count-1 of splitmix64(seed)."""
    k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uni(seed, count, start=0):
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(n, seed, extent=1.0):
    """x = extent * uni(), then y = extent * uni(), per point (qdt_bench.cpp:637-648)."""
    u = uni(seed, 2 * n).reshape(n, 2)
    return u * extent if extent != 1.0 else u


def _gauss_pairs(u):
    """Box-Muller with the second value discarded (qdt_bench.cpp:283-292); u is (k, 2)."""
    u1 = np.maximum(u[:, 0], 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u[:, 1])


def gaussian(n, seed, mean=0.5, sigma=0.1):
    """Each coordinate mean + sigma * gauss(), x then y (qdt_bench.cpp:660-670)."""
    u = uni(seed, 4 * n).reshape(n, 2, 2)
    gx = _gauss_pairs(u[:, 0, :])
    gy = _gauss_pairs(u[:, 1, :])
    return np.stack([mean + sigma * gx, mean + sigma * gy], axis=1)


def clustered(n, seed, k=32, sigma=0.005):
    """k Gaussian clusters with centres uniform in [0,1)^2, each point picks a cluster."""
    rng = np.random.default_rng(seed)
    centres = rng.random((k, 2))
    which = rng.integers(0, k, size=n)
    return centres[which] + sigma * rng.standard_normal((n, 2))


def lattice_dups_circle(side=2000, ndup=200000, ncircle=1000000, seed=2, step=1e-3, jitter=1e-12):
    """Degenerate-heavy stress set: a side x side lattice (i*step, j*step) in
    row-major order, then ndup exact copies of lattice points, then ncircle
    points on the unit circle (centre 0) with radius 1 + U[-jitter, jitter]."""
    i = np.arange(side, dtype=np.float64)
    gx, gy = np.meshgrid(i * step, i * step, indexing="ij")
    lat = np.stack([gx.ravel(), gy.ravel()], axis=1)
    z = splitmix64(seed, ndup + 2 * ncircle)
    pick = (z[:ndup] % np.uint64(lat.shape[0])).astype(np.int64)
    dups = lat[pick]
    u = (z[ndup:] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    a = 2.0 * math.pi * u[0::2]
    r = 1.0 + (2.0 * u[1::2] - 1.0) * jitter
    circ = np.stack([r * np.cos(a), r * np.sin(a)], axis=1)
    return np.concatenate([lat, dups, circ], axis=0)


def lattice(side, step=1.0):
    i = np.arange(side, dtype=np.float64)
    gx, gy = np.meshgrid(i * step, i * step, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel()], axis=1)


def log_uniform_sizes(count, lo, hi, seed):
    """floor(exp(U[ln lo, ln hi])) per set."""
    u = uni(seed, count)
    return np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64).clip(lo, hi)


def mixed_set(n, seed):
    """One of uniform [0,1)^2, Gaussian (0.5, 0.1), 32-cluster mixture, chosen by the seed."""
    kind = int(splitmix64(seed, 1)[0] % np.uint64(3))
    if kind == 0:
        return uniform(n, seed + 1), "uniform"
    if kind == 1:
        return gaussian(n, seed + 1), "gaussian"
    return clustered(n, seed + 1), "clustered"


def batch_sets(count=4096, lo=64, hi=65536, seed=3):
    """The batched workload: count sets of log-uniform size, mixed distributions."""
    sizes = log_uniform_sizes(count, lo, hi, seed)
    return [mixed_set(int(s), seed * 1000003 + i)[0] for i, s in enumerate(sizes)]
