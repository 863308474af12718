"""Inputs of the golden fixtures (tests/golden/fixtures.json).

All of them use exact IEEE operations only, so every machine rebuilds the
same points: splitmix64 uniforms, integer lattices, sums of uniforms.
"""
import numpy as np

from libqdt_amd import workloads as W


def irwin_hall_clusters(n, seed, k=8, spread=0.02):
    u = W.uni(seed, 2 * k + 8 * n)
    cx, cy = u[0:k], u[k:2 * k]
    r = u[2 * k:]
    which = (W.splitmix64(seed + 7, n) % np.uint64(k)).astype(np.int64)
    dx = (r[0:4 * n:4] + r[1:4 * n:4] + r[2:4 * n:4] + r[3:4 * n:4] - 2.0) * spread
    dy = (r[4 * n:8 * n:4] + r[4 * n + 1:8 * n:4] + r[4 * n + 2:8 * n:4] + r[4 * n + 3:8 * n:4] - 2.0) * spread
    return np.stack([cx[which] + dx, cy[which] + dy], axis=1)


def dups_of(base, extra, seed):
    pick = (W.splitmix64(seed, extra) % np.uint64(len(base))).astype(np.int64)
    return np.concatenate([base, base[pick]], axis=0)


def shuffled(p, seed):
    key = W.splitmix64(seed, len(p))
    return p[np.argsort(key, kind="stable")]


def inputs():
    d = {}
    d["readme quad"] = np.array([[0, 0], [4, 0], [3, 1], [1, 1]], dtype=np.float64)
    d["unit square"] = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.float64)
    d["square and centre"] = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5], [0.25, 0.75]], dtype=np.float64)
    d["duplicates 5"] = np.array([[0, 0], [1, 0], [0, 0], [0, 1], [1, 0]], dtype=np.float64)
    d["grid 40x40"] = W.lattice(40)
    d["grid 17x17 shuffled"] = shuffled(W.lattice(17), 3)
    d["two rows"] = np.array([[i, r] for i in range(100) for r in (0, 1)], dtype=np.float64)
    d["collinear plus one"] = np.array([[i, 7] for i in range(50)] + [[25, 8]], dtype=np.float64)
    d["uniform 100k seed 0"] = W.uniform(100000, 0)
    d["uniform 1M seed 5"] = W.uniform(1000000, 5)
    d["bench uniform 100000"] = W.uniform(100000, 20260607 + 100000, 1000.0)
    d["clusters 200k"] = irwin_hall_clusters(200000, 9)
    d["lattice 300x300 step 1e-3 + dups"] = shuffled(dups_of(W.lattice(300, 1e-3), 9000, 4), 6)
    d["lattice 1000x1000 step 1e-3 + 5% dups"] = dups_of(W.lattice(1000, 1e-3), 50000, 8)
    d["band 200k"] = np.stack([W.uni(12, 200000), W.uni(13, 200000) * 1e-6], axis=1)
    return d


def large_inputs():
    """Fixture cases that only the GPU tests run (the oracle takes ~10 s each)."""
    return {"C2 uniform 10M seed 1": W.uniform(10_000_000, 1)}
