"""Device-input triangulation against the oracle over input families that
stress different parts of the path, two sizes each (fixed seeds): flat bands
(1e6:1), presorted input, half duplicates, jittered and exact lattices
(cocircular quads, collinear rows), points on a circle with its centre (a
vertex adjacent to every other one), a lattice with a circle through it,
coordinates scaled by 2^-600 and 2^600, and points on a few horizontal
lines.  The canonical triangle lists must be identical (the reference's
algorithm fixes the triangulation of degenerate inputs too)."""
import math

import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def family(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "band":
        return np.stack([rng.random(n), rng.random(n) * 1e-6], 1)
    if kind == "presorted":
        p = W.uniform(n, seed)
        return p[np.argsort(p[:, 0], kind="stable")]
    if kind == "half duplicates":
        p = W.uniform(n // 2, seed)
        return np.concatenate([p, p[rng.integers(0, len(p), n - len(p))]])
    if kind == "jittered lattice":
        s = int(math.sqrt(n)) + 1
        p = W.lattice(s, 1.0)[:n]
        return p + rng.uniform(-1e-9, 1e-9, p.shape)
    if kind == "shuffled lattice":
        s = int(math.sqrt(n)) + 1
        p = W.lattice(s, 0.5)[:n]
        return p[rng.permutation(len(p))]
    if kind == "circle and centre":
        a = rng.random(n - 1) * 2 * math.pi
        r = 1 + rng.uniform(-1e-12, 1e-12, n - 1)
        return np.concatenate([np.stack([r * np.cos(a), r * np.sin(a)], 1), [[0.0, 0.0]]])
    if kind == "lattice and circle":
        s = max(4, int(math.sqrt(n * 0.7)))
        lat = W.lattice(s, 2.0 / s)
        a = rng.random(n - len(lat)) * 2 * math.pi
        r = 1 + rng.uniform(-1e-12, 1e-12, len(a))
        return np.concatenate([lat, np.stack([r * np.cos(a), r * np.sin(a)], 1)])
    if kind == "tiny scale":
        return W.uniform(n, seed) * 2.0 ** -600
    if kind == "huge scale":
        return (W.uniform(n, seed) - 0.5) * 2.0 ** 600
    if kind == "lines":
        return np.stack([rng.random(n), rng.integers(0, 8, n) * 0.125], 1)
    raise ValueError(kind)


KINDS = ["band", "presorted", "half duplicates", "jittered lattice", "shuffled lattice", "circle and centre",
         "lattice and circle", "tiny scale", "huge scale", "lines"]


@pytest.mark.parametrize("kind", KINDS)
def test_family_matches_oracle(gpu, kind):
    import torch

    with Q.Triangulation() as t:
        for n, seed in ((3000, 101), (60000, 202)):
            p = np.ascontiguousarray(family(kind, n, seed), dtype=np.float64)
            print("case: %s, %d points, seed %d" % (kind, n, seed), flush=True)
            dev = torch.from_numpy(p).to("cuda:0")
            assert t.execute_device(dev.data_ptr(), len(p)) == Q.QDT_OK
            want = O.triangulate(p)
            got = t.triangles()
            assert np.array_equal(O.canonical(got), O.canonical(want.tris)), (kind, n)
            used = np.array([t.point_is_used(i) for i in range(0, len(p), max(1, len(p) // 500))])
            want_used = (want.rep == np.arange(len(p)))[::max(1, len(p) // 500)].astype(used.dtype)
            assert np.array_equal(used, want_used), (kind, n)
