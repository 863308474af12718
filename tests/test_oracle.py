"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

- against the reference's published triangle counts (README.md:505-528);
- against golden fixtures written by the compiled reference
  (tests/golden/make_fixtures.py): canonical triangle sets, duplicate use;
- against the compiled reference live, where oracle/_ref exists;
- predicates against exact integer closed forms and the reference's own
  predicates; the exact checker against planted mistakes.
"""
import json
import math
import os
import random

import numpy as np
import pytest

import pyoracle as O
from libqdt_amd import workloads as W

import _fixture_inputs

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))


def sgn(v):
    return (v > 0) - (v < 0)


def orient_int(a, b, c):
    return sgn((a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0]))


def incircle_int(a, b, c, d):
    adx, ady = a[0] - d[0], a[1] - d[1]
    bdx, bdy = b[0] - d[0], b[1] - d[1]
    cdx, cdy = c[0] - d[0], c[1] - d[1]
    return sgn((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
               + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))


@pytest.mark.parametrize("n", [100, 1000, 10000, 100000])
def test_published_counts(n):
    pts = W.uniform(n, 20260607 + n, 1000.0)
    assert len(O.triangulate(pts).tris) == FIX["published_counts"]["bench_uniform"][str(n)]


def test_published_dist_counts_exact_generators():
    # the uniform, grid and band cases of README.md:518-528 use only integer ops and
    # uni(); regenerate the uniform one (Rng(seed + 1000 + 0), kExtent = 1000)
    pts = W.uniform(1000000, 20260607 + 1000, 1000.0)
    assert len(O.triangulate(pts).tris) == FIX["published_counts"]["dist_1M"]["uniform random"]
    w = 1000
    g = np.stack([np.arange(1000000) % w, np.arange(1000000) // w], axis=1).astype(np.float64)
    assert len(O.triangulate(g).tris) == FIX["published_counts"]["dist_1M"]["regular grid (shuffled)"]


@pytest.mark.parametrize("name", sorted(_fixture_inputs.inputs().keys()))
def test_fixtures(name):
    pts = _fixture_inputs.inputs()[name]
    case = FIX["cases"][name]
    r = O.triangulate(pts)
    assert len(r.tris) == case["triangles"]
    assert "%016x" % O.canon_hash(r.tris) == case["fnv"]
    if "canonical" in case:
        assert O.canonical(r.tris).tolist() == case["canonical"]
        if case.get("used") is not None:
            used = [int(r.rep[i] == i and len(r.tris) > 0) for i in range(len(pts))]
            assert used == case["used"]


def test_readme_kats():
    # README.md:326-331 and test/qdt_verify.cpp:749-760
    quad = FIX["cases"]["readme quad"]["canonical"]
    assert quad == O.canonical(np.array([[3, 0, 1], [1, 2, 3]])).tolist()
    # test/qdt_verify.cpp:913-923: rep/used of the duplicate set
    assert FIX["cases"]["duplicates 5"]["used"] == [1, 1, 0, 1, 0]


def _fnv1a64_py(rows):
    """Textbook FNV-1a-64 (offset basis 0xcbf29ce484222325, prime
    0x100000001b3) over every index as 8 little-endian bytes, in pure
    Python: an implementation independent of the oracle's."""
    h = 0xcbf29ce484222325
    for row in rows:
        for v in row:
            for b in int(v).to_bytes(8, "little"):
                h ^= b
                h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_canonical_form_and_hash_independently():
    """orc_canonicalize / orc_canon_hash against hand-written canonical lists
    and a pure-Python FNV-1a-64: the rotation puts the smallest index first
    (keeping the cyclic order, so orientation is kept), then rows sort
    lexicographically."""
    tris = np.array([[3, 0, 1], [1, 2, 3], [9, 4, 7], [7, 9, 4], [2, 8, 5]], np.uint32)
    want = [[0, 1, 3], [1, 2, 3], [2, 8, 5], [4, 7, 9], [4, 7, 9]]
    assert O.canonical(tris).tolist() == want
    assert O.canon_hash(tris) == _fnv1a64_py(want)
    assert O.canon_hash(np.zeros((0, 3), np.uint32)) == 0xcbf29ce484222325
    # the fixtures' small canonical lists, written by the compiled reference
    n = 0
    for name, case in FIX["cases"].items():
        if "canonical" in case and case["canonical"]:
            assert "%016x" % _fnv1a64_py(case["canonical"]) == case["fnv"], name
            n += 1
    assert n >= 5
    # a larger random list: canonicalise in numpy by hand, hash in Python
    rng = np.random.default_rng(5)
    t = rng.integers(0, 1000, size=(500, 3)).astype(np.uint32)
    r = np.argmin(t, axis=1)
    rot = np.stack([t[np.arange(len(t)), (r + k) % 3] for k in range(3)], axis=1)
    rot = rot[np.lexsort((rot[:, 2], rot[:, 1], rot[:, 0]))]
    assert O.canonical(t).tolist() == rot.tolist()
    assert O.canon_hash(t) == _fnv1a64_py(rot.tolist())


def test_dedupe_rule_closed_form():
    rng = np.random.default_rng(3)
    pts = rng.integers(0, 20, size=(3000, 2)).astype(np.float64)
    pts[::7] = -0.0  # -0.0 and 0.0 are the same coordinate
    r = O.triangulate(pts)
    order = np.lexsort((np.arange(len(pts)), pts[:, 1], pts[:, 0]))
    rep = np.empty(len(pts), np.int64)
    first = None
    for k, i in enumerate(order):
        if k == 0 or not (pts[i, 0] == pts[order[k - 1], 0] and pts[i, 1] == pts[order[k - 1], 1]):
            first = i
        rep[i] = first
    assert np.array_equal(r.rep.astype(np.int64), rep)
    assert r.unique == len(np.unique(pts + 0.0, axis=0))


def test_predicates_closed_forms():
    rng = random.Random(11)
    L = O.lib()
    for it in range(6000):
        R = [3, 30, 1 << 14, 1 << 28][it % 4]
        p = [(rng.randint(-R, R), rng.randint(-R, R)) for _ in range(4)]
        if it % 5 == 0:  # exactly collinear
            t = rng.randint(-3, 3)
            p[2] = (p[0][0] + t * (p[1][0] - p[0][0]), p[0][1] + t * (p[1][1] - p[0][1]))
        if it % 5 == 1:  # axis-aligned rectangle: exactly cocircular
            x0, x1, y0, y1 = p[0][0], p[1][0], p[0][1], p[1][1]
            p = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
        a, b, c, d = p
        for k in (0, -1000, 600):
            s = math.ldexp(1.0, k)
            f = [(x * s, y * s) for x, y in p]
            o = L.orc_orient2d(*f[0], *f[1], *f[2])
            assert sgn(o) == orient_int(a, b, c), (p, k)
            ic = L.orc_incircle(*f[0], *f[1], *f[2], *f[3])
            assert sgn(ic) == incircle_int(a, b, c, d), (p, k)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
def test_predicates_vs_reference():
    rng = random.Random(5)
    R = O.ref()
    L = O.lib()
    for it in range(20000):
        mode = it % 3
        if mode == 0:
            p = [(rng.random(), rng.random()) for _ in range(4)]
        elif mode == 1:
            base = rng.random()
            p = [(base + rng.randint(-2, 2) * 1e-16, base + rng.randint(-2, 2) * 1e-16) for _ in range(4)]
        else:
            p = [(rng.randint(0, 3) * 0.1, rng.randint(0, 3) * 0.1) for _ in range(4)]
        assert sgn(L.orc_orient2d(*p[0], *p[1], *p[2])) == sgn(R.ref_orient2d(*p[0], *p[1], *p[2]))
        assert sgn(L.orc_incircle(*p[0], *p[1], *p[2], *p[3])) == sgn(R.ref_incircle(*p[0], *p[1], *p[2], *p[3]))


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
def test_oracle_vs_reference_small_sets():
    # like test/qdt_verify.cpp:1165 (sc_batch): small sets with tiny ranges,
    # many duplicates and cocircular points
    rng = np.random.default_rng(1000)
    for it in range(600):
        n = int(rng.integers(0, 70))
        R = [2, 3, 4, 6, 10, 30, 100, 10000, 1 << 20][int(rng.integers(0, 9))]
        pts = rng.integers(-R, R + 1, size=(n, 2)).astype(np.float64)
        got = O.triangulate(pts)
        want, used = O.ref_triangulate(pts)
        assert np.array_equal(O.canonical(got.tris), O.canonical(want)), (it, n, R)
        if n:
            mine = (got.rep == np.arange(n)).astype(np.uint8) if len(got.tris) else np.zeros(n, np.uint8)
            assert np.array_equal(mine, used), (it, n, R)


def test_checker_accepts_oracle_output():
    for pts in (W.uniform(5000, 3), W.lattice(30, 1e-3), _fixture_inputs.inputs()["two rows"]):
        r = O.triangulate(pts)
        code, msg, info = O.check(pts, r.tris)
        assert code == 0, msg
        assert info["T"] == 2 * info["m"] - 2 - info["b"]


def test_checker_catches_planted_mistakes():
    pts = W.uniform(2000, 4)
    tris = O.triangulate(pts).tris.copy()
    assert O.check(pts, tris)[0] == 0
    # reversed orientation
    bad = tris.copy()
    bad[0] = bad[0][::-1]
    assert O.check(pts, bad)[0] != 0
    # a missing triangle
    assert O.check(pts, tris[1:])[0] != 0
    # a non-Delaunay diagonal: flip the shared edge of two adjacent triangles
    edge = {}
    for t, (a, b, c) in enumerate(tris.tolist()):
        for u, v, w in ((a, b, c), (b, c, a), (c, a, b)):
            edge[(u, v)] = (t, w)
    for (u, v), (t, w) in edge.items():
        if (v, u) in edge:
            t2, w2 = edge[(v, u)]
            # triangles (u, v, w) and (v, u, w2) -> (w, w2, v) and (w2, w, u) if convex
            if O.orient2d(pts[w], pts[w2], pts[v]) > 0 and O.orient2d(pts[w2], pts[w], pts[u]) > 0:
                bad = tris.copy()
                bad[t] = (w, w2, v)
                bad[t2] = (w2, w, u)
                code, msg, _ = O.check(pts, bad)
                assert code != 0 and "Delaunay" in msg
                break
    # a duplicate (non-lowest) index
    dup = np.concatenate([pts, pts[:1]], axis=0)
    bad = tris.copy()
    k = int(np.argwhere(bad == 0)[0][0])
    bad[k][bad[k] == 0] = len(pts)
    assert O.check(dup, bad)[0] != 0
