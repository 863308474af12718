"""The GPU predicate code (libqdt_amd/csrc/predicates.hpp) compiled for the
host: its exact paths (62-bit integer fast path and the wide accumulator)
against exact rational arithmetic and the oracle, including exponent spreads
of 2^2000 and subnormals."""
import ctypes
import os
import random
from fractions import Fraction as F

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "libpredhost.so")

pytestmark = pytest.mark.skipif(not os.path.exists(SO), reason="host predicate build missing")


def lib():
    L = ctypes.CDLL(SO)
    d = ctypes.c_double
    for f, na in (("ph_orient2d_exact", 6), ("ph_orient2d_general", 6), ("ph_incircle_exact", 8),
                  ("ph_incircle_general", 8)):
        getattr(L, f).restype = ctypes.c_int
        getattr(L, f).argtypes = [d] * na
    L.ph_orient2d.restype = d
    L.ph_orient2d.argtypes = [d] * 6 + [ctypes.POINTER(ctypes.c_uint)]
    L.ph_incircle.restype = d
    L.ph_incircle.argtypes = [d] * 8 + [ctypes.POINTER(ctypes.c_uint)]
    return L


def sgn(x):
    return (x > 0) - (x < 0)


def orient_true(a, b, c):
    ax, ay, bx, by, cx, cy = map(F, (a[0], a[1], b[0], b[1], c[0], c[1]))
    return sgn((ax - cx) * (by - cy) - (ay - cy) * (bx - cx))


def incircle_true(a, b, c, d):
    ax, ay, bx, by, cx, cy, dx, dy = map(F, (a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1]))
    adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
    return sgn((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
               + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))


def points(rng, mode):
    if mode == 0:
        return [(float(rng.randint(-10, 10)), float(rng.randint(-10, 10))) for _ in range(4)]
    if mode == 1:
        return [(rng.random(), rng.random()) for _ in range(4)]
    if mode == 2:  # wildly different exponents, subnormals included
        return [(rng.uniform(-1, 1) * 2.0 ** rng.randint(-1074, 1000), rng.uniform(-1, 1) * 2.0 ** rng.randint(-1074, 1000))
                for _ in range(4)]
    if mode == 3:  # lattice spacing 1e-3: collinear and cocircular ties
        return [(rng.randint(0, 2000) * 1e-3, rng.randint(0, 2000) * 1e-3) for _ in range(4)]
    e = rng.randint(-60, 60)
    return [(rng.randint(-3, 3) * 2.0 ** e, rng.randint(-3, 3) * 2.0 ** e) for _ in range(4)]


def test_exact_paths_against_rationals():
    L = lib()
    rng = random.Random(1)
    for it in range(15000):
        p = points(rng, it % 5)
        if it % 7 == 0:
            p[2] = (p[1][0], p[1][1] + (p[1][1] - p[0][1]))
        a, b, c, d = p
        t = orient_true(a, b, c)
        assert L.ph_orient2d_exact(*a, *b, *c) == t
        assert L.ph_orient2d_general(*a, *b, *c) == t
        t = incircle_true(a, b, c, d)
        assert L.ph_incircle_exact(*a, *b, *c, *d) == t
        assert L.ph_incircle_general(*a, *b, *c, *d) == t


def test_coordinates_near_dbl_max():
    """Coordinates up to +-1.7e308 with opposite signs: the rounded
    differences overflow, the exact paths must still give the true sign."""
    L = lib()
    rng = random.Random(3)
    big = 1.7976931348623157e308
    for it in range(4000):
        p = []
        for _ in range(4):
            mode = rng.randint(0, 3)
            if mode == 0:
                v = (rng.choice([-1, 1]) * big * rng.uniform(0.5, 1.0), rng.choice([-1, 1]) * big * rng.uniform(0.5, 1.0))
            elif mode == 1:
                v = (rng.choice([-1.0, 1.0]) * big, rng.uniform(-1, 1) * 2.0 ** rng.randint(900, 1023))
            elif mode == 2:
                v = (rng.randint(-3, 3) * 2.0 ** 1022, rng.randint(-3, 3) * 2.0 ** 1022)
            else:
                v = (rng.uniform(-1, 1) * 2.0 ** rng.randint(-1074, 1023), rng.uniform(-1, 1))
            p.append(v)
        if it % 5 == 0:  # collinear: the midpoint of two representable points
            p[2] = (p[0][0] / 2 + p[1][0] / 2, p[0][1] / 2 + p[1][1] / 2)
        a, b, c, d = p
        cnt = ctypes.c_uint(0)
        t = orient_true(a, b, c)
        assert L.ph_orient2d_exact(*a, *b, *c) == t
        assert L.ph_orient2d_general(*a, *b, *c) == t
        assert sgn(L.ph_orient2d(*a, *b, *c, ctypes.byref(cnt))) == t
        t = incircle_true(a, b, c, d)
        assert L.ph_incircle_exact(*a, *b, *c, *d) == t
        assert L.ph_incircle_general(*a, *b, *c, *d) == t
        assert sgn(L.ph_incircle(*a, *b, *c, *d, ctypes.byref(cnt))) == t


def test_filtered_matches_exact_and_counts():
    L = lib()
    rng = random.Random(2)
    cnt = ctypes.c_uint(0)
    for it in range(5000):
        p = points(rng, 0 if it % 2 else 1)
        a, b, c, d = p
        assert sgn(L.ph_orient2d(*a, *b, *c, ctypes.byref(cnt))) == orient_true(a, b, c)
        assert sgn(L.ph_incircle(*a, *b, *c, *d, ctypes.byref(cnt))) == incircle_true(a, b, c, d)
    assert cnt.value > 0  # the small-integer half sends ties to the exact path


# ---- bounds fixed per cell (the LDS cells' filters) ----

K_EPS = 2.0 ** -53
K_ERR_ORIENT = (3.0 + 16.0 * K_EPS) * K_EPS
K_ERR_INCIRCLE = (10.0 + 96.0 * K_EPS) * K_EPS


def static_lib():
    L = lib()
    d = ctypes.c_double
    L.ph_static_bounds.restype = None
    L.ph_static_bounds.argtypes = [d, d, ctypes.POINTER(d), ctypes.POINTER(d)]
    L.ph_orient2d_det.restype = d
    L.ph_orient2d_det.argtypes = [d] * 6
    L.ph_incircle_det.restype = d
    L.ph_incircle_det.argtypes = [d] * 8
    return L


def box_points(rng, k, x0, y0, wx, wy, near):
    """k points in [x0, x0 + wx] x [y0, y0 + wy]; with `near`, nearly
    collinear / cocircular constructions (points of a line or circle, rounded)."""
    import math
    if not near:
        return [(x0 + rng.random() * wx, y0 + rng.random() * wy) for _ in range(k)]
    cx, cy = x0 + wx / 2, y0 + wy / 2
    r = min(wx, wy) / 2
    if rng.random() < 0.5:
        return [(cx + r * math.cos(t), cy + r * math.sin(t)) for t in (rng.random() * 6.283 for _ in range(k))]
    return [(x0 + t * wx, y0 + t * wy) for t in (rng.random() for _ in range(k))]


def test_static_bounds_dominate_and_decide_right():
    """Over boxes of many sizes and aspect ratios, with random and nearly
    degenerate points: the fixed bound is at least the run-time filter's
    bound (computed here in IEEE doubles as the filter does), and every sign
    it lets through is the exact one (rationals)."""
    L = static_lib()
    rng = random.Random(7)
    bo, bi = ctypes.c_double(), ctypes.c_double()
    decided = 0
    for trial in range(3000):
        e = rng.choice([1e-9, 1e-3, 1.0, 1e3, 1e12])
        wx = e * rng.choice([1.0, 0.5, 0.01, 1e-4])
        wy = e * rng.choice([1.0, 0.3, 0.01, 1e-4])
        x0, y0 = rng.uniform(-10, 10) * e, rng.uniform(-10, 10) * e
        pts = box_points(rng, 4, x0, y0, wx, wy, near=trial % 2 == 1)
        # the extent as the cell computes it
        xs, ys = [p[0] for p in pts], [p[1] for p in pts]
        L.ph_static_bounds(max(xs) - min(xs), max(ys) - min(ys), ctypes.byref(bo), ctypes.byref(bi))
        (ax, ay), (bx, by), (cx, cy), (dx, dy) = pts
        # orient2d of a, b, c
        dl = (ax - cx) * (by - cy)
        dr = (ay - cy) * (bx - cx)
        det = L.ph_orient2d_det(ax, ay, bx, by, cx, cy)
        assert det == dl - dr
        if bo.value > 0:
            assert bo.value >= K_ERR_ORIENT * (abs(dl) + abs(dr))
            if det > bo.value or -det > bo.value:
                decided += 1
                assert sgn(det) == orient_true(pts[0], pts[1], pts[2])
        # incircle of a, b, c, d
        adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
        al, bl, cl = adx * adx + ady * ady, bdx * bdx + bdy * bdy, cdx * cdx + cdy * cdy
        perm = (abs(bdx * cdy) + abs(cdx * bdy)) * al + (abs(cdx * ady) + abs(adx * cdy)) * bl + \
            (abs(adx * bdy) + abs(bdx * ady)) * cl
        det = L.ph_incircle_det(ax, ay, bx, by, cx, cy, dx, dy)
        if bi.value > 0:
            assert bi.value >= K_ERR_INCIRCLE * perm
            if det > bi.value or -det > bi.value:
                decided += 1
                assert sgn(det) == incircle_true(*pts)
    assert decided > 2000
