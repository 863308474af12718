"""The predicate code compiled for the GPU (tests/native/predicates_dev.hip
wraps libqdt_amd/csrc/predicates.hpp in a kernel): its exact paths, the
62-bit integer path and the wide accumulator, against exact rational
arithmetic, including coordinates scaled to 2^-1000 and 2^1000, exponent
spreads of 2^2000 and subnormals.  The host build of the same code is checked
in test_predicates_host.py; this catches device code generation errors."""
import ctypes
import os
import random
from fractions import Fraction as F

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "libpreddev.so")


def sgn(x):
    return (x > 0) - (x < 0)


def truth(r):
    ax, ay, bx, by, cx, cy, dx, dy = map(F, r)
    o = sgn((ax - cx) * (by - cy) - (ay - cy) * (bx - cx))
    adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
    i = sgn((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
            + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))
    return o, i


def rows(rng, count):
    out = []
    for it in range(count):
        mode = it % 7
        if mode == 0:
            p = [float(rng.randint(-10, 10)) for _ in range(8)]
        elif mode == 1:
            p = [rng.random() for _ in range(8)]
        elif mode == 2:
            p = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-1074, 1000) for _ in range(8)]
        elif mode == 3:
            p = [rng.randint(0, 2000) * 1e-3 for _ in range(8)]
        elif mode == 4:
            e = rng.randint(-60, 60)
            p = [rng.randint(-3, 3) * 2.0 ** e for _ in range(8)]
        else:
            # one point set scaled to the ends of the exponent range
            s = 2.0 ** (-1000 if mode == 5 else 1000)
            p = [rng.randint(0, 6) * s * 0.25 for _ in range(8)]
        if it % 5 == 0:  # c on the line ab: orient2d ties
            p[4], p[5] = p[2], p[3] + (p[3] - p[1])
        out.append(p)
    return out


def test_device_exact_paths_against_rationals(gpu):
    if not os.path.exists(SO):
        pytest.fail("device predicate build missing: run __graft_entry__.build()")
    L = ctypes.CDLL(SO)
    L.pd_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.pd_run.restype = ctypes.c_int
    rng = random.Random(7)
    data = rows(rng, 7000)
    inp = np.ascontiguousarray(np.array(data, dtype=np.float64))
    out = np.zeros((len(data), 4), dtype=np.int32)
    assert L.pd_run(inp.ctypes.data, len(data), out.ctypes.data) == 0
    bad = []
    for k, r in enumerate(data):
        o, i = truth(r)
        got = out[k].tolist()
        if got != [o, o, i, i]:
            bad.append((k, r, got, (o, i)))
    assert not bad, "%d mismatches, first: %r" % (len(bad), bad[:3])
