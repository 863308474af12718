"""Insertion and constraints through the C ABI (qdt_insert_point*,
qdt_insert_constraint*, libqdt/qdt.h:184-314) after a GPU batch call, against
the compiled reference doing the same calls on the same points
(libqdt/GuibasStolfiDwyer_Wrapper.cpp:543-876, :1157-1285).

Points in general position: the (constrained) Delaunay triangulation is
unique, so the canonical triangle lists must be identical, and so must
every status, index and duplicate representative.  Every case prints its
name before it runs."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")]


def same(t, rc, name):
    got, want = t.triangles(), rc.triangles()
    assert len(got) == len(want), name
    assert np.array_equal(O.canonical(got), O.canonical(want)), name


@pytest.mark.parametrize("n,parallel", [(20000, False), (300000, True)])
def test_insertions_match_reference(gpu, n, parallel):
    """After a batch call of n points (300k: the mesh's twins are paired on
    the device), 1,500 insertions inside and outside the hull."""
    rng = np.random.default_rng(n)
    base = W.uniform(n, 40 + n)
    extra = np.concatenate([rng.random((1000, 2)), rng.random((500, 2)) * 3 - 1])
    print("case: %d points + %d insertions" % (n, len(extra)), flush=True)
    with Q.Triangulation() as t, O.RefContext() as rc:
        assert t.execute(base) == Q.QDT_OK and rc.execute(base) == 0
        for i, (x, y) in enumerate(extra):
            assert t.insert_point(x, y, parallel) == rc.insert(x, y), i
        same(t, rc, "insertions")
        assert t.points_size() == n + len(extra)
        v = t.verify()
        assert v["ok"], v


def test_insertions_with_duplicates_on_a_lattice(gpu):
    rng = np.random.default_rng(7)
    base = rng.integers(0, 30, size=(400, 2)).astype(np.float64)
    print("case: lattice insertions", flush=True)
    with Q.Triangulation() as t, O.RefContext() as rc:
        assert t.execute(base) == Q.QDT_OK and rc.execute(base) == 0
        for i in range(600):
            x, y = rng.integers(-5, 36, size=2).astype(np.float64)
            assert t.insert_point(x, y) == rc.insert(x, y), i
        assert t.triangles_size() == len(rc.triangles())
        used = rc.used()
        assert [t.point_is_used(i) for i in range(t.points_size())] == used.tolist()
        v = t.verify()
        assert v["ok"], v


def test_constraints_match_reference(gpu):
    rng = np.random.default_rng(11)
    base = W.uniform(30000, 12)
    print("case: 30k points, 400 constraints, then 300 insertions", flush=True)
    with Q.Triangulation() as t, O.RefContext() as rc:
        assert t.execute(base) == Q.QDT_OK and rc.execute(base) == 0
        acc = 0
        for i in range(400):
            a = int(rng.integers(0, len(base)))
            d = ((base - base[a]) ** 2).sum(axis=1)
            b = int(np.argsort(d)[int(rng.integers(3, 60))])
            s1, s2 = t.insert_constraint(a, b), rc.constrain(a, b)
            assert s1 == s2, (i, a, b, s1, s2)
            acc += s1 == Q.QDT_OK
        assert acc > 100
        same(t, rc, "constraints")
        for i in range(300):
            x, y = rng.random(2)
            assert t.insert_point(x, y) == rc.insert(x, y), i
        same(t, rc, "constraints then insertions")


def test_rejected_constraint_keeps_the_context(gpu):
    """A rejected constraint changes nothing, and the context stays usable
    for later insertions and constraints (qdt.h:61-77)."""
    pts = np.array([(i, j) for i in range(8) for j in range(8)], np.float64)
    print("case: rejections", flush=True)
    with Q.Triangulation() as t, O.RefContext() as rc:
        assert t.execute(pts) == Q.QDT_OK and rc.execute(pts) == 0
        before = O.canonical(t.triangles())
        for a, b in ((0, 63), (0, 0), (0, 64), (0, 9)):
            assert t.insert_constraint(a, b) == rc.constrain(a, b)
        assert t.insert_constraint(0, 63) == Q.QDT_ERR_THROUGH_POINT
        assert t.insert_constraint(3, 3) == Q.QDT_ERR_SAME_POINT
        assert t.insert_constraint(0, 64) == Q.QDT_ERR_INVALID_INDEX
        # (0, 9), a cell diagonal, was accepted above: the other diagonal crosses it
        assert t.insert_constraint(0, 9) == rc.constrain(0, 9) == Q.QDT_OK
        assert t.insert_constraint(1, 8) == rc.constrain(1, 8) == Q.QDT_ERR_CROSSES_CONSTRAINT
        assert t.triangles_size() == len(before)
        assert t.insert_point(3.5, 3.25) == rc.insert(3.5, 3.25)
        assert t.insert_constraint(10, 64) == rc.constrain(10, 64)
        assert t.triangles_size() == len(rc.triangles())
        assert t.points_size() == 65


def test_batch_call_replaces_edits(gpu):
    pts = W.uniform(1000, 3)
    with Q.Triangulation() as t:
        assert t.execute(pts) == Q.QDT_OK
        assert t.insert_point(0.5, 0.5)[0] == Q.QDT_OK
        assert t.insert_constraint(0, 1) in (Q.QDT_OK, Q.QDT_ERR_CROSSES_CONSTRAINT, Q.QDT_ERR_THROUGH_POINT)
        assert t.execute(pts) == Q.QDT_OK
        assert t.points_size() == 1000
        want = O.triangulate(pts)
        assert np.array_equal(O.canonical(t.triangles()), O.canonical(want.tris))
