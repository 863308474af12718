"""qdt_gpu_execute_noalloc (include/qdt_gpu.h), the entry point that
allocates no device memory, against qdt_gpu_execute_device on the same
device-resident input: identical canonical triangle lists and duplicate
representatives, on two calls in a row (the second smaller, so stale
contents of the shared scratch array from the first call would show), plus
its size limit and argument checks.  Every case prints its name first."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


def noalloc(pts):
    import torch

    n = len(pts)
    d = torch.from_numpy(np.ascontiguousarray(pts, np.float64)).to("cuda:0")
    tris = torch.full((max(2 * n, 1), 3), -1, dtype=torch.int32, device="cuda:0")
    rep = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    st, T = Q.execute_noalloc(d.data_ptr(), n, tris.data_ptr(), 2 * n, rep.data_ptr())
    torch.cuda.synchronize()
    assert st == Q.QDT_OK, Q.status_string(st)
    return tris[:T].cpu().numpy().view(np.uint32), rep[:n].cpu().numpy().view(np.uint32)


def plain(pts):
    import torch

    d = torch.from_numpy(np.ascontiguousarray(pts, np.float64)).to("cuda:0")
    with Q.Triangulation() as t:
        assert t.execute_device(d.data_ptr(), len(pts)) == Q.QDT_OK
        tris = t.triangles()
        used = np.array([t.point_is_used(i) for i in range(len(pts))], np.uint8)
    return tris, used


def check(name, pts):
    print("case:", name, len(pts), flush=True)
    got, rep = noalloc(pts)
    want, used = plain(pts)
    assert len(got) == len(want), name
    assert np.array_equal(O.canonical(got), O.canonical(want)), name
    exp_used = (rep == np.arange(len(pts), dtype=np.uint32)).astype(np.uint8) if len(got) else \
        np.zeros(len(pts), np.uint8)
    assert np.array_equal(exp_used, used), name
    want_rep = O.triangulate(pts).rep
    assert np.array_equal(rep, want_rep), name


def test_noalloc_matches_device_entry_two_calls(gpu):
    check("uniform at the size limit", W.uniform(MAX_POINTS, 31))
    check("gaussian 50k after it", W.gaussian(50000, 32))
    check("lattice + copies", W.lattice_dups_circle(side=200, ndup=5000, ncircle=8000, seed=5))
    check("clustered 3k", W.clustered(3000, 33))
    check("two points", np.array([[0.0, 0.0], [1.0, 1.0]]))


def test_noalloc_limits_and_arguments(gpu):
    import torch

    print("case: limits", flush=True)
    n = MAX_POINTS + 1
    d = torch.zeros((n, 2), dtype=torch.float64, device="cuda:0")
    tris = torch.zeros((2 * n, 3), dtype=torch.int32, device="cuda:0")
    rep = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    assert Q.execute_noalloc(d.data_ptr(), n, tris.data_ptr(), 2 * n, rep.data_ptr())[0] == Q.QDT_ERR_TOO_MANY_POINTS
    assert Q.execute_noalloc(0, 10, tris.data_ptr(), 20, rep.data_ptr())[0] == Q.QDT_ERR_NULL_ARGUMENT
    assert Q.execute_noalloc(d.data_ptr(), 10, tris.data_ptr(), 19, rep.data_ptr())[0] == Q.QDT_ERR_OUT_OF_MEMORY
    assert Q.execute_noalloc(d.data_ptr(), 0, 0, 0, 0) == (Q.QDT_OK, 0)
    bad = torch.from_numpy(np.array([[0.0, 0.0], [1.0, np.nan], [0.0, 1.0]])).to("cuda:0")
    assert Q.execute_noalloc(bad.data_ptr(), 3, tris.data_ptr(), 6, rep.data_ptr())[0] == Q.QDT_ERR_NOT_FINITE
    # still usable after the rejections
    check("after rejections", W.uniform(1000, 34))
