"""The batched form's HIP graphs and autotuner (qdt_gpu_batch_set_graphs,
qdt_gpu_batch_set_tuning).

The first batch of a shape runs as usual and is captured; later batches of
that shape are one graph launch.  Every replay is checked against the eager
path (graphs off) on the same sets, with eager work and synchronisations
between the replays; batches whose data needs other launches than the
recorded run (duplicates, a non-finite set) must fall back to the eager path
with correct results.  Every case prints its name before it runs."""
import os

import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _tune_file(tmp_path, monkeypatch):
    monkeypatch.setenv("QDT_TUNE_FILE", str(tmp_path / "tune.txt"))


def concat(sets):
    offs = np.zeros(len(sets) + 1, dtype=np.uint64)
    for i, p in enumerate(sets):
        offs[i + 1] = offs[i] + len(p)
    return np.ascontiguousarray(np.concatenate(sets)), offs


def hashes(b, k):
    allt = b.all_triangles()
    return [O.canon_hash(b.triangles(i, allt)) for i in range(k)], allt


def eager(sets):
    pts, offs = concat(sets)
    b = Q.Batch()
    b.set_graphs(False)
    b.set_tuning(256, 16)
    assert b.execute(pts, offs) == Q.QDT_OK
    h, _ = hashes(b, len(sets))
    st = [b.status(i) for i in range(len(sets))]
    b.close()
    return h, st


def test_replays_equal_the_eager_path(gpu):
    import torch

    sets = W.batch_sets(count=512, seed=3)
    pts, offs = concat(sets)
    want, _ = eager(sets)
    d = torch.from_numpy(pts).to("cuda:0")
    b = Q.Batch()
    b.set_tuning(256, 16)
    print("case: first batch of the shape (recorded and captured)", flush=True)
    assert b.execute_device(d.data_ptr(), offs) == Q.QDT_OK
    st = b.stats()
    assert st["graph_captured"] == 1 and st["graph_replay"] == 0, st
    assert hashes(b, len(sets))[0] == want
    other = W.uniform(20000, 9)
    for rep in range(4):
        print("case: replay %d" % rep, flush=True)
        assert b.execute_device(d.data_ptr(), offs) == Q.QDT_OK
        st = b.stats()
        assert st["graph_replay"] == 1 and st["graph_fallback"] == 0, st
        got, _ = hashes(b, len(sets))
        assert got == want, rep
        # eager work in between, on another context
        with Q.Triangulation() as t:
            assert t.execute(other) == Q.QDT_OK
        torch.cuda.synchronize()
    print("case: the same points in another device buffer", flush=True)
    d2 = d.clone()
    del d
    assert b.execute_device(d2.data_ptr(), offs) == Q.QDT_OK
    assert b.stats()["graph_replay"] == 1
    assert hashes(b, len(sets))[0] == want
    b.close()


def test_new_data_same_shape(gpu):
    """Other points with the same set sizes replay the graph (the guards
    pass) and give the eager path's results."""
    import torch

    sets = W.batch_sets(count=256, seed=5)
    sizes = [len(p) for p in sets]
    b = Q.Batch()
    b.set_tuning(256, 16)
    pts, offs = concat(sets)
    assert b.execute(pts, offs) == Q.QDT_OK
    for seed in (6, 7):
        print("case: same shape, seed %d" % seed, flush=True)
        sets2 = [W.mixed_set(n, seed * 1000003 + i)[0] for i, n in enumerate(sizes)]
        want, _ = eager(sets2)
        p2, o2 = concat(sets2)
        d = torch.from_numpy(p2).to("cuda:0")
        assert b.execute_device(d.data_ptr(), o2) == Q.QDT_OK
        st = b.stats()
        assert st["graph_replay"] + st["graph_fallback"] == 1, st
        assert hashes(b, len(sets2))[0] == want
    b.close()


def test_guard_failures_fall_back(gpu):
    """Duplicates (other distinct counts) and a non-finite set under a
    captured shape: the replay's guards fail, the batch runs eagerly, the
    results and statuses equal the eager path's; a later clean batch replays
    again."""
    rng = np.random.default_rng(3)
    sets = [rng.random((n, 2)) for n in (3000, 70, 12000, 500, 6000, 64, 1000)]
    b = Q.Batch()
    b.set_tuning(128, 32)
    pts, offs = concat(sets)
    assert b.execute(pts, offs) == Q.QDT_OK
    assert b.stats()["graph_captured"] == 1
    dup = [p.copy() for p in sets]
    dup[2][100:200] = dup[2][:100]
    bad = [p.copy() for p in sets]
    bad[4][17, 0] = np.nan
    tie = [p.copy() for p in sets]
    tie[0][:, 1] = np.round(tie[0][:, 1] * 8) / 8   # long runs of equal y
    for name, s2 in (("duplicates", dup), ("non-finite set", bad), ("equal y runs", tie)):
        print("case:", name, flush=True)
        want, wst = eager(s2)
        p2, o2 = concat(s2)
        assert b.execute(p2, o2) == Q.QDT_OK
        st = b.stats()
        assert st["graph_fallback"] == 1 and st["graph_replay"] == 0, (name, st)
        assert [b.status(i) for i in range(len(s2))] == wst
        assert hashes(b, len(s2))[0] == want
    print("case: clean batch after the fallbacks", flush=True)
    assert b.execute(pts, offs) == Q.QDT_OK
    assert b.stats()["graph_replay"] == 1
    b.close()


def test_every_tuning_gives_the_same_result(gpu):
    sets = W.batch_sets(count=128, seed=11)
    pts, offs = concat(sets)
    want, _ = eager(sets)
    for c in (128, 256, 512):
        for lanes in (16, 32, 64):
            print("case: cell %d, lanes %d" % (c, lanes), flush=True)
            b = Q.Batch()
            b.set_graphs(False)
            b.set_tuning(c, lanes)
            assert b.execute(pts, offs) == Q.QDT_OK
            st = b.stats()
            assert st["cell_points"] == c and st["merge_lanes"] == lanes
            assert hashes(b, len(sets))[0] == want
            b.close()


def test_autotuner_picks_and_caches(gpu):
    sets = W.batch_sets(count=64, seed=13)
    pts, offs = concat(sets)
    want, _ = eager(sets)
    print("case: autotune", flush=True)
    b = Q.Batch()
    assert b.execute(pts, offs) == Q.QDT_OK
    st = b.stats()
    assert st["tune_source"] == 1 and st["tune_ms"] > 0, st
    assert st["cell_points"] in (128, 256, 512) and st["merge_lanes"] in (16, 32, 64)
    assert hashes(b, len(sets))[0] == want
    b.close()
    assert os.path.exists(os.environ["QDT_TUNE_FILE"])
    b = Q.Batch()
    assert b.execute(pts, offs) == Q.QDT_OK
    assert b.stats()["tune_source"] == 2
    assert hashes(b, len(sets))[0] == want
    b.close()
