"""GPU parity of mixed range batches: rectangles routed by area across FSG,
the GT-quadtree and STIG, run in groups whose three device pipelines are one
captured graph.  Every rectangle's count and digest must equal the oracle's:
Grid::rangeQueryBatch restated (Grid.cpp:341-405) for the quadtree, inclusive
brute force for FSG and STIG."""
import numpy as np
import pytest

import oracle as O
from gat_amd import api, workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu):
    w = W.CFG4
    db = w.db(50_000)
    rects = w.rects(3000)
    g = api.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db)
    s = api.StigIndex(w.leaf).build(db)
    f = api.FsgIndex(W.UNIT_BOX, w.fsg_cell).build(db)
    og = O.Grid(W.UNIT_BOX, w.grid_cell, w.leaf).add(db)
    oc, od = og.range(rects, "digest", nthreads=16)
    bc, bd = O.brute_range(db, rects, "digest", nthreads=16)
    return db, rects, g, s, f, (oc, od), (bc, bd)


def _expected(route, grid_ref, brute_ref):
    c = np.where(route == 1, grid_ref[0], brute_ref[0])
    d = np.where(route == 1, grid_ref[1], brute_ref[1])
    return c, d


@pytest.mark.parametrize("graph", [True, False])
def test_mixed_matches_oracle(setup, graph):
    db, rects, g, s, f, gref, bref = setup
    m = api.MixedRange(g, s, f)
    counts, digests, route, st = m.run(rects, group_size=200, graph=graph)
    assert set(np.unique(route)) == {0, 1, 2}
    c, d = _expected(route, gref, bref)
    assert (counts == c).all() and (digests == d).all()
    assert st["groups"] >= 3 and st["pairs"] == counts.sum()
    if graph:
        assert st["replays"] >= 1


def test_mixed_routing_thresholds(setup):
    """Everything to one index, then to another: the same answers."""
    db, rects, g, s, f, gref, bref = setup
    m = api.MixedRange(g, s, f)
    for fsg_max, grid_max, want in ((1.0, 2.0, 0), (0.0, 1.0, 1), (0.0, 0.0, 2)):
        counts, digests, route, _ = m.run(rects[:500], fsg_max, grid_max, group_size=128)
        assert (route == want).all()
        c, d = _expected(route, (gref[0][:500], gref[1][:500]), (bref[0][:500], bref[1][:500]))
        assert (counts == c).all() and (digests == d).all(), want


def test_mixed_recaptures_after_engine_growth(setup):
    """A mixed run captures its groups' pipelines over the three engines'
    workspaces.  When an index's own batched call grows its engine's
    workspace between two mixed runs, the buffers the graph holds are freed:
    the next group of the same shape must be run step by step and recaptured,
    not replayed, and the answers must still equal the oracle's."""
    db, rects, g, s, f, gref, bref = setup
    m = api.MixedRange(g, s, f)
    sub = rects[:1200]
    # Everything to the GT-quadtree, twelve groups of one shape.
    c1, d1, route, st1 = m.run(sub, 0.0, 1.0, group_size=100)
    assert (route == 1).all()
    assert st1["captures"] == 1 and st1["replays"] == st1["groups"] - 1
    c, d = gref[0][:1200], gref[1][:1200]
    assert (c1 == c).all() and (d1 == d).all()
    # Grow the quadtree engine's workspace (3,000 rectangles > 100).
    big = g.range_batched(rects)
    assert (big.counts() == gref[0]).all()
    c2, d2, _, st2 = m.run(sub, 0.0, 1.0, group_size=100)
    assert st2["captures"] == 2, "the first group after the growth must be recaptured"
    assert st2["replays"] == st2["groups"] - 1
    assert (c2 == c).all() and (d2 == d).all()
    # Nothing moved since: the third run replays every group.
    c3, d3, _, st3 = m.run(sub, 0.0, 1.0, group_size=100)
    assert st3["captures"] == 2 and st3["replays"] == st3["groups"]
    assert (c3 == c).all() and (d3 == d).all()
