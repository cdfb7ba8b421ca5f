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
