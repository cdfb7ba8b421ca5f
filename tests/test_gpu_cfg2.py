"""Full-size GPU parity on the 1M-trajectory workload: 1,000,000 random walks
(lengths lognormal, median 64, sigma 0.8, clipped to [8, 2048]; 88.2M points),
a 1024 x 1024 GT-quadtree with 256-point leaves, its MortonGrid form, FSG with
1/1024-degree cells and STIG with 256-point leaves, and 10,000 squares with
log-uniform side in [1e-4, 1e-1].

Every index answers the whole batch through the plain drop-in (host leaf
selection, per-task refine) and through the batched form (device selection,
leaf-major join, a captured graph replayed), and every rectangle's pair count
and order-independent digest must equal the oracle's restatement of that
index: Grid::rangeQueryBatch (Grid.cpp:341-405), MortonGrid's walk
(MortonGrid.cpp:355-393), FSG's cell range (FSG.cpp:183-312) and the corrected
STIG search (STIG.cpp:179-252), run on 16 host threads.  On 200 sampled
rectangles the full (trajectory id, point id) sets are compared, for the
plain path and both batched forms.  Bar: bit-exact."""
import numpy as np
import pytest

import oracle as O
from gat_amd import api, workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

THREADS = 16
SAMPLE = 200


@pytest.fixture(scope="module")
def cfg2(gpu):
    w = W.CFG2
    db = w.db(threads=THREADS)
    rects = w.rects()
    assert db.ntraj == 1_000_000 and len(rects) == 10_000
    rng = np.random.default_rng(2002)
    sample = np.sort(rng.choice(len(rects), SAMPLE, replace=False))
    print(f"[cfg2] {db.ntraj:,} trajectories, {db.npoints:,} points, {len(rects):,} rectangles", flush=True)
    return db, rects, sample


def _keys(per_query):
    """Each query's (trajectory id, point id) pairs as a sorted int64 array:
    equal arrays are equal sets (no query holds a pair twice)."""
    out = []
    for r in per_query:
        a = np.asarray(r, np.int64).reshape(-1, 2)
        out.append(np.sort((a[:, 0] << 32) | a[:, 1]))
    return out


def _same_sets(got, want):
    return len(got) == len(want) and all(np.array_equal(g, w) for g, w in zip(got, want))


def _check_full(ix, oix, db, rects, sample, name):
    want_c, want_d = oix.range(rects, "digest", nthreads=THREADS)
    assert want_c.sum() > 600_000_000, name  # the batch really is the full workload
    print(f"[cfg2] {name}: oracle counts and digests", flush=True)
    plain = ix.range_pairs(rects)
    assert (plain.counts() == want_c).all(), f"{name}: plain counts"
    assert (plain.digests() == want_d).all(), f"{name}: plain digests"
    print(f"[cfg2] {name}: plain path equals the oracle ({int(want_c.sum()):,} pairs)", flush=True)
    first = ix.range_batched(rects, graph=True)
    assert (first.counts() == want_c).all() and (first.digests() == want_d).all(), f"{name}: batched"
    again = ix.range_batched(rects, graph=True)
    assert again.replayed, f"{name}: the second batch must replay the captured graph"
    assert (again.counts() == want_c).all() and (again.digests() == want_d).all(), f"{name}: graph replay"
    print(f"[cfg2] {name}: batched form and its graph replay equal the oracle", flush=True)
    # Full pair sets on the sample: plain, direct form, leaf-major form.
    sub = rects[sample]
    want = _keys(oix.range(sub, nthreads=THREADS))
    print(f"[cfg2] {name}: oracle pair sets of {SAMPLE} rectangles ({sum(map(len, want)):,} pairs)", flush=True)
    assert _same_sets(_keys(ix.range_pairs(sub).per_query()), want), f"{name}: plain sample sets"
    ix.set_batch_params(16, 1 << 30)
    got_direct = ix.range_batched(sub)
    assert _same_sets(_keys(got_direct.per_query()), want), f"{name}: direct-form sample sets"
    ix.set_batch_params(16, 0)
    got_leaf = ix.range_batched(sub)
    assert got_leaf.work_items > 0
    assert _same_sets(_keys(got_leaf.per_query()), want), f"{name}: leaf-major sample sets"
    ix.set_batch_params(16, 256)
    print(f"[cfg2] {name}: {SAMPLE} sampled pair sets equal the oracle's", flush=True)


def test_cfg2_gt_quadtree(cfg2):
    db, rects, sample = cfg2
    w = W.CFG2
    g = api.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    info = g.info()
    og = O.Grid(W.UNIT_BOX, w.grid_cell, w.leaf).add(db)
    oi = og.info()
    assert oi["axis"] == 1024 and oi["points"] == db.npoints
    assert int(info[0]) == db.npoints
    _check_full(g, og, db, rects, sample, "GT-quadtree")


def test_cfg2_morton_grid(cfg2):
    db, rects, sample = cfg2
    w = W.CFG2
    mg = api.MortonGridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    om = O.MortonGrid(W.UNIT_BOX, w.grid_cell, w.leaf).add(db)
    _check_full(mg, om, db, rects, sample, "MortonGrid")


def test_cfg2_fsg(cfg2):
    db, rects, sample = cfg2
    w = W.CFG2
    f = api.FsgIndex(W.UNIT_BOX, w.fsg_cell).build(db, device=True)
    of = O.FSG(W.UNIT_BOX, w.fsg_cell).add(db)
    _check_full(f, of, db, rects, sample, "FSG")


def test_cfg2_stig(cfg2):
    db, rects, sample = cfg2
    w = W.CFG2
    s = api.StigIndex(w.leaf).build(db, device=True)
    os_ = O.STIG().add(db, w.leaf)
    _check_full(s, os_, db, rects, sample, "STIG")
