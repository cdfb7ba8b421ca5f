"""GPU parity of the batched range form (device selection, leaf-major join,
captured-graph replay) against the oracle's restatement of
Grid::rangeQueryBatch (Grid.cpp:341-405) and inclusive brute force for STIG.

Bar: bit-exact result sets per query, and per-query digests equal to the
plain path's and the oracle's."""
import ctypes

import numpy as np
import pytest

import oracle as O
from gat_amd import _native as N, api, workloads as W

from tests.devbuf import DevBuf
from tests.test_gpu_range import _check_pairs, _edge_rects, _sets

pytestmark = pytest.mark.gpu


# Both forms of the batched pipeline: leaf-major join (direct_max_queries 0)
# and direct per-task refine (every batch below the threshold).
FORMS = {"leaf": 0, "direct": 1 << 30}


def _form(ix, form):
    return ix.set_batch_params(32, FORMS[form])


@pytest.fixture(scope="module")
def cfg1(gpu):
    db = W.CFG1.db()
    return db, W.CFG1.rects()


def _check_digests(res, counts, dig):
    assert (res.counts() == counts).all()
    assert (res.digests() == dig).all()


@pytest.mark.parametrize("form", FORMS)
def test_cfg1_grid_batched(cfg1, form):
    """Point-for-point on the 100 squares; per-query counts and digests on
    rectangles along cell edges, through data points and over everything
    (about 20M pairs, too many to compare as sets)."""
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = _form(api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db), form)
    res = g.range_batched(rects)
    _check_pairs(res, _sets(og.range(rects)), rects)
    edge = _edge_rects(db, W.CFG1.grid_cell, 60, 21)
    res = g.range_batched(edge)
    _check_digests(res, *og.range(edge, "digest"))
    plain = g.range_pairs(edge)
    assert plain.candidates == res.candidates  # same leaves selected
    assert res.leaf_queries > 0 and (res.work_items >= 1 or form == "direct")


@pytest.mark.parametrize("form", FORMS)
def test_cfg1_stig_batched(cfg1, form):
    db, rects = cfg1
    s = _form(api.StigIndex(W.CFG1.leaf).build(db), form)
    res = s.range_batched(rects)
    _check_pairs(res, _sets(O.brute_range(db, rects)), rects)
    edge = _edge_rects(db, W.CFG1.grid_cell, 40, 22)
    res = s.range_batched(edge)
    _check_digests(res, *O.brute_range(db, edge, "digest"))
    assert s.range_pairs(edge).candidates == res.candidates


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("cell", [0.01, 1 / 16, 1 / 1024])
def test_fsg_batched(cfg1, cell, form):
    """The flat grid's cells as leaves: every cell of the clamped row/column
    range (FSG.cpp:183-312) against inclusive brute force."""
    db, rects = cfg1
    f = _form(api.FsgIndex(W.UNIT_BOX, cell).build(db), form)
    res = f.range_batched(rects)
    _check_pairs(res, _sets(O.brute_range(db, rects)), rects)
    edge = _edge_rects(db, cell, 40, 24)
    res = f.range_batched(edge)
    _check_digests(res, *O.brute_range(db, edge, "digest"))
    assert f.range_pairs(edge).candidates == res.candidates


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("cell,leaf", [(1 / 1024, 8), (1 / 16, 4096)])
def test_tree_shapes(gpu, cell, leaf, form):
    """A deep tree (tiny leaves on a 1024^2 grid) and a shallow one whose
    leaves exceed the 1024 points a work item loads at once."""
    db = W.CFG1.db(3000)
    rects = np.concatenate([W.CFG1.rects(), _edge_rects(db, cell, 30, 23),
                            api.synth_rects(23, 50, 1e-4, 0.5, True)]).astype(np.float32)
    og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    g = _form(api.GridIndex(W.UNIT_BOX, cell, leaf).build(db), form)
    _check_digests(g.range_batched(rects), *og.range(rects, "digest"))


def test_hot_leaves_split_into_items(gpu):
    """Hundreds of rectangles over the same leaves: each leaf's list is split
    into work items of at most qb rectangles, for several qb."""
    db = W.CFG1.db(4000)
    rng = np.random.default_rng(3)
    c = 0.5 + rng.normal(0, 0.01, size=(600, 2))
    half = 10 ** rng.uniform(-3.5, -1.5, size=(600, 1))
    rects = np.concatenate([c - half, c + half], 1).astype(np.float32)
    og = O.Grid(W.UNIT_BOX, 0.01, 256).add(db)
    counts, dig = og.range(rects, "digest")
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    for qb in (64, 32, 5, 1):
        res = g.set_batch_params(qb, 0).range_batched(rects)
        assert (res.counts() == counts).all() and (res.digests() == dig).all(), qb
        assert res.work_items >= -(-res.leaf_queries // qb)
    assert res.work_items == res.leaf_queries  # one rectangle per item


def test_graph_replay(gpu):
    """Capture on the first call, replay on the next ones; a larger batch
    that outgrows the captured buffers runs step by step and is recaptured.
    Every call is checked against the oracle."""
    db = W.CFG1.db(5000)
    og = O.Grid(W.UNIT_BOX, 0.01, 256).add(db)
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db).set_batch_params(32, 0)
    small = api.synth_rects(31, 300, 1e-4, 0.05, True)
    other = api.synth_rects(32, 300, 1e-4, 0.05, True)
    large = api.synth_rects(33, 300, 0.05, 0.3, True)
    seen = []
    for rects in (small, small, other, small, large, large, small):
        res = g.range_batched(rects, graph=True)
        counts, dig = og.range(rects, "digest")
        assert (res.counts() == counts).all()
        assert (res.digests() == dig).all()
        seen.append(res.replayed)
    assert seen[0] is False and seen[1] is True
    assert seen[4] is False  # outgrew the captured capacities


@pytest.mark.parametrize("form", FORMS)
def test_empty_and_outside(gpu, form):
    db = W.CFG1.db(500)
    g = _form(api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db), form)
    res = g.range_batched(np.zeros((0, 4), np.float32))
    assert res.nq == 0 and res.pairs_total == 0
    out = np.array([[1.5, 1.5, 2.0, 2.0], [-1, -1, -0.5, -0.5]], np.float32)
    res = g.range_batched(out)
    assert (res.counts() == 0).all() and res.candidates == 0
    f = _form(api.FsgIndex(W.UNIT_BOX, 0.01).build(db), form)
    assert f.range_batched(np.zeros((0, 4), np.float32)).nq == 0
    res = f.range_batched(out)  # clamped to the border cells: candidates, no hits
    assert (res.counts() == 0).all()


def test_autotuned_batches_give_the_same_answers(cfg1):
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db)
    tuned = g.tune_batched(rects, reps=1)
    assert tuned["queries_per_item"] in (4, 8, 16, 32, 64)
    _check_digests(g.range_batched(rects), *og.range(rects, "digest"))


def _tiles(runs, total):
    """The run records cover the batch's pairs once each: runs with hits are
    disjoint, start on their work item's line-aligned region and hold
    pairs_total pairs in all."""
    r = np.sort(runs[runs["count"] > 0], order="first")
    ends = r["first"].astype(np.int64) + r["count"].astype(np.int64)
    return int(r["count"].sum()) == total and (len(r) < 2 or (r["first"][1:].astype(np.int64) >= ends[:-1]).all())


@pytest.mark.parametrize("kind", ["grid", "stig", "fsg"])
def test_result_layouts_agree(cfg1, kind):
    """Run-indexed and per-query layouts of the leaf-major form: the same sets,
    counts and digests; the run index tiles the compacted pairs and sums to the
    per-query counts."""
    db, rects = cfg1
    edge = _edge_rects(db, W.CFG1.grid_cell, 60, 31)
    batch = np.concatenate([rects, edge])
    make = {"grid": lambda: api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf),
            "stig": lambda: api.StigIndex(W.CFG1.leaf),
            "fsg": lambda: api.FsgIndex(W.UNIT_BOX, W.CFG1.fsg_cell)}[kind]
    ix = _form(make().build(db), "leaf")
    per = ix.set_result_layout(False).range_batched(batch)
    pc, pd = per.counts(), per.digests()
    assert len(per.runs()) == 0
    res = ix.set_result_layout(True).range_batched(batch)
    assert (res.counts() == pc).all() and (res.digests() == pd).all()
    runs = res.runs()
    assert len(runs) > 0 and _tiles(runs, res.pairs_total)
    assert (np.bincount(runs["q"], weights=runs["count"], minlength=len(batch)).astype(np.int64) == pc).all()
    want = _sets(O.brute_range(db, rects)) if kind != "grid" else \
        _sets(O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db).range(rects))
    got = res.per_query()
    for q in range(len(rects)):
        assert set(map(tuple, got[q].tolist())) == want[q], q


def test_runs_over_multichunk_leaves(gpu):
    """Leaves of thousands of points (whole trajectories inside one cell of the
    deepest level): each (item, rectangle) still makes one run, its hits
    counted over every chunk first."""
    lengths = np.full(300, 2048, np.int32)
    db = api.synth_walks(3, lengths, step_sigma=1e-7)
    rects = np.concatenate([np.array([[0, 0, 1, 1], [0.25, 0.25, 0.75, 0.75], [0.1, 0.1, 0.6, 0.3]], np.float32),
                            W.CFG1.rects()[:400]])
    og = O.Grid(W.UNIT_BOX, 1 / 16, 256).add(db)
    g = _form(api.GridIndex(W.UNIT_BOX, 1 / 16, 256).build(db), "leaf")
    nodes, _ = g.tree_nodes()
    leaves = nodes[nodes["first_child"] < 0]
    assert leaves["count"].max() > 1024  # several chunks per leaf
    res = g.set_result_layout(True).range_batched(rects)
    _check_pairs(res, _sets(og.range(rects)), rects)
    runs = res.runs()
    assert _tiles(runs, res.pairs_total)
    assert len(runs) == res.leaf_queries  # one record per (leaf, rectangle) entry
    counts, dig = og.range(rects, "digest")
    assert (res.digests() == dig).all()


def test_run_indexed_graph_replay(cfg1):
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = _form(api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db), "leaf").set_result_layout(True)
    want = og.range(rects, "digest")
    for i in range(4):
        res = g.range_batched(rects, graph=True)
        assert res.replayed == (i > 0)
        assert (res.counts() == want[0]).all() and (res.digests() == want[1]).all()
        assert _tiles(res.runs(), res.pairs_total)


# ------------------------------------------------------------- MortonGrid

@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("cell,leaf", [(0.01, 256), (1 / 1024, 8), (1 / 16, 4096)])
def test_mgrid_batched(cfg1, cell, leaf, form):
    """MortonGrid's implicit tree walked on the device (gat_morton_select):
    result sets identical to the oracle's MortonGrid restatement, digests on
    edge rectangles equal to it and to the plain path, the same candidates as
    the explicit GT-quadtree."""
    db, rects = cfg1
    om = O.MortonGrid(W.UNIT_BOX, cell, leaf).add(db)
    m = _form(api.MortonGridIndex(W.UNIT_BOX, cell, leaf).build(db), form)
    res = m.range_batched(rects)
    _check_pairs(res, _sets(om.range(rects)), rects)
    edge = _edge_rects(db, cell, 40, 25)
    res = m.range_batched(edge)
    counts, dig = om.range(edge, "digest")
    _check_digests(res, counts, dig)
    plain = m.range_pairs(edge)
    _check_digests(plain, counts, dig)
    assert plain.candidates == res.candidates
    g = _form(api.GridIndex(W.UNIT_BOX, cell, leaf).build(db), form)
    gres = g.range_batched(edge)
    assert gres.candidates == res.candidates and gres.leaf_queries == res.leaf_queries


@pytest.mark.parametrize("form", FORMS)
def test_mgrid_root_leaf_and_empty(gpu, form):
    """A tree that is only its root (fewer points than the leaf capacity): the
    root is taken untested (MortonGrid.cpp:355-360), so a rectangle that misses
    every point still scans them all; empty batches and rectangles outside the
    grid give empty answers."""
    db = W.CFG1.db(300)
    om = O.MortonGrid(W.UNIT_BOX, 1 / 16, 1 << 20).add(db)
    assert om.info()["leaves"] == 1
    m = _form(api.MortonGridIndex(W.UNIT_BOX, 1 / 16, 1 << 20).build(db), form)
    rects = np.array([[0.1, 0.1, 0.3, 0.3], [2, 2, 3, 3], [0, 0, 1, 1]], np.float32)
    res = m.range_batched(rects)
    _check_pairs(res, _sets(om.range(rects)), rects)
    assert res.candidates == 3 * db.npoints
    assert m.range_batched(np.zeros((0, 4), np.float32)).pairs_total == 0


def test_mgrid_graph_replay_and_device_build(gpu):
    """The store built on the device feeds the implicit tree too; captured
    pipelines replay with the implicit select and agree with the oracle."""
    db = W.CFG1.db(5000)
    om = O.MortonGrid(W.UNIT_BOX, 1 / 1024, 32).add(db)
    m = api.MortonGridIndex(W.UNIT_BOX, 1 / 1024, 32).build(db, device=True).set_batch_params(32, 0)
    small = api.synth_rects(41, 300, 1e-4, 0.05, True)
    other = api.synth_rects(42, 300, 1e-4, 0.05, True)
    seen = []
    for rects in (small, small, other, small):
        res = m.range_batched(rects, graph=True)
        _check_digests(res, *om.range(rects, "digest"))
        seen.append(res.replayed)
    assert seen[0] is False and seen[1] is True


def test_plain_after_batched_keeps_its_layout(cfg1):
    """The plain path after a run-indexed batched run on the same index (one
    result record per index): its per-query regions are read as such, so its
    digests and fetched pairs are the oracle's, and a batched run after it is
    run-indexed again."""
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = _form(api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db), "leaf")
    edge = _edge_rects(db, W.CFG1.grid_cell, 30, 27)
    counts, dig = og.range(edge, "digest")
    for _ in range(2):
        res = g.range_batched(edge)
        assert len(res.runs()) > 0
        _check_digests(res, counts, dig)
        _check_digests(g.range_pairs(edge), counts, dig)
    g.range_batched(rects)
    _check_pairs(g.range_pairs(rects), _sets(og.range(rects)), rects)


@pytest.mark.parametrize("nqueries", [3, 20000])
def test_run_indexed_abi_skips_malformed_items(gpu, nqueries):
    """The run-indexed join through the C ABI on hand-built work items that
    cover the list in order, three of them malformed (a run past the store, no
    rectangles, a list slice past the list): the refine skips them and writes
    empty run records for their entries inside the record array, where garbage
    lay before, so the count pass over the list's records counts only the good
    items.  Both count paths: LDS-aggregated (3 rectangles) and direct
    (20,000)."""
    h = N.hip()
    xy = np.array([[0.1, 0.1], [0.2, 0.2], [0.3, 0.3], [0.9, 0.9], [0.5, 0.5], [0.6, 0.6]], np.float32)
    tid = np.array([1, 1, 2, 3, 4, 4], np.uint32)
    pid = np.array([0, 1, 0, 0, 0, 1], np.uint32)
    dxy, did = DevBuf.of(xy), DevBuf.of((tid << 3) | pid)
    store = N.PointStore(dxy.ptr, did.ptr, None, 6, 3)
    sp = ctypes.byref(store)
    boxes = np.zeros((nqueries, 4), np.float32)
    boxes[:3] = [[0, 0, 0.25, 0.25], [0.25, 0.25, 1, 1], [0.55, 0.55, 0.7, 0.7]]
    qlist = np.array([0, 1, 2, 0, 1, 0], np.uint32)
    items = np.array([[0, 4, 0, 2],    # rectangles 0 and 1 over points 0..3
                      [4, 2, 2, 1],    # rectangle 2 over points 4..5
                      [5, 4, 3, 2],    # run past the store
                      [0, 2, 5, 0],    # no rectangles
                      [0, 2, 5, 3]],   # list slice past the list
                     np.uint32)
    dboxes, dq, ditems = DevBuf.of(boxes), DevBuf.of(qlist), DevBuf.of(items)
    total = DevBuf.of(np.array([5], np.uint32))
    base, bound = DevBuf(5 * 8).zero(), DevBuf(8).zero()
    tiles = DevBuf(8 * int(h.gat_scan_tiles(5))).zero()
    N.check_hip(h.gat_leaf_item_regions(ditems.ptr, total.ptr, 5, base.ptr, tiles.ptr, bound.ptr, None))
    # Regions: rectangles x points rounded up to 16 pairs, malformed items too.
    assert base.get(np.uint64, 5).tolist() == [0, 16, 32, 48, 48] and bound.get(np.uint64, 1)[0] == 64
    garbage = np.zeros(6, api.RUN)
    garbage["q"], garbage["count"] = 1, 77
    runs, pairs = DevBuf.of(garbage), DevBuf(64 * 8).zero()
    N.check_hip(h.gat_leaf_refine_runs(sp, ditems.ptr, total.ptr, 5, dq.ptr, 6, dboxes.ptr, nqueries, pairs.ptr, 64,
                                       runs.ptr, 6, base.ptr, None))
    got = runs.get(api.RUN, 6)
    assert [(int(r["q"]), int(r["count"]), int(r["first"])) for r in got[:3]] == [(0, 2, 0), (1, 2, 2), (2, 1, 16)]
    assert (got["count"][3:] == 0).all()  # entries of skipped items read as empty
    out = pairs.get(np.int32, 128).reshape(64, 2)
    assert set(map(tuple, out[0:2].tolist())) == {(1, 0), (1, 1)}
    assert set(map(tuple, out[2:4].tolist())) == {(2, 0), (3, 0)}
    assert tuple(out[16]) == (4, 1)
    qcount, ptotal = DevBuf(4 * nqueries).zero(), DevBuf(8).zero()
    N.check_hip(h.gat_run_counts(sp, ditems.ptr, total.ptr, 5, 6, runs.ptr, 6, nqueries, qcount.ptr, ptotal.ptr, None))
    counts = qcount.get(np.uint32, nqueries)
    assert counts[:3].tolist() == [2, 2, 1] and not counts[3:].any()
    assert ptotal.get(np.uint64, 1)[0] == 5
    # Argument errors come back as a status, not a crash.
    assert h.gat_run_counts(sp, ditems.ptr, total.ptr, 5, 6, runs.ptr, 6, nqueries, None, None, None) == 2
    assert h.gat_leaf_refine_runs(sp, None, total.ptr, 5, dq.ptr, 6, dboxes.ptr, nqueries, pairs.ptr, 64, runs.ptr,
                                  6, base.ptr, None) == 2


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 3 * 4096 + 5, 200_000])
def test_scan_u64_over_tiles(gpu, n):
    """The exclusive scan the pipeline's regions and lists come from: tiles
    of 4,096 values, each tile's carry summed from the tile totals before it
    in the apply pass; the total from the first workgroup (zero for n = 0)."""
    h = N.hip()
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 1 << 40, max(n, 1), dtype=np.uint64)[:n]
    din = DevBuf.of(vals) if n else DevBuf(8)
    out, total = DevBuf(8 * max(n, 1)).zero(), DevBuf.of(np.array([12345], np.uint64))
    tiles = DevBuf(8 * int(h.gat_scan_tiles(n))).zero()
    N.check_hip(h.gat_scan_u64(din.ptr, n, out.ptr, total.ptr, tiles.ptr, None))
    want = np.zeros(n + 1, np.uint64)  # integer throughout: sums pass 2^53
    want[1:] = np.cumsum(vals, dtype=np.uint64)
    assert (out.get(np.uint64, n) == want[:n]).all()
    assert total.get(np.uint64, 1)[0] == want[n]


def test_fill_spans(gpu):
    """gat_fill_u32_spans clears up to four spans in one launch (empty spans,
    with or without a pointer, included) and nothing beside them."""
    h = N.hip()
    lens = [5000, 0, 1, 100_003]
    bufs = [DevBuf.of(np.full(m + 2, 7, np.uint32)) for m in lens]
    # Each span starts one word into its buffer, so the words on both sides stay 7.
    ptrs = (ctypes.c_void_p * 4)(*[b.ptr.value + 4 if m else None for b, m in zip(bufs, lens)])
    n = (ctypes.c_int64 * 4)(*lens)
    N.check_hip(h.gat_fill_u32_spans(4, ptrs, n, 0xABCD, None))
    for b, m in zip(bufs, lens):
        got = b.get(np.uint32, m + 2)
        assert got[0] == 7 and got[-1] == 7 and (got[1:-1] == 0xABCD).all()
    assert h.gat_fill_u32_spans(5, ptrs, n, 0, None) == 2
