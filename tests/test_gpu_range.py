"""GPU parity of the range path: every index's GPU answer against the oracle.

Bar: bit-exact.  The GT-quadtree answer must equal the oracle's restatement of
Grid::rangeQueryBatch (Grid.cpp:341-405) point for point; FSG and STIG must
equal inclusive brute-force containment (their reference has no CPU path)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from gat_amd import GatError, _native as N, api, workloads as W

from tests.devbuf import DevBuf

pytestmark = pytest.mark.gpu


def _sets(per_query):
    return [set(map(tuple, np.asarray(r).tolist())) for r in per_query]


def _check_pairs(res, want_sets, rects):
    got = res.per_query()
    counts = res.counts()
    assert len(got) == len(rects)
    for q, (g, w) in enumerate(zip(got, want_sets)):
        assert counts[q] == len(g)
        s = set(map(tuple, g.tolist()))
        assert len(s) == len(g), f"query {q}: duplicate pairs"
        assert s == w, f"query {q}: {len(s - w)} extra, {len(w - s)} missing"


@pytest.fixture(scope="module")
def cfg1(gpu):
    db = W.CFG1.db()
    rects = W.CFG1.rects()
    return db, rects


def test_cfg1_grid_pairs_and_bitmap(cfg1):
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db)
    want = og.range(rects)
    res = g.range_pairs(rects)
    _check_pairs(res, _sets(want), rects)
    assert res.candidates >= res.pairs_total > 0
    counts, dig = og.range(rects, "digest")
    assert (res.digests() == dig).all()
    assert (g.range_bitmap(rects) == og.range(rects, "bitmap")).all()


def test_cfg1_fsg_and_stig(cfg1):
    db, rects = cfg1
    brute = O.brute_range(db, rects)
    f = api.FsgIndex(W.UNIT_BOX, W.CFG1.fsg_cell).build(db)
    _check_pairs(f.range_pairs(rects), _sets(brute), rects)
    s = api.StigIndex(W.CFG1.leaf).build(db)
    _check_pairs(s.range_pairs(rects), _sets(brute), rects)
    bm = O.brute_range(db, rects, "bitmap")
    assert (f.range_bitmap(rects) == bm).all()
    assert (s.range_bitmap(rects) == bm).all()


def _edge_rects(db, cell, n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, int(round(1 / cell)), size=(n, 2)).astype(np.float64)
    span = rng.integers(1, 5, size=(n, 2))
    on_cells = np.stack([k[:, 0] * cell, 1 - (k[:, 1] + span[:, 1]) * cell, (k[:, 0] + span[:, 0]) * cell,
                         1 - k[:, 1] * cell], 1)
    i = rng.integers(0, db.npoints, n)
    j = rng.integers(0, db.npoints, n)
    on_points = np.stack([np.minimum(db.lon[i], db.lon[j]), np.minimum(db.lat[i], db.lat[j]),
                          np.maximum(db.lon[i], db.lon[j]), np.maximum(db.lat[i], db.lat[j])], 1)
    point_boxes = np.stack([db.lon[i], db.lat[i], db.lon[i], db.lat[i]], 1)  # zero-area rectangles
    outside = np.array([[1.5, 1.5, 2.0, 2.0], [-1, -1, -0.5, -0.5], [0.2, 1.2, 0.4, 1.3]])
    whole = np.array([[0, 0, 1, 1], [-1, -1, 2, 2]])
    return np.concatenate([on_cells, on_points, point_boxes, outside, whole]).astype(np.float32)


@pytest.mark.parametrize("cell,leaf", [(1 / 16, 8), (0.01, 256), (0.001, 32)])
def test_edge_rectangles_all_indexes(gpu, cell, leaf):
    db = api.synth_walks(7, api.synth_lengths(7, 3000, "uniform", 1, 80, 1, 80), step_sigma=3e-3)
    rects = _edge_rects(db, cell, 120, int(1 / cell))
    og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    brute = _sets(O.brute_range(db, rects))
    g = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db)
    _check_pairs(g.range_pairs(rects), _sets(og.range(rects)), rects)
    f = api.FsgIndex(W.UNIT_BOX, cell).build(db)
    _check_pairs(f.range_pairs(rects), brute, rects)
    s = api.StigIndex(leaf).build(db)
    _check_pairs(s.range_pairs(rects), brute, rects)


def test_dense_cells_are_chunked(gpu):
    # Every point of a trajectory in one FSG cell: runs far longer than one task.
    lengths = np.full(400, 2048, np.int32)
    db = api.synth_walks(3, lengths, step_sigma=1e-7)
    rects = np.array([[0, 0, 1, 1], [0.25, 0.25, 0.75, 0.75], [0.1, 0.1, 0.6, 0.3]], np.float32)
    brute = _sets(O.brute_range(db, rects))
    f = api.FsgIndex(W.UNIT_BOX, 0.5).build(db)
    res = f.range_pairs(rects)
    assert res.tasks >= db.npoints // 1024
    _check_pairs(res, brute, rects)
    g = api.GridIndex(W.UNIT_BOX, 0.5, 256).build(db)
    _check_pairs(g.range_pairs(rects), _sets(O.Grid(W.UNIT_BOX, 0.5, 256).add(db).range(rects)), rects)


def test_empty_batches(gpu):
    db = W.CFG1.db(50)
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    res = g.range_pairs(np.zeros((0, 4), np.float32))
    assert res.pairs_total == 0 and res.counts().shape == (0,)
    far = np.array([[5, 5, 6, 6]], np.float32)
    res = g.range_pairs(far)
    assert res.pairs_total == 0 and res.candidates == 0
    assert g.range_bitmap(far).sum() == 0


def test_repeated_batches_reuse_buffers(gpu, cfg1):
    db, rects = cfg1
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    og = O.Grid(W.UNIT_BOX, 0.01, 256).add(db)
    for r in (rects, rects[:7], rects[::-1].copy(), rects[:1]):
        res = g.range_pairs(r)
        c, d = og.range(r, "digest")
        assert (res.counts() == c).all() and (res.digests() == d).all()


@pytest.mark.parametrize("packed", [True, False])
def test_refine_abi_skips_malformed_tasks(gpu, packed):
    """The C ABI on its own: hand-built tasks, including ones that point past
    the store or at a missing query, which the kernel must skip; both id
    layouts of gat_point_store."""
    h = N.hip()
    xy = np.array([[0.1, 0.1], [0.2, 0.2], [0.3, 0.3], [0.9, 0.9]], np.float32)
    tid = np.array([1, 1, 2, 3], np.uint32)
    pid = np.array([0, 1, 0, 0], np.uint16)
    boxes = np.array([[0, 0, 0.25, 0.25], [0.25, 0.25, 1, 1]], np.float32)
    tasks = np.array([[0, 4, 0], [0, 4, 1], [2, 5, 0], [0, 2, 7]], np.uint32)  # last two malformed
    dxy, db_, dt = DevBuf.of(xy), DevBuf.of(boxes), DevBuf.of(tasks)
    if packed:
        did, dpid = DevBuf.of((tid << 3) | pid.astype(np.uint32)), None
        store = N.PointStore(dxy.ptr, did.ptr, None, 4, 3)
    else:
        did, dpid = DevBuf.of(tid), DevBuf.of(pid)
        store = N.PointStore(dxy.ptr, did.ptr, dpid.ptr, 4, 0)
    sp = ctypes.byref(store)
    qbase = DevBuf.of(np.array([0, 4], np.uint64))
    qcount = DevBuf(8).zero()
    pairs = DevBuf(8 * 8).zero()
    N.check_hip(h.gat_range_refine_pairs(sp, dt.ptr, 4, db_.ptr, 2, qbase.ptr, qcount.ptr, pairs.ptr, 8, None))
    cnt = qcount.get(np.uint32, 2)
    out = pairs.get(np.int32, 16).reshape(8, 2)
    assert cnt.tolist() == [2, 2]
    assert set(map(tuple, out[:2].tolist())) == {(1, 0), (1, 1)}
    assert set(map(tuple, out[4:6].tolist())) == {(2, 0), (3, 0)}
    bitmap = DevBuf(2 * 4).zero()
    N.check_hip(h.gat_range_refine_bitmap(sp, dt.ptr, 4, db_.ptr, 2, bitmap.ptr, 4, None))
    assert bitmap.get(np.uint8, 8).reshape(2, 4).tolist() == [[0, 1, 0, 0], [0, 0, 1, 1]]
    # The reference-signature handler: synchronous, host result.
    host = np.full(8, 7, np.uint8)
    N.check_hip(h.gatRangeQueryTestHandler(sp, dt.ptr, 4, db_.ptr, api._ptr(host), 4, 2, None))
    assert host.reshape(2, 4).tolist() == [[0, 1, 0, 0], [0, 0, 1, 1]]
    # Argument errors come back as a status, not a crash.
    assert h.gat_range_refine_pairs(sp, None, 4, db_.ptr, 2, qbase.ptr, qcount.ptr, pairs.ptr, 8, None) == 2
    assert h.gat_range_refine_pairs(None, dt.ptr, 4, db_.ptr, 2, qbase.ptr, qcount.ptr, pairs.ptr, 8, None) == 2


def test_system_test_entry_points(cfg1):
    """The reference's SystemTest range entry points (SystemTest.cpp:24-102,
    251-384, 386-445) run the plain GPU path of each index: every rectangle's
    result pairs equal the oracle's as sets (GT-quadtree and MortonGrid: the
    Grid restatement; STIG and FSG: brute force), with the oracle's per-query
    counts and digests, over the same candidates as the index's own plain
    call; a MortonGrid answers rangeQueryTest too (it is a Grid)."""
    db, rects = cfg1
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    want_sets, want_dig = _sets(og.range(rects)), og.range(rects, "digest")
    brute_sets, brute_dig = _sets(O.brute_range(db, rects)), O.brute_range(db, rects, "digest")
    cases = [("grid", api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf), want_sets, want_dig),
             ("mgrid", api.MortonGridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf), want_sets, want_dig),
             ("stig", api.StigIndex(W.CFG1.leaf), brute_sets, brute_dig),
             ("fsg", api.FsgIndex(W.UNIT_BOX, W.CFG1.grid_cell), brute_sets, brute_dig)]
    for which, ix, sets, (counts, dig) in cases:
        ix.build(db)
        cand, got = ix.system_range_test(which, rects)
        assert got == sum(len(w) for w in sets), which
        res = ix.system_range_result()
        assert (res.counts() == counts).all() and (res.digests() == dig).all(), which
        _check_pairs(res, sets, rects)
        assert cand == ix.range_pairs(rects).candidates, which
    mg = cases[1][1]
    assert mg.system_range_test("grid", rects)[1] == sum(len(w) for w in want_sets)
    assert (mg.system_range_result().digests() == want_dig[1]).all()
    with pytest.raises(GatError):
        cases[0][1].system_range_test("mgrid", rects)
