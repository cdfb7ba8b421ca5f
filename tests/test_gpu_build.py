"""Device construction of the cell-ordered point store (gat_cell_store_build)
against the oracle's store, which restates Grid::addDatasetToGrid
(Grid.cpp:202-261: cells in Morton order, sub-trajectories in id order, points
in trajectory order), and against the host bucketing of Grid and FSG
(FSG.cpp:95-143).

Bar: bit-exact store (coordinates compared as bit patterns, trajectory and
point ids), identical cell offsets and tree, and range answers equal to the
oracle's on the device-built index."""
import numpy as np
import pytest

import oracle as O
from gat_amd import GatError, api, workloads as W

from tests.test_gpu_range import _check_pairs, _edge_rects, _sets

pytestmark = pytest.mark.gpu


def _ids(ids, pid, bits):
    if bits > 0:
        return (ids >> np.uint32(bits)).astype(np.int64), (ids & np.uint32((1 << bits) - 1)).astype(np.int64)
    return ids.astype(np.int64), pid.astype(np.int64)


def _on_lines(cell, n, seed):
    """Trajectories whose points sit exactly on cell lines, on the grid's
    edges and corners, and in between: the cases where the float formula's
    truncation decides the cell."""
    rng = np.random.default_rng(seed)
    k = int(round(1 / cell))
    lines = np.arange(k + 1, dtype=np.float32) * np.float32(cell)
    lines = np.concatenate([lines, np.float32([0.0, 1.0, 0.5, np.nextafter(np.float32(1), np.float32(0))])])
    lens = rng.integers(1, 40, size=n)
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    m = int(offsets[-1])
    pick = rng.random(m) < 0.6
    lon = np.where(pick, rng.choice(lines, m), rng.random(m)).astype(np.float32)
    lat = np.where(rng.random(m) < 0.6, rng.choice(lines, m), rng.random(m)).astype(np.float32)
    return api.TrajSet(offsets, lon, lat)


def _dbs():
    walks = api.synth_walks(7, api.synth_lengths(7, 3000, "uniform", 1, 80, 1, 80), step_sigma=3e-3)
    return [("cfg1", W.CFG1.db(4000)), ("walks", walks), ("lines16", _on_lines(1 / 16, 2000, 3)),
            ("lines100", _on_lines(0.01, 2000, 4))]


@pytest.mark.parametrize("cell,leaf", [(1 / 16, 8), (0.01, 64), (1 / 1024, 256)])
def test_grid_device_store_equals_oracle(gpu, cell, leaf):
    for name, db in _dbs():
        og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
        x, y, t, p = og.store()
        dev = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, device=True)
        xy, ids, pid, bits = dev.fetch_store()
        assert len(xy) == len(x) == db.npoints, name
        assert (xy[:, 0].view(np.uint32) == x.view(np.uint32)).all(), name
        assert (xy[:, 1].view(np.uint32) == y.view(np.uint32)).all(), name
        tt, pp = _ids(ids, pid, bits)
        assert (tt == t).all() and (pp == p).all(), name
        host = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db)
        hxy, hids, hpid, hbits = host.fetch_store()
        assert hbits == bits and (hids == ids).all() and (hxy.view(np.uint32) == xy.view(np.uint32)).all(), name
        assert (dev.cell_starts() == host.cell_starts()).all(), name
        dn, dd = dev.tree_nodes()
        hn, hd = host.tree_nodes()
        assert dd == hd and dn.tobytes() == hn.tobytes(), name


@pytest.mark.parametrize("cell", [1 / 16, 0.01, 1 / 1024])
def test_fsg_device_store_equals_host(gpu, cell):
    for name, db in _dbs():
        dev = api.FsgIndex(W.UNIT_BOX, cell).build(db, device=True)
        host = api.FsgIndex(W.UNIT_BOX, cell).build(db)
        xy, ids, pid, bits = dev.fetch_store()
        hxy, hids, hpid, hbits = host.fetch_store()
        assert len(xy) == db.npoints and bits == hbits, name
        assert (xy.view(np.uint32) == hxy.view(np.uint32)).all() and (ids == hids).all(), name
        assert (dev.cell_starts() == host.cell_starts()).all(), name
        # Store order: cells ascending (oracle cell function), then (tid, pid).
        of = O.FSG(W.UNIT_BOX, cell)
        cells = np.array([of.cell_of(a, b) for a, b in zip(xy[:, 0], xy[:, 1])], np.int64)
        tt, pp = _ids(ids, pid, bits)
        key = (cells << 42) | (tt << 12) | pp
        assert (np.diff(key) > 0).all(), name


def test_device_built_indexes_answer_like_the_oracle(gpu):
    db = W.CFG1.db()
    rects = W.CFG1.rects()
    og = O.Grid(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).add(db)
    g = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db, device=True)
    _check_pairs(g.range_pairs(rects), _sets(og.range(rects)), rects)
    _check_pairs(g.range_batched(rects), _sets(og.range(rects)), rects)
    edge = _edge_rects(db, W.CFG1.grid_cell, 60, 21)
    counts, dig = og.range(edge, "digest")
    res = g.range_batched(edge)
    assert (res.counts() == counts).all() and (res.digests() == dig).all()
    f = api.FsgIndex(W.UNIT_BOX, W.CFG1.fsg_cell).build(db, device=True)
    _check_pairs(f.range_pairs(rects), _sets(O.brute_range(db, rects)), rects)
    # The trajectory store for EDR is the same as after a host build.
    q = api.synth_queries(5, np.array([40, 128], np.int32), db, copy_mode=2)
    h = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db)
    assert (g.edr_all(q.xy(1), 5e-4) == h.edr_all(q.xy(1), 5e-4)).all()


def test_unpacked_ids(gpu):
    """More than 2^21 trajectories and 2048-point ones: trajectory and point ids
    do not share 32 bits, so point ids go to the 16-bit array."""
    n = (1 << 21) + 2
    lens = np.ones(n, np.int64)
    lens[-1] = 2048
    lens[n // 2] = 1500
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    rng = np.random.default_rng(11)
    m = int(offsets[-1])
    db = api.TrajSet(offsets, rng.random(m).astype(np.float32), rng.random(m).astype(np.float32))
    dev = api.GridIndex(W.UNIT_BOX, 1 / 256, 256).build(db, device=True)
    host = api.GridIndex(W.UNIT_BOX, 1 / 256, 256).build(db)
    xy, ids, pid, bits = dev.fetch_store()
    hxy, hids, hpid, hbits = host.fetch_store()
    assert bits == 0 and hbits == 0
    assert (ids == hids).all() and (pid == hpid).all() and (xy.view(np.uint32) == hxy.view(np.uint32)).all()
    assert pid.max() == 2047


def test_points_off_the_grid_are_rejected(gpu):
    db = W.CFG1.db(50)
    lon = db.lon.copy()
    lon[17] = 1.5  # beyond the 128-cell Morton grid of [0, 1] at 0.01
    bad = api.TrajSet(db.offsets, lon, db.lat)
    with pytest.raises(GatError, match="point 17 .* outside the grid"):
        api.GridIndex(W.UNIT_BOX, 0.01, 64).build(bad, device=True)
    lon[17] = 1.02  # inside the Morton grid's padding, outside FSG's 101 columns
    bad = api.TrajSet(db.offsets, lon, db.lat)
    with pytest.raises(GatError, match="point 17 .* outside the grid"):
        api.FsgIndex(W.UNIT_BOX, 0.01).build(bad, device=True)


def test_empty_dataset(gpu):
    db = api.TrajSet(np.zeros(1, np.int64), np.zeros(0, np.float32), np.zeros(0, np.float32))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 64).build(db, device=True)
    assert (g.cell_starts() == 0).all()
    res = g.range_batched(np.array([[0, 0, 1, 1]], np.float32))
    assert res.pairs_total == 0


# ------------------------------------------------------------- STIG kd-split

def _subtree_ranges(nodes):
    """Store range [lo, hi) of every node of a breadth-first tree (children
    follow their parent)."""
    lo = nodes["start"].astype(np.int64).copy()
    hi = lo + nodes["count"].astype(np.int64)
    for i in range(len(nodes) - 1, -1, -1):
        c = int(nodes["first_child"][i])
        if c >= 0:
            lo[i] = min(lo[c], lo[c + 1])
            hi[i] = max(hi[c], hi[c + 1])
    return lo, hi


@pytest.mark.parametrize("block", [8, 256])
def test_stig_device_build(gpu, block):
    """STIG's kd-split on the device (gat_kd_store_build): the host build's
    tree shape (the split positions depend only on node sizes, STIG.cpp:71-168
    corrected), every point once, every leaf box the exact box of its points,
    every split half-plane holding its whole subtree (what the search's
    inclusive tests rely on, STIG.cpp:179-252), deterministic, and answers
    equal to inclusive brute force on both query forms."""
    for name, db in _dbs():
        _check_stig(db, block, name)


def _check_stig(db, block, name, rects=None):
    host = api.StigIndex(block).build(db)
    dev = api.StigIndex(block).build(db, device=True)
    hn, hd = host.tree_nodes()
    dn, dd = dev.tree_nodes()
    assert hd == dd and len(hn) == len(dn), name
    for f in ("first_child", "start", "count", "flags"):
        assert (hn[f] == dn[f]).all(), (name, f)
    xy, ids, pid, bits = dev.fetch_store()
    t, p = _ids(ids, pid, bits)
    tids = np.repeat(np.arange(1, db.ntraj + 1), np.diff(db.offsets))
    pids = np.arange(db.npoints) - np.repeat(db.offsets[:-1], np.diff(db.offsets))
    order = np.lexsort((p, t))
    assert (t[order] == tids).all() and (p[order] == pids).all(), name
    # The coordinates travel with their ids.
    src = db.offsets[t - 1] + p
    assert (xy[:, 0] == db.lon[src]).all() and (xy[:, 1] == db.lat[src]).all(), name
    lo, hi = _subtree_ranges(dn)
    for i, nd in enumerate(dn):
        x, y = xy[lo[i]:hi[i], 0], xy[lo[i]:hi[i], 1]
        b = nd["box"]
        if nd["first_child"] < 0:
            assert list(b) == [x.min(), y.min(), x.max(), y.max()], (name, i)
        else:
            assert (x >= b[0]).all() and (x <= b[2]).all() and (y >= b[1]).all() and (y <= b[3]).all(), (name, i)
    again = api.StigIndex(block).build(db, device=True).fetch_store()
    assert (again[0].view(np.uint32) == xy.view(np.uint32)).all() and (again[1] == ids).all(), name
    if rects is None:
        rects = np.concatenate([api.synth_rects(5, 60, 1e-3, 0.3, True), _edge_rects(db, 1 / 16, 30, 6)])
    counts, dig = O.brute_range(db, rects, "digest")
    # A result reads its index's last batch: check each before the next call.
    for run in (lambda: dev.range_pairs(rects), lambda: dev.set_batch_params(32, 0).range_batched(rects)):
        res = run()
        assert (res.counts() == counts).all() and (res.digests() == dig).all(), name


def test_stig_device_build_root_leaf_and_cfg1(gpu):
    """A dataset no larger than a block is one leaf (no split depth); cfg1's
    2.7M points give the same answers as the host-built index."""
    small = W.CFG1.db(3)
    dev = api.StigIndex(1 << 20).build(small, device=True)
    nodes, depth = dev.tree_nodes()
    assert len(nodes) == 1 and depth == 0 and nodes["count"][0] == small.npoints
    db, rects = W.CFG1.db(), W.CFG1.rects()
    host = api.StigIndex(W.CFG1.leaf).build(db)
    dev = api.StigIndex(W.CFG1.leaf).build(db, device=True)
    a = host.range_batched(rects)
    ca, da = a.counts(), a.digests()
    b = dev.range_batched(rects)
    assert (b.counts() == ca).all() and (b.digests() == da).all()
    _check_pairs(dev.range_pairs(rects), _sets(O.brute_range(db, rects)), rects)


def _sized(n, seed, spot):
    """n points in trajectories of 1 to 12 points: spread over the unit
    square, or all inside one 1/1024 cell (every sort key equal)."""
    rng = np.random.default_rng(seed)
    lens = []
    while sum(lens) < n:
        lens.append(int(min(rng.integers(1, 13), n - sum(lens))))
    offsets = np.zeros(len(lens) + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    if spot:
        lon = (np.float32(0.3001) + rng.random(n).astype(np.float32) * np.float32(5e-4)).astype(np.float32)
        lat = (np.float32(0.6001) + rng.random(n).astype(np.float32) * np.float32(5e-4)).astype(np.float32)
    else:
        lon, lat = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    return api.TrajSet(offsets, lon, lat)


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 12289])
def test_device_sort_tile_edges(gpu, n):
    """The builds' radix sort works in tiles of 4,096 pairs: sizes just below,
    at and above a tile, a single pair, and a dataset whose points all share
    one cell (one digit everywhere, so every pass's stability across tiles
    decides the store order) — the cell store equals the oracle's (20 key bits,
    three passes; 16 cells, one pass) and the kd store passes the kd checks
    (64-bit keys)."""
    for spot in (False, True):
        db = _sized(n, n, spot)
        name = f"n={n} spot={spot}"
        for cell, leaf in ((1 / 1024, 256), (1 / 16, 8)):
            og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
            x, y, t, p = og.store()
            xy, ids, pid, bits = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, device=True).fetch_store()
            assert (xy[:, 0].view(np.uint32) == x.view(np.uint32)).all(), name
            assert (xy[:, 1].view(np.uint32) == y.view(np.uint32)).all(), name
            tt, pp = _ids(ids, pid, bits)
            assert (tt == t).all() and (pp == p).all(), name
        _check_stig(db, 8, name, api.synth_rects(n + 1, 40, 1e-4, 0.3, True))
