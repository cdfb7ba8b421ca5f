"""GPU parity of the batched EDR forms: the persistent wave-systolic kernel, the
all-mismatch table, the candidate filter and the pruned top-k, all against the
oracle's restatement of Grid::SimilarityExecuter (Grid.cpp:1222-1318) and
brute-force top-k by (distance, id).  Bar: bit-exact."""
import ctypes

import numpy as np
import pytest

import oracle as O
from gat_amd import _native as N, api, workloads as W

from tests.devbuf import DevBuf

pytestmark = pytest.mark.gpu
EPS = 5e-4
THREADS = 16


def _brute(db, queries, eps=EPS):
    return np.stack([O.edr_batch(queries.xy(i + 1), db, eps, nthreads=THREADS) for i in range(queries.ntraj)])


def test_mismatch_table_matches_oracle(gpu):
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(W.CFG1.db(20))
    assert (g.mismatch_table() == O.ftable(api.MAX_TRAJ_LEN)).all()


def test_batched_all_cfg1(gpu):
    db = W.CFG1.db()
    g = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db)
    q = api.synth_queries(W.CFG1.seed, np.array([128, 128], np.int32), db, copy_mode=2)
    dist, st = g.edr_batched_all(q, EPS)
    assert st["pairs"] == 2 * db.ntraj
    assert (dist == _brute(db, q)).all()


def _varied_db(seed, lengths):
    rng = np.random.default_rng(seed)
    L = np.array(lengths, np.int32)
    db = api.synth_walks(seed, L, step_sigma=2e-4)
    # Pull a third of the trajectories onto one spot so many pairs match.
    for t in range(0, db.ntraj, 3):
        b, e = db.offsets[t], db.offsets[t + 1]
        db.lon[b:e] = np.float32(0.5) + (db.lon[b:e] - db.lon[b]) * np.float32(0.2)
        db.lat[b:e] = np.float32(0.5) + (db.lat[b:e] - db.lat[b]) * np.float32(0.2)
    return db


@pytest.mark.parametrize("maxlen", [512, 2048])
def test_batched_all_lengths(gpu, maxlen):
    base = [1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 255, 256, 257, 511, 512]
    if maxlen > 512:
        base += [513, 1000, 1023, 1024, 1025, 2047, 2048]
    rng = np.random.default_rng(maxlen)
    lengths = base + list(rng.integers(1, maxlen + 1, 60))
    db = _varied_db(maxlen, lengths)
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    qlen = np.array([1, 16, 17, 64, 129, min(maxlen, 700), maxlen] + list(rng.integers(1, maxlen + 1, 5)), np.int32)
    q = api.synth_queries(maxlen + 1, qlen, db, copy_mode=2)
    dist, st = g.edr_batched_all(q, EPS)
    want = _brute(db, q)
    bad = np.argwhere(dist != want)
    assert len(bad) == 0, [(int(a), int(b), int(qlen[a]), int(db.lengths[b]), int(dist[a, b]), int(want[a, b]))
                           for a, b in bad[:10]]
    # Another eps, to check the threshold rounding path.
    dist2, _ = g.edr_batched_all(q, 3e-3)
    assert (dist2 == _brute(db, q, 3e-3)).all()


def test_similarity_batched_topk(gpu):
    # A 4000-trajectory slice of the 200k-candidate workload (lognormal lengths,
    # median 128); queries U{16..1024}, every other one a noisy copy.
    w = W.CFG3
    db = w.db(4000)
    g = api.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db)
    qlen = api.synth_lengths(w.seed + 100, 12, "uniform", 16, 1024, 16, 1024)
    q = api.synth_queries(w.seed, qlen, db, copy_mode=2)
    ids, dists, st = g.similarity_batched(q, w.k, w.eps)
    assert st["pruned"] > 0 and st["pairs"] + st["pruned"] == q.ntraj * db.ntraj
    full = _brute(db, q, w.eps)
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    for i in range(q.ntraj):
        oi, od = O.topk(full[i], tid, w.k)
        assert (ids[i] == oi).all() and (dists[i] == od).all(), i
    # The plain per-query path gives the same answer.
    pi, pd = g.similarity(q, w.k, w.eps)
    assert (pi == ids).all() and (pd == dists).all()


def test_device_topk_ties_and_sizes(gpu):
    """The device top-k merge (k up to 1024) against the oracle where the cut
    falls inside large groups of equal distance (600 trajectories of five
    lengths, a third of them on one spot, so many survivors), for k = 1, a
    few, 50 and above the database size; k = 1025 takes the host merge and
    gives the same answers."""
    rng = np.random.default_rng(17)
    db = _varied_db(17, rng.choice([20, 40, 60, 64, 100], 600).astype(np.int32))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(18, np.array([20, 64, 100, 300], np.int32), db, copy_mode=2)
    full = _brute(db, q)
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    answers = {}
    for k in (1, 7, 50, 1024, 1025):
        ids, dists, st = g.similarity_batched(q, k, EPS)
        assert st["pairs"] > 0
        for i in range(q.ntraj):
            oi, od = O.topk(full[i], tid, k)
            n = len(oi)
            assert n == min(k, db.ntraj)
            assert (ids[i, :n] == oi).all() and (dists[i, :n] == od).all(), (k, i)
            assert (ids[i, n:] == 0).all() and (dists[i, n:] == -1).all(), (k, i)
        answers[k] = (ids[:, :db.ntraj], dists[:, :db.ntraj])
    assert (answers[1024][0] == answers[1025][0]).all() and (answers[1024][1] == answers[1025][1]).all()


def test_similarity_batched_cfg1(gpu):
    db = W.CFG1.db()
    g = api.GridIndex(W.UNIT_BOX, W.CFG1.grid_cell, W.CFG1.leaf).build(db)
    q = api.synth_queries(W.CFG1.seed, np.array([128, 128], np.int32), db, copy_mode=2)
    ids, dists, _ = g.similarity_batched(q, W.CFG1.k, EPS)
    full = _brute(db, q)
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    for i in range(2):
        oi, od = O.topk(full[i], tid, W.CFG1.k)
        assert (ids[i] == oi).all() and (dists[i] == od).all()


def test_filter_boundary_is_exact(gpu):
    """A candidate box exactly u (the float threshold) beyond the query box is
    dropped, one float closer is kept; the dropped pair's points cannot match."""
    h = N.hip()
    u = np.float32(api.eps_threshold(EPS))
    below = np.nextafter(u, np.float32(0))
    # The query box ends at x = 0, so a candidate starting at x has difference x exactly.
    q = np.array([[-0.25, 0.25, 0.0, 0.5]], np.float32)
    cb = np.array([[u, 0.3, u + np.float32(0.1), 0.4], [below, 0.3, below + np.float32(0.1), 0.4]], np.float32)
    dq, dc = DevBuf.of(q), DevBuf.of(cb)
    out = DevBuf(8 * 4)
    cnt = DevBuf(8)
    N.check_hip(h.gat_edr_filter(dq.ptr, 1, dc.ptr, 2, EPS, out.ptr, 4, cnt.ptr, None))
    n = int(cnt.get(np.uint64, 1)[0])
    pairs = out.get(np.int32, 2 * n).reshape(-1, 2)
    assert n == 1 and pairs.tolist() == [[0, 1]]
    # The boxes' nearest points: the dropped pair does not match, the kept one does.
    assert O.edr(np.array([[0.0, 0.35]], np.float32), np.array([[u, 0.35]], np.float32), EPS) == 1
    assert O.edr(np.array([[0.0, 0.35]], np.float32), np.array([[below, 0.35]], np.float32), EPS) == 0


def test_filter_many_queries_matches_the_box_rule(gpu):
    """The filter over 37 queries (blocks of 16 queries and a partial one)
    against 1,500 candidate boxes keeps exactly the pairs the float32 box rule
    keeps, each once (output order is not defined)."""
    h = N.hip()
    rng = np.random.default_rng(11)
    u = np.float32(api.eps_threshold(EPS))

    def boxes(n, side):
        lo = rng.random((n, 2), dtype=np.float32) * np.float32(0.05)
        hi = lo + rng.random((n, 2), dtype=np.float32) * np.float32(side)
        return np.ascontiguousarray(np.concatenate([lo, hi], 1))

    qb, cb = boxes(37, 0.01), boxes(1500, 0.002)
    keep = ~(((cb[None, :, 0] - qb[:, None, 2]) >= u) | ((qb[:, None, 0] - cb[None, :, 2]) >= u) |
             ((cb[None, :, 1] - qb[:, None, 3]) >= u) | ((qb[:, None, 1] - cb[None, :, 3]) >= u))
    want = {(int(a), int(b)) for a, b in np.argwhere(keep)}
    assert 0 < len(want) < keep.size
    dq, dc = DevBuf.of(qb), DevBuf.of(cb)
    cap = keep.size
    out = DevBuf(8 * cap)
    cnt = DevBuf(8)
    N.check_hip(h.gat_edr_filter(dq.ptr, 37, dc.ptr, 1500, EPS, out.ptr, cap, cnt.ptr, None))
    n = int(cnt.get(np.uint64, 1)[0])
    got = [tuple(p) for p in out.get(np.int32, 2 * n).reshape(-1, 2).tolist()]
    assert n == len(want) and len(set(got)) == n and set(got) == want


class _CandGrid(ctypes.Structure):
    _fields_ = [("cells_per_side", ctypes.c_int32), ("x0", ctypes.c_float), ("y0", ctypes.c_float),
                ("inv_x", ctypes.c_float), ("inv_y", ctypes.c_float), ("cell_start", ctypes.c_void_p),
                ("entry_id", ctypes.c_void_p), ("entry_box", ctypes.c_void_p), ("always", ctypes.c_void_p),
                ("nalways", ctypes.c_int32)]


def _grid_filter(qb, cb, grid, eps=EPS):
    h = N.hip()
    bufs = [DevBuf.of(grid[k]) for k in ("start", "id", "box", "always")]
    s = _CandGrid(grid["g"], grid["x0"], grid["y0"], grid["invx"], grid["invy"], *[b.ptr.value for b in bufs],
                  grid["nalways"])
    dq, dc = DevBuf.of(np.ascontiguousarray(qb)), DevBuf.of(np.ascontiguousarray(cb))
    cap = len(qb) * len(cb)
    out, cnt = DevBuf(8 * cap), DevBuf.of(np.array([12345], np.uint64))
    N.check_hip(h.gat_edr_filter_grid(dq.ptr, len(qb), dc.ptr, len(cb), ctypes.byref(s), eps, out.ptr, cap, cnt.ptr,
                                      None))
    n = int(cnt.get(np.uint64, 1)[0])
    return [tuple(p) for p in out.get(np.int32, 2 * n).reshape(-1, 2).tolist()]


@pytest.mark.parametrize("g", [1, 7, 64, 313])
def test_grid_filter_matches_the_box_rule(gpu, g):
    """gat_edr_filter_grid over a candidate grid of g x g cells keeps exactly
    the pairs the float32 box rule keeps, each once — with NaN, inverted and
    wide candidate boxes (tested by every query) and queries whose boxes are
    NaN or inverted (tested against every candidate), and a candidate a
    float-difference of exactly u from a query (dropped) and one float closer
    (kept) — and the same set as the every-pair filter."""
    from tests import cand_grid as CG
    u = np.float32(api.eps_threshold(EPS))
    qb, cb = CG.cases(g + 100, 70, 3000)
    qb[0] = np.array([0.3, 0.3, 0.31, 0.31], np.float32)
    edge = np.float32(qb[0, 2] + u)
    while edge - qb[0, 2] < u:
        edge = np.nextafter(edge, np.float32(1))
    cb[0] = np.array([edge, 0.3, edge + np.float32(0.001), 0.31], np.float32)
    cb[1] = np.array([np.nextafter(edge, np.float32(0)), 0.3, edge + np.float32(0.001), 0.31], np.float32)
    want = {(i, int(j)) for i in range(len(qb)) for j in np.nonzero(CG.keep(qb[i], cb, u))[0]}
    got = _grid_filter(qb, cb, CG.build(cb, g))
    assert len(got) == len(set(got)) and set(got) == want
    assert (0, 1) in want and (0, 0) not in want
    h = N.hip()
    dq, dc = DevBuf.of(qb), DevBuf.of(cb)
    out, cnt = DevBuf(8 * len(qb) * len(cb)), DevBuf(8)
    N.check_hip(h.gat_edr_filter(dq.ptr, len(qb), dc.ptr, len(cb), EPS, out.ptr, len(qb) * len(cb), cnt.ptr, None))
    n = int(cnt.get(np.uint64, 1)[0])
    assert {tuple(p) for p in out.get(np.int32, 2 * n).reshape(-1, 2).tolist()} == want


def test_topk_survivors_equal_the_box_rule(gpu):
    """The top-k batch's filter (the index's candidate grid) passes exactly as
    many pairs as the every-pair box rule on the trajectories' boxes."""
    db = _varied_db(91, np.array(list(np.random.default_rng(91).integers(2, 300, 800)), np.int32))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(92, np.array([5, 40, 128, 300, 64, 17], np.int32), db, copy_mode=2)
    _, _, st = g.similarity_batched(q, 10, EPS)

    def boxes(ts):
        return np.array([[ts.xy(i + 1)[:, 0].min(), ts.xy(i + 1)[:, 1].min(), ts.xy(i + 1)[:, 0].max(),
                          ts.xy(i + 1)[:, 1].max()] for i in range(ts.ntraj)], np.float32)

    from tests import cand_grid as CG
    u = np.float32(api.eps_threshold(EPS))
    qb, cb = boxes(q), boxes(db)
    want = sum(int(CG.keep(qb[i], cb, u).sum()) for i in range(len(qb)))
    assert st["pairs"] == want and st["pairs"] + st["pruned"] == q.ntraj * db.ntraj


def test_autotuned_edr_gives_the_same_distances(gpu):
    db = _varied_db(77, [1, 17, 64, 300, 511, 513, 700, 1024, 1025, 2048] + list(range(20, 1200, 37)))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(78, np.array([33, 600, 1500], np.int32), db, copy_mode=2)
    want = _brute(db, q)
    tuned = g.tune_edr(q, EPS, reps=1)
    assert tuned["split"] in (256, 512, 1024, 2048) and tuned["evaluations"] == 21
    assert tuned["waves_per_long_pair"] in (1, 2, 4, 8)
    dist, _ = g.edr_batched_all(q, EPS)
    assert (dist == want).all()
    for split in (256, 2048):
        g.set_edr_params(split, 6, 24)
        dist, _ = g.edr_batched_all(q, EPS)
        assert (dist == want).all(), split


@pytest.mark.parametrize("split", [512, 1])
def test_waves_per_long_pair(gpu, split):
    """Long pairs of a one-round batch over 1, 2, 4 or 8 waves in lockstep
    (edr_pipe_kernel): bit-exact against the oracle on rows fewer than one
    strip per wave, rows not a multiple of 64, the longest rows and columns,
    equal lengths (orientation rule); split 1 sends every pair to the
    long-pair launches.  Every long pair over several waves, and the default
    split between pairs over several waves and pairs on one."""
    lengths = [1, 2, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 511, 700, 1000, 1023, 1024, 1025, 1500, 2047,
               2048]
    db = _varied_db(91, lengths + [600, 900, 1300, 1800])
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(92, np.array([1, 64, 65, 300, 1024, 1500, 2048], np.int32), db, copy_mode=2)
    want = _brute(db, q)
    pairs = np.array([(a, b) for a in range(q.ntraj) for b in range(db.ntraj)], np.int32)
    g.set_edr_params(split, 0, 0)
    for waves, every in ((1, False), (2, True), (4, True), (8, True), (2, False), (4, False), (8, False)):
        g.set_edr_waves(waves, every_long_pair=every)
        dist, st = g.edr_batched_pairs(q, pairs, EPS)
        got = dist.reshape(q.ntraj, db.ntraj)
        bad = np.argwhere(got != want)
        assert len(bad) == 0, (waves, every, [(int(a), int(b), int(q.lengths[a]), int(db.lengths[b]),
                                               int(got[a, b]), int(want[a, b])) for a, b in bad[:8]])
        assert st["pairs"] == len(pairs)
    # A shuffled subset: pair order and duplicates do not matter.
    rng = np.random.default_rng(5)
    sub = pairs[rng.integers(0, len(pairs), 97)]
    g.set_edr_waves(4)
    dist, _ = g.edr_batched_pairs(q, sub, EPS)
    assert (dist == want[sub[:, 0], sub[:, 1]]).all()


@pytest.mark.parametrize("k", [10, 1100])
def test_topk_rejects_overlong_database_trajectory(gpu, k):
    """A database trajectory longer than the EDR kernels take (2049 points)
    is rejected before any work, by both top-k paths and for both merges
    (device merge for k <= 1024, host merge above): never a distance of -1
    ranked first, never a batch failing with an unclear status."""
    lengths = np.array([64, 2049, 100, 300], np.int32)
    db = api.synth_walks(7, lengths)
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(8, np.array([64, 128], np.int32), db, copy_mode=2)
    with pytest.raises(RuntimeError, match="2049 points"):
        g.similarity_batched(q, k, EPS)
    with pytest.raises(RuntimeError, match="2049 points"):
        g.similarity(q, k, EPS)


def test_topk_device_chain_matches_host_built_path(gpu):
    """The top-k batch chained on the device (gat_edr_jobs groups the filter's
    survivors into buckets and jobs, the DP launches read their job windows
    on the device, gat_edr_survivor_lists builds the per-query lists) gives
    the oracle's answers, as the host-built path does, under every wave
    setting of the long pairs (the device job builder applies the same
    one-round rules), on lengths that fill all four buckets."""
    rng = np.random.default_rng(23)
    lengths = list(rng.choice([16, 40, 100, 300, 700, 1100, 1600, 2048], 500))
    db = _varied_db(23, np.array(lengths, np.int32))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    for qlen in ([64, 900], [16, 100, 700, 1024, 1500, 2048, 300, 40]):
        q = api.synth_queries(24 + len(qlen), np.array(qlen, np.int32), db, copy_mode=2)
        full = _brute(db, q)
        for waves, every in ((1, False), (4, False), (4, True), (8, False)):
            g.set_edr_waves(waves, every_long_pair=every)
            got = {}
            for chain in (True, False):
                g.set_topk_device_chain(chain)
                ids, dists, st = g.similarity_batched(q, 25, EPS)
                assert st["pairs"] + st["pruned"] == q.ntraj * db.ntraj
                for i in range(q.ntraj):
                    oi, od = O.topk(full[i], tid, 25)
                    assert (ids[i] == oi).all() and (dists[i] == od).all(), (qlen, waves, every, chain, i)
                got[chain] = st["pairs"]
            assert got[True] == got[False]
    g.set_topk_device_chain(True)
    g.set_edr_waves(4)


def test_topk_device_chain_falls_back_for_long_lists(gpu):
    """Survivors past the chain's first capacity (1,024) grow it and run the
    batch again; then a query with more survivors than the device's per-query
    sort holds (2,600 trajectories on one spot, all within eps of the query)
    sends the batch to the host-built lists, with the oracle's answers."""
    db = api.synth_walks(31, np.full(2600, 20, np.int32), step_sigma=1e-5)
    db.lon[:] = np.float32(0.5) + (db.lon - db.lon.mean()) * np.float32(1e-3)
    db.lat[:] = np.float32(0.5) + (db.lat - db.lat.mean()) * np.float32(1e-3)
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(32, np.array([20, 12], np.int32), db, copy_mode=2)
    ids, dists, st = g.similarity_batched(q, 10, 0.01)
    assert st["pairs"] > 2048  # above the chain's first capacity and one query's sortable list
    full = _brute(db, q, 0.01)
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    for i in range(q.ntraj):
        oi, od = O.topk(full[i], tid, 10)
        assert (ids[i] == oi).all() and (dists[i] == od).all(), i


def test_device_jobs_large_batches_match_host_built_path(gpu):
    """Batches above 131,072 pairs build their job list over the whole device
    (lengths, bin totals, one-workgroup scan, per-workgroup placement, jobs)
    instead of in one workgroup: the all-pairs, given-pairs and top-k forms
    give the host-built path's distances, and sampled rows the oracle's."""
    rng = np.random.default_rng(41)
    lengths = list(rng.integers(1, 400, 2400)) + [513, 1024, 1500, 2048]
    db = _varied_db(41, np.array(lengths, np.int32))
    g = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db)
    q = api.synth_queries(42, np.array(list(rng.integers(1, 600, 58)) + [1024, 2048], np.int32), db, copy_mode=2)
    assert q.ntraj * db.ntraj > 131072
    dist, pairs_dist = {}, {}
    pairs = np.stack([rng.integers(0, q.ntraj, 140000), rng.integers(0, db.ntraj, 140000)], 1).astype(np.int32)
    for chain in (True, False):
        g.set_topk_device_chain(chain)
        dist[chain], st = g.edr_batched_all(q, EPS)
        assert st["pairs"] == q.ntraj * db.ntraj
        pairs_dist[chain], _ = g.edr_batched_pairs(q, pairs, EPS)
    assert (dist[True] == dist[False]).all()
    assert (pairs_dist[True] == pairs_dist[False]).all()
    assert (pairs_dist[True] == dist[True][pairs[:, 0], pairs[:, 1]]).all()
    for i in (0, 31, q.ntraj - 1):
        assert (dist[True][i] == O.edr_batch(q.xy(i + 1), db, EPS, nthreads=THREADS)).all(), i
    # Top-k with a chain capacity above the bound (64 per query, 2,100 queries).
    small = api.synth_walks(43, rng.integers(8, 80, 300).astype(np.int32), step_sigma=2e-4)
    gs = api.GridIndex(W.UNIT_BOX, 0.01, 256).build(small)
    qs = api.synth_queries(44, rng.integers(8, 64, 2100).astype(np.int32), small, copy_mode=2)
    got = {}
    for chain in (True, False):
        gs.set_topk_device_chain(chain)
        got[chain] = gs.similarity_batched(qs, 10, EPS)
    assert (got[True][0] == got[False][0]).all() and (got[True][1] == got[False][1]).all()
    assert got[True][2]["pairs"] == got[False][2]["pairs"]
    g.set_topk_device_chain(True)
    gs.set_topk_device_chain(True)


@pytest.mark.parametrize("capacity", [1000, 140_000])
def test_edr_jobs_overflow_empties_the_windows(gpu, capacity):
    """gat_edr_jobs with more pairs than its capacity, in both forms (one
    workgroup up to 131,072 pairs of capacity, the device-wide launches
    above): nothing is placed, every bucket window is empty and status bit 1
    is set; with the count at the capacity every pair lands in one window."""
    h = N.hip()
    nq, nc = 50, 4000
    rng = np.random.default_rng(capacity)
    qlen = rng.integers(1, 2049, nq).astype(np.int32)
    clen = rng.integers(1, 2049, nc).astype(np.int32)
    pairs = np.stack([rng.integers(0, nq, capacity), rng.integers(0, nc, capacity)], 1).astype(np.int32)
    dp, dq, dcl = DevBuf.of(pairs), DevBuf.of(qlen), DevBuf.of(clen)
    sorted_, origin = DevBuf(8 * capacity), DevBuf(4 * capacity)
    scratch = DevBuf(4 * (2 * capacity + 32768))
    jobs, windows, status, cells = DevBuf(8 * (capacity + 4)), DevBuf(64).zero(), DevBuf(4).zero(), DevBuf(8)
    for count, overflow in ((capacity + 1, True), (capacity, False)):
        cnt = DevBuf.of(np.array([count], np.uint64))
        N.check_hip(h.gat_edr_jobs(dp.ptr, cnt.ptr, capacity, dq.ptr, nq, dcl.ptr, nc, 256, 4, 0, sorted_.ptr,
                                   origin.ptr, scratch.ptr, jobs.ptr, windows.ptr, status.ptr, cells.ptr, None))
        win = windows.get(np.int32, 8)
        st = int(status.get(np.int32, 1)[0])
        if overflow:
            assert st & 1 and not win[1::2].any(), (count, win.tolist())
        else:
            assert st == 0
            got = np.sort(origin.get(np.int32, capacity))
            assert (got == np.arange(capacity)).all()  # every pair placed once
            n, m = np.minimum(qlen[pairs[:, 0]], clen[pairs[:, 1]]), np.maximum(qlen[pairs[:, 0]], clen[pairs[:, 1]])
            assert int(cells.get(np.uint64, 1)[0]) == int((n.astype(np.int64) * m).sum())
