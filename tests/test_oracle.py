"""Pins the oracle (CPU restatement of the reference) before anything is checked
against it: against the reference's own compiled MBB.cpp, against closed forms
and hand-derived known answers of the EDR recurrence, against independent
restatements, and against brute force on small inputs.  The reference ships no
tests or golden data (its only check is the CHECK_CORRECT dump,
ConstDefine.h:8, SystemTest.cpp:79-89)."""
import itertools

import numpy as np
import pytest

import oracle as O
from gat_amd import api, workloads as W

EPS = 5e-4


# ------------------------------------------------------------- geometry
def _edge_boxes(rng, n):
    """Boxes on a coarse lattice so that edges coincide, touch and nest often."""
    v = rng.integers(0, 5, size=(n, 4)).astype(np.float32) * np.float32(0.25)
    b = np.stack([np.minimum(v[:, 0], v[:, 2]), np.minimum(v[:, 1], v[:, 3]),
                  np.maximum(v[:, 0], v[:, 2]), np.maximum(v[:, 1], v[:, 3])], 1)
    return b.astype(np.float32)


def test_pinbox_inclusive_edges():
    box = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
    for x, y in [(0.25, 0.5), (0.75, 1.0), (0.25, 1.0), (0.75, 0.5), (0.5, 0.75)]:
        assert O.pinbox(box, x, y) == 1
    for x, y in [(np.nextafter(np.float32(0.25), np.float32(0)), 0.75), (0.5, np.nextafter(np.float32(1.0), np.float32(2)))]:
        assert O.pinbox(box, x, y) == 0


def test_geometry_matches_reference_mbb():
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (the reference sources are absent here)")
    rng = np.random.default_rng(7)
    boxes = _edge_boxes(rng, 400)
    c = O._p
    for a, b in itertools.product(boxes[:60], boxes[:60]):
        assert O.box_intersect(a, b) == ref.ref_box_intersect(c(a), c(b)), (a, b)
    pts = rng.integers(0, 5, size=(200, 2)).astype(np.float32) * np.float32(0.25)
    for box in boxes[:50]:
        for x, y in pts:
            assert O.pinbox(box, x, y) == ref.ref_pinbox(c(box), float(x), float(y))


def test_box_intersect_closed_form():
    # Touching boxes are disjoint; containment codes depend on argument order.
    a = np.array([0, 0, 1, 1], np.float32)
    assert O.box_intersect(a, np.array([1, 0, 2, 1], np.float32)) == 0   # touch on x
    assert O.box_intersect(a, np.array([0, 1, 1, 2], np.float32)) == 0   # touch on y
    assert O.box_intersect(a, np.array([0.5, 0.5, 2, 2], np.float32)) == 1
    inner = np.array([0.25, 0.25, 0.75, 0.75], np.float32)
    assert O.box_intersect(a, inner) == 2
    assert O.box_intersect(inner, a) == 3


def test_morton_known_answers():
    assert O.morton(0, 0) == 0
    assert O.morton(1, 0) == 1
    assert O.morton(0, 1) == 2
    assert O.morton(1, 1) == 3
    assert O.morton(3, 5) == 39
    # Closed form: x's bits in even positions, y's in odd positions.
    rng = np.random.default_rng(1)
    for x, y in rng.integers(0, 4096, size=(500, 2)):
        want = 0
        for b in range(13):
            want |= ((int(x) >> b) & 1) << (2 * b)
            want |= ((int(y) >> b) & 1) << (2 * b + 1)
        assert O.morton(int(x), int(y)) == want


# ------------------------------------------------------------- EDR
def pts(*coords):
    return np.array(coords, np.float32).reshape(-1, 2)


A, B, C, D = (0.1, 0.1), (0.3, 0.3), (0.5, 0.5), (0.7, 0.7)


def test_edr_known_answers():
    assert O.edr(pts(A), pts(A), EPS) == 0
    assert O.edr(pts(A), pts(B), EPS) == 1
    assert O.edr(pts(A), pts(A, B), EPS) == 1           # D[1][2] = L = 1
    assert O.edr(pts(A), pts(B, C), EPS) == 3           # M == L < U returns U = 3 (textbook: 2)
    assert O.edr_textbook(pts(A), pts(B, C), EPS) == 2
    assert O.edr(pts(A, B), pts(A, B), EPS) == 0
    assert O.edr(pts(), pts(A, B, C), EPS) == 3         # D[0][m] = m


def allmiss(n, m):
    q = np.stack([np.linspace(0.0, 0.1, n), np.zeros(n)], 1).astype(np.float32)
    c = np.stack([np.linspace(0.5, 0.6, m), np.ones(m) * 0.5], 1).astype(np.float32)
    return q, c


def test_edr_all_mismatch_closed_form_one_row():
    # With one row and no match: D[1][c] = c when c is odd, c + 1 when c is even
    # (by induction on c: L ties M exactly when D[1][c-1] = c - 1).
    for m in range(1, 40):
        q, c = allmiss(1, m)
        assert O.edr(q, c, EPS) == m + (1 - m % 2)


def test_edr_all_mismatch_table_values():
    want = {(1, k): v for k, v in zip(range(1, 10), [1, 3, 3, 5, 5, 7, 7, 9, 9])}
    want.update({(2, k): v for k, v in zip(range(2, 11), [2, 3, 6, 6, 6, 7, 10, 10, 10])})
    want.update({(3, k): v for k, v in zip(range(3, 12), [3, 7, 7, 7, 7, 11, 11, 11, 11])})
    want.update({(4, k): v for k, v in zip(range(4, 13), [4, 5, 6, 7, 12, 12, 12, 12, 12])})
    want.update({(16, k): k for k in range(16, 29)})
    want[(20, 32)] = 52
    F = O.ftable(40)
    for (n, m), v in want.items():
        q, c = allmiss(n, m)
        assert O.edr(q, c, EPS) == v, (n, m)
        assert F[n, m] == v, (n, m)


def test_edr_with_matches_can_exceed_the_all_mismatch_value():
    """The cell rule ((M == L) ? U : min(M, L, U), Grid.cpp:1297-1300) is not a
    minimum, so a pair with matching points can end above the distance of
    the same two lengths with no match at all: two pairs of the 20k-trajectory
    workload against a 511-point query at eps 0.02.  The device top-k merge
    keeps survivors at their own distance for this reason."""
    db = W.CFG1.db()
    q = api.synth_queries(5, np.array([511], np.int32), db, copy_mode=2)
    F = O.ftable(600)
    for cand, want in ((4580, 512), (8919, 529)):
        c = db.xy(cand + 1)
        assert O.edr(q.xy(1), c, 0.02) == want
        n, m = sorted((511, len(c)))
        assert F[n, m] == 511 < want


def test_edr_eps_boundary():
    below = np.float32(4.9999997e-4)
    at = np.float32(5e-4)  # 5.0000002e-4 > 5e-4 (double)
    assert float(at) > 5e-4 and float(below) < 5e-4
    assert O.edr(pts((0, 0)), pts((below, 0)), EPS) == 0
    assert O.edr(pts((0, 0)), pts((at, 0)), EPS) == 1
    assert O.edr(pts((0, 0)), pts((0, -below)), EPS) == 0
    assert O.edr(pts((0, at)), pts((0, 0)), EPS) == 1
    # The difference is rounded to float before the comparison: 0.25 + below
    # rounds to a float whose distance from 0.25 is not `below`.
    x = np.float32(0.25) + below
    d = np.float32(x - np.float32(0.25))
    assert O.edr(pts((0.25, 0.25)), pts((x, 0.25)), EPS) == (0 if float(d) < 5e-4 else 1)


def _levenshtein(a, b):
    prev = list(range(len(b) + 1))
    for i, x in enumerate(a, 1):
        cur = [i]
        for j, y in enumerate(b, 1):
            cur.append(min(prev[j - 1] + (x != y), prev[j] + 1, cur[j - 1] + 1))
        prev = cur
    return prev[-1]


def test_edr_textbook_is_levenshtein_on_symbols():
    # Points far apart act as symbols; textbook EDR is then Levenshtein distance.
    rng = np.random.default_rng(3)
    sym = np.array([[0.1 * k, 0.05 * k] for k in range(6)], np.float32)
    for _ in range(200):
        a = rng.integers(0, 6, rng.integers(1, 9))
        b = rng.integers(0, 6, rng.integers(1, 9))
        assert O.edr_textbook(sym[a], sym[b], EPS) == _levenshtein(list(a), list(b))


def _walk_pair(rng, n, m, noisy):
    q = (rng.random(2) * 0.5 + 0.25) + np.cumsum(rng.normal(0, 1e-4, (n, 2)), 0)
    if noisy:
        idx = np.minimum(np.arange(m) * n // m, n - 1)
        c = q[idx] + rng.normal(0, 2e-4, (m, 2))
    else:
        c = (rng.random(2) * 0.5 + 0.25) + np.cumsum(rng.normal(0, 1e-4, (m, 2)), 0)
    return q.astype(np.float32), c.astype(np.float32)


def test_edr_restatements_agree_and_bound():
    rng = np.random.default_rng(11)
    orient_diff = 0
    for i in range(300):
        n, m = rng.integers(1, 70, 2)
        q, c = _walk_pair(rng, int(n), int(m), noisy=i % 3 != 0)
        d = O.edr(q, c, EPS)
        assert d == O.edr_antidiag(q, c, EPS)
        assert d >= O.edr_textbook(q, c, EPS) >= abs(int(n) - int(m))
        assert d <= int(n) + int(m)
        if n == m and O.edr(c, q, EPS) != d:
            orient_diff += 1
    # Zero-match pairs take the F-table value.
    F = O.ftable(80)
    for _ in range(50):
        n, m = (int(x) for x in rng.integers(1, 80, 2))
        q, c = allmiss(n, m)
        assert O.edr(q, c, EPS) == F[min(n, m), max(n, m)]


def test_edr_orientation_rule():
    # Rows are the shorter trajectory, the query on equal lengths
    # (Grid.cpp:1234-1247); the rule is asymmetric, so the choice is visible.
    rng = np.random.default_rng(5)
    found = False
    for _ in range(2000):
        n = int(rng.integers(2, 8))
        q, c = _walk_pair(rng, n, n, noisy=True)
        if O.edr(q, c, EPS) != O.edr(c, q, EPS):
            found = True
            break
    assert found
    # Unequal lengths: the shorter is the rows whichever argument it is.
    for _ in range(100):
        q, c = _walk_pair(rng, 5, 9, noisy=True)
        assert O.edr(q, c, EPS) == O.edr(c, q, EPS)


def test_topk_order():
    rng = np.random.default_rng(2)
    d = rng.integers(0, 20, 500).astype(np.int32)
    ids = np.arange(1, 501, dtype=np.int32)
    oi, od = O.topk(d, ids, 37)
    order = np.lexsort((ids, d))[:37]
    assert (oi == ids[order]).all() and (od == d[order]).all()


# ------------------------------------------------------------- range
def _small_db(seed, ntraj=300):
    lengths = api.synth_lengths(seed, ntraj, "uniform", 1, 60, 1, 60)
    return api.synth_walks(seed, lengths, step_sigma=2e-2)


def _sets(res):
    return [set(map(tuple, r.tolist())) for r in res]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_indexes_against_brute_force(seed):
    db = _small_db(seed)
    rng = np.random.default_rng(seed)
    rects = api.synth_rects(seed, 60, 0.01, 0.5, True)
    # Rectangles whose edges sit exactly on cell edges and on data points.
    cell = np.float32(1 / 16)
    lat = rng.integers(0, 16, size=(30, 4)).astype(np.float32) * cell
    rects = np.concatenate([rects, np.stack([np.minimum(lat[:, 0], lat[:, 2]), np.minimum(lat[:, 1], lat[:, 3]),
                                             np.maximum(lat[:, 0], lat[:, 2]) + cell,
                                             np.maximum(lat[:, 1], lat[:, 3]) + cell], 1)])
    k = rng.integers(0, db.npoints, 20)
    rects = np.concatenate([rects, np.stack([db.lon[k], db.lat[k], db.lon[k] + 0.05, db.lat[k] + 0.05], 1)])
    rects = rects.astype(np.float32)
    brute = _sets(O.brute_range(db, rects))
    grid = _sets(O.Grid(W.UNIT_BOX, float(cell), 8).add(db).range(rects))
    fsg = _sets(O.FSG(W.UNIT_BOX, float(cell)).add(db).range(rects))
    stig = _sets(O.STIG().add(db, 8).range(rects))
    for q in range(len(rects)):
        assert fsg[q] == brute[q]
        assert stig[q] == brute[q]
        assert grid[q] <= brute[q]
        # Any point the quadtree misses lies on the rectangle's boundary.
        for tid, pid in brute[q] - grid[q]:
            i = db.offsets[tid - 1] + pid
            x, y = db.lon[i], db.lat[i]
            r = rects[q]
            assert x in (r[0], r[2]) or y in (r[1], r[3])


def test_oracle_range_modes_agree():
    db = _small_db(4)
    rects = api.synth_rects(4, 40, 0.02, 0.4, True)
    g = O.Grid(W.UNIT_BOX, 1 / 16, 8).add(db)
    pairs = g.range(rects)
    bitmap = g.range(rects, "bitmap")
    counts, dig = g.range(rects, "digest")
    for q, p in enumerate(pairs):
        assert counts[q] == len(p)
        assert set(np.flatnonzero(bitmap[q])) == set(p[:, 0].tolist())
        assert int(dig[q]) == sum(O.pair_digest(int(t), int(i)) for t, i in p) % (1 << 64)
    # Multi-threaded slicing (Grid.cpp:407-426) gives the same answer.
    assert all((a == b).all() for a, b in zip(g.range(rects, nthreads=5), pairs))


# ------------------------------------------------------------- MortonGrid

def test_mgrid_bitmap_index_known_answers():
    """findIdxInBitmap (MortonGrid.cpp:8-13): levels above hold (4^l - 1) / 3
    nodes.  Level 10 starts at 349,525, the size of the reference's flag set
    (MortonGrid.h:17), which a 1024 x 1024 grid's single-cell leaves overrun."""
    want = {(0, 0): 0, (1, 0): 1, (1, 3): 4, (2, 0): 5, (2, 15): 20, (3, 0): 21, (3, 63): 84, (10, 0): 349525}
    for (level, nid), idx in want.items():
        assert O.MortonGrid.bitmap_index(level, nid) == idx


def test_mgrid_geometry_known_answers():
    """Range [0, 4]^2 with unit cells: 5 cells needed per axis, rounded to 8
    (MortonGrid.cpp:29-33), so 3 levels; node boxes by hand from their first
    and last cells (generateMBBfromNode, MortonGrid.cpp:263-268), rows counted
    down from the range's top edge y = 4 (the grid extends below y = 0)."""
    m = O.MortonGrid((0, 0, 4, 4), 1.0, 2)
    db = api.TrajSet(np.array([0, 1], np.int64), np.array([0.5], np.float32), np.array([3.5], np.float32))
    m.add(db)
    info = m.info()
    assert (info["axis"], info["cellnum"], info["total_level"]) == (8, 64, 3)
    # Cell (row i, col j) spans [j, j+1] x [4 - (i+1), 4 - i].
    assert list(m.node_box(1, 0)) == [0, 0, 4, 4]      # UL: cells 0..15 = rows 0..3, cols 0..3
    assert list(m.node_box(1, 1)) == [4, 0, 8, 4]      # UR: cols 4..7
    assert list(m.node_box(1, 2)) == [0, -4, 4, 0]     # DL: rows 4..7
    assert list(m.node_box(2, 5)) == [6, 2, 8, 4]      # cells 20..23: cols 6..7, rows 0..1
    assert list(m.node_box(3, 20)) == [6, 3, 7, 4]     # cell 20 = (col 6, row 0)


def test_mgrid_leaf_rule_known_answer():
    """Three points in one cell, maxPointsInNode 2 (MortonGrid.cpp:238-261):
    the cell's ancestors hold 3 >= 2 points and split down to the cell, a
    single cell is a leaf at the last level, and every empty sibling is a
    leaf where it appears: 3 + 3 + 4 leaves."""
    m = O.MortonGrid((0, 0, 4, 4), 1.0, 2)
    db = api.TrajSet(np.array([0, 3], np.int64), np.array([0.2, 0.5, 0.7], np.float32),
                     np.array([3.9, 3.5, 3.1], np.float32))
    m.add(db)
    assert m.info()["leaves"] == 10
    assert not m.is_leaf(0, 0) and not m.is_leaf(1, 0) and not m.is_leaf(2, 0)
    assert m.is_leaf(3, 0) and m.is_leaf(3, 1) and m.is_leaf(2, 1) and m.is_leaf(1, 3)
    # A rectangle over the point's cell only: the leaf (3, 0) with its 3 points.
    got = m.leaves((0.1, 3.1, 0.9, 3.9))
    assert [tuple(r) for r in got.tolist()] == [(3, 0, 0, 3)]
    # Touching the cell from outside selects nothing there (BoxIntersect is strict).
    assert all(tuple(r[:2]) != (3, 0) for r in m.leaves((1.0, 3.1, 1.5, 3.9)).tolist())


@pytest.mark.parametrize("cell,leaf", [(1 / 16, 8), (0.01, 64), (1 / 1024, 8), (0.25, 100000)])
def test_mgrid_matches_grid_restatement(cell, leaf):
    """MortonGrid and Grid are two restatements of two reference files
    (MortonGrid.cpp, Grid.cpp) with the same geometry, store and leaf rule, so
    they select the same leaves in the same order and answer alike; the last
    shape makes the root a leaf (taken untested by both)."""
    db = _small_db(7)
    rects = np.concatenate([api.synth_rects(7, 80, 1e-3, 0.5, True),
                            np.array([[0, 0, 1, 1], [0.5, 0.5, 0.5, 0.5], [-1, -1, -0.5, -0.5]], np.float32)])
    c = np.float32(cell)
    rng = np.random.default_rng(7)
    k = rng.integers(0, int(1 / cell), size=(40, 2)).astype(np.float32)
    rects = np.concatenate([rects, np.stack([k[:, 0] * c, 1 - (k[:, 1] + 2) * c, (k[:, 0] + 2) * c,
                                             1 - k[:, 1] * c], 1)]).astype(np.float32)
    g = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    m = O.MortonGrid(W.UNIT_BOX, cell, leaf).add(db)
    assert m.info()["leaves"] == g.info()["leaves"]
    for r in rects:
        assert [tuple(x) for x in m.leaves(r)[:, 2:].tolist()] == [tuple(x) for x in g.leaves(r).tolist()]
    gc, gd = g.range(rects, "digest")
    mc, md = m.range(rects, "digest")
    assert (gc == mc).all() and (gd == md).all()
    assert _sets(m.range(rects[:30])) == _sets(g.range(rects[:30]))


# ---- the reference's own CPU top-k (FVTable.cpp:18-207, Grid.cpp:1021-1220):
# a timed baseline, pinned here by hand-derived answers and by the cases the
# loop's rules fix.

def test_morton_x86_known_answers():
    """getIdxFromXY for the negative coordinates findNeighbor passes at the
    grid's edge (FVTable.cpp:67-73), as the reference's x86-64 build computes
    it: a negative coordinate's bit length is INT_MIN + 1 (int(log2(-1)) is
    the conversion of NaN), so only the other coordinate's length counts, and
    2 * (INT_MIN + 1) wraps to 2 when both are negative.  By hand:
    (-1, -1): 2 bits, x0 = 1, y0 = 1 -> 3; (-1, 5): 6 bits, x bits all 1 at
    positions 0, 2, 4 (21) plus y = 101b at 1, 3, 5 (2 + 32) -> 55;
    (0, -1): 2 bits, y0 = 1 -> 2; (2, -1): 4 bits, x = 10b at 0, 2 (4) plus
    y at 1, 3 (2 + 8) -> 14."""
    assert O.morton_x86(-1, -1) == 3
    assert O.morton_x86(-1, 5) == 55
    assert O.morton_x86(0, -1) == 2
    assert O.morton_x86(2, -1) == 14
    assert O.morton_x86(3, -1) == 15
    for x, y in itertools.product(range(0, 70, 3), range(0, 70, 5)):
        assert O.morton_x86(x, y) == O.morton(x, y)


def test_fv_neighbors_known_answers():
    """FVTable::findNeighbor (FVTable.cpp:53-75): decode the cell (even bits x,
    odd bits y), then the eight cells around it with x outermost.  Inside a
    64 x 64 virtual grid the cell (5, 5) gives the eight real neighbours; on a
    4 x 4 grid the corner cell 0 gives its three real neighbours 1, 2, 3 and
    five edge ids that happen to name them again, and cell 4 = (2, 0) names
    cells 3, 14 and 15 for its three off-grid neighbours (2, -1) -> 14 lies
    on the opposite edge."""
    m = O.morton
    want = [m(4, 4), m(4, 5), m(4, 6), m(5, 4), m(5, 6), m(6, 4), m(6, 5), m(6, 6)]
    assert O.fv_neighbors(m(5, 5), 4096).tolist() == want
    assert O.fv_neighbors(0, 16).tolist() == [3, 1, 3, 2, 2, 3, 1, 3]
    assert O.fv_neighbors(4, 16).tolist() == [3, 1, 3, 14, m(2, 1), 15, m(3, 0), m(3, 1)]


def test_fd_value_known_answers():
    """The frequency distance of FVTable::formPriorityQueue (FVTable.cpp:86-198)
    on hand-built vectors over the 64 x 64 virtual grid, a = cell (5, 5):"""
    m = O.morton
    a = m(5, 5)
    C = 4096
    # No overlap: all of t is positive, all of q negative.
    assert O.fd_value({a: 3}, {}, C) == 3
    assert O.fd_value({}, {a: 2}, C) == 2
    assert O.fd_value({a: 3, m(40, 40): 4}, {m(20, 20): 5}, C) == 7
    # Same cell: the difference.
    assert O.fd_value({a: 3}, {a: 1}, C) == 2
    # A neighbour cancels part of a positive cell.
    assert O.fd_value({a: 3}, {m(6, 5): 1}, C) == 2
    # Equal values cancel both; a larger negative neighbour keeps the rest.
    assert O.fd_value({a: 2}, {m(5, 6): 2}, C) == 0
    assert O.fd_value({a: 1}, {m(4, 5): 3}, C) == 2
    # The positive cell's value is read once before its neighbour loop
    # (FVTable.cpp:132): two neighbours of 2 each both cancel against the
    # original 5, leaving 5 - 2 = 3, not 1.
    assert O.fd_value({a: 5}, {m(4, 4): 2, m(6, 6): 2}, C) == 3
    # The negative pass (FVTable.cpp:160-187) acts where the positive pass did
    # not see the pair: on the 4 x 4 grid cell 14 does not list cell 4, but
    # cell 4 lists 14 as its off-grid neighbour (2, -1), so the negative cell 4
    # (1 point) cancels against the positive cell 14 (2 points): max(1, 0).
    assert O.fd_value({14: 2}, {4: 1}, 16) == 1
    # The other way round the positive pass sees it: 4 cancels into 14.
    assert O.fd_value({4: 1}, {14: 2}, 16) == 1


def _fd_db(rng, centres, length, spread=2e-5):
    lengths = np.full(len(centres), length, np.int32)
    db = api.TrajSet.from_lengths(lengths)
    for t, (x, y) in enumerate(centres):
        b, e = db.offsets[t], db.offsets[t + 1]
        db.lon[b:e] = np.float32(x) + rng.normal(0, spread, e - b).astype(np.float32)
        db.lat[b:e] = np.float32(y) + rng.normal(0, spread, e - b).astype(np.float32)
    return db


def _query(db, t):
    b, e = db.offsets[t], db.offsets[t + 1]
    return api.TrajSet(np.array([0, e - b], np.int64), db.lon[b:e].copy(), db.lat[b:e].copy())


def test_fd_topk_one_round_evaluates_everything():
    """A database of at most 80 trajectories is one round of KSIMILARITY = 80
    candidates (Grid.cpp:1046-1051): every distance is computed, so the heap
    of the k best (Grid.cpp:1082-1106) holds the exact top-k distances, popped
    worst first (Grid.cpp:1202-1209), and each id carries its own distance."""
    rng = np.random.default_rng(11)
    centres = rng.uniform(0.2, 0.8, size=(60, 2))
    centres[:20] = centres[0] + rng.normal(0, 3e-4, size=(20, 2))
    db = _fd_db(rng, centres, 24)
    q = _query(db, 0)
    fdk = O.FdTopk(W.UNIT_BOX, 0.01, db)
    ids, dist, st = fdk.topk(q, 10, 5e-4)
    assert st[0].tolist() == [1, 60]
    exact = O.edr_batch(q.xy(1), db, 5e-4)
    assert (dist[0] == exact[ids[0] - 1]).all()
    assert dist[0].tolist() == sorted(dist[0].tolist(), reverse=True)
    assert sorted(dist[0].tolist()) == sorted(exact.tolist())[:10]


def test_fd_topk_stop_rule():
    """The loop stops when the worst kept distance no longer exceeds the next
    frequency distance (Grid.cpp:1040).  80 near copies of the query (FD 0)
    come out first and fill the heap with small distances; the 80 far
    trajectories have FD = max(24, 24) = 24 > every kept distance: one round.
    With far trajectories twice the query's length instead, every FD is
    max(48, 24) = 48 and every distance F(24, 48) > 48 (the all-mismatch
    value), so the rule never stops the loop and the queue runs dry:
    ceil(200 / 80) = 3 rounds over all 200."""
    rng = np.random.default_rng(12)
    near = np.tile([[0.505, 0.505]], (80, 1))  # mid-cell: every copy in the query's one cell
    far = rng.uniform(0.05, 0.3, size=(80, 2))
    db = _fd_db(rng, np.concatenate([near, far]), 24, spread=1e-5)
    q = _query(db, 0)
    fdk = O.FdTopk(W.UNIT_BOX, 0.01, db)
    fd = fdk.fd(q.xy(1))
    assert fd[:80].max() == 0 and fd[80:].min() == 24
    ids, dist, st = fdk.topk(q, 10, 5e-4)
    assert st[0].tolist() == [1, 80]
    assert dist[0].max() < 24 and (ids[0] <= 80).all()
    F = O.ftable(64)
    assert F[24, 48] > 48
    far_db = _fd_db(rng, rng.uniform(0.05, 0.3, size=(200, 2)), 48, spread=1e-5)
    q2 = _fd_db(rng, np.array([[0.705, 0.705]]), 24, spread=1e-5)
    fdk2 = O.FdTopk(W.UNIT_BOX, 0.01, far_db)
    assert (fdk2.fd(q2.xy(1)) == 48).all()
    ids2, dist2, st2 = fdk2.topk(q2, 10, 5e-4)
    assert st2[0].tolist() == [3, 200]
    assert (dist2[0] == F[24, 48]).all()
