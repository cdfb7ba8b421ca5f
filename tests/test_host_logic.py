"""CPU tests of the product's host logic (libgat_host.so without a GPU):
geometry, cell numbering, index construction and candidate selection checked
against the oracle; the ABI surface of both libraries; the workload generator."""
import itertools
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from gat_amd import _native as N, api, workloads as W

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path) -> set[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(gath?_[a-z0-9_]+|gat[A-Z][A-Za-z0-9_]*)\s*\(", text))
    return {n for n in names if not n.startswith("gat_range_task") and not n.startswith("gat_edr_task")}


def test_libraries_export_every_declared_symbol():
    hip_decl = _declared(ROOT / "include" / "gat_hip.h")
    host_decl = _declared(ROOT / "include" / "gat_host.h")
    assert "gatRangeQueryTestHandler" in hip_decl and "gatEDRDistanceBatchHandler" in hip_decl
    hip_syms = N.exported_symbols(N.HIP_LIB)
    host_syms = N.exported_symbols(N.HOST_LIB)
    assert hip_decl <= hip_syms, hip_decl - hip_syms
    assert host_decl <= host_syms, host_decl - host_syms
    # Both libraries load and bind every declared function without a GPU.
    N.hip()
    N.host()
    assert set(N.HIP_FUNCS) == hip_decl
    assert set(N.HOST_FUNCS) == host_decl


def test_eps_threshold_is_smallest_float_at_or_above():
    for eps in [5e-4, 0.005, 1e-3, 0.1, 0.5, 1.0, 3e-7]:
        u = np.float32(api.eps_threshold(eps))
        assert float(u) >= eps
        assert float(np.nextafter(u, np.float32(0))) < eps


def test_geometry_matches_oracle():
    h = N.host()
    rng = np.random.default_rng(9)
    v = rng.integers(0, 5, size=(120, 4)).astype(np.float32) * np.float32(0.25)
    boxes = np.stack([np.minimum(v[:, 0], v[:, 2]), np.minimum(v[:, 1], v[:, 3]),
                      np.maximum(v[:, 0], v[:, 2]), np.maximum(v[:, 1], v[:, 3])], 1).astype(np.float32)
    boxes = np.concatenate([boxes, rng.random((60, 4)).astype(np.float32)])  # includes inverted boxes
    for a, b in itertools.product(boxes, boxes[:40]):
        assert h.gath_box_intersect(api._ptr(a), api._ptr(b)) == O.box_intersect(a, b)
    pts = rng.integers(0, 5, size=(100, 2)).astype(np.float32) * np.float32(0.25)
    for box in boxes[:60]:
        for x, y in pts:
            assert h.gath_pinbox(api._ptr(box), float(x), float(y)) == O.pinbox(box, x, y)


def test_morton_matches_oracle():
    h = N.host()
    for x in range(0, 300, 7):
        for y in range(0, 300, 11):
            assert h.gath_morton(x, y) == O.morton(x, y)
    rng = np.random.default_rng(4)
    for x, y in rng.integers(0, 32768, size=(2000, 2)):
        assert h.gath_morton(int(x), int(y)) == O.morton(int(x), int(y))


def _edge_points(cell, n, rng):
    """Points on and next to cell edges, where float assignment is delicate."""
    k = rng.integers(0, int(round(1 / cell)) + 1, size=(n, 2)).astype(np.float64)
    base = (k * cell).astype(np.float32)
    jitter = rng.integers(-2, 3, size=(n, 2))
    out = base.copy()
    for i in range(n):
        for d in range(2):
            for _ in range(abs(int(jitter[i, d]))):
                out[i, d] = np.nextafter(out[i, d], np.float32(np.inf if jitter[i, d] > 0 else -np.inf))
    return np.clip(out, 0, 1).astype(np.float32)


@pytest.mark.parametrize("cell", [0.01, 0.001, 1 / 16, 0.00025])
def test_cell_assignment_matches_oracle(cell):
    rng = np.random.default_rng(int(1 / cell))
    g = api.GridIndex(W.UNIT_BOX, cell, 256)
    og = O.Grid(W.UNIT_BOX, cell, 256)
    f = api.FsgIndex(W.UNIT_BOX, cell)
    of = O.FSG(W.UNIT_BOX, cell)
    pts = np.concatenate([_edge_points(cell, 3000, rng), rng.random((3000, 2)).astype(np.float32)])
    for x, y in pts:
        assert g.cell_of(x, y) == og.cell_of(x, y)
        assert f.cell_of(x, y) == of.cell_of(x, y)


def test_grid_geometry_matches_oracle():
    for cell in [0.01, 0.001, 0.00025, 0.3]:
        db = api.synth_walks(1, np.full(10, 5, np.int32))
        g = api.GridIndex(W.UNIT_BOX, cell, 256).build(db, upload=False).info()
        o = O.Grid(W.UNIT_BOX, cell, 256).add(db).info()
        assert g[3] == o["axis"] and g[2] == o["cellnum"] and g[6] == o["total_level"]


def _merged_runs(leaves):
    runs = []
    for s, c in leaves:
        if c == 0:
            continue
        if runs and runs[-1][0] + runs[-1][1] == s:
            runs[-1][1] += int(c)
        else:
            runs.append([int(s), int(c)])
    return runs


def _task_runs(tasks, q):
    runs = []
    for s, c, _ in tasks[tasks[:, 2] == q]:
        assert c <= 1024
        if runs and runs[-1][0] + runs[-1][1] == s:
            runs[-1][1] += int(c)
        else:
            runs.append([int(s), int(c)])
    return runs


@pytest.mark.parametrize("cell,leaf", [(0.01, 256), (1 / 16, 8), (0.001, 32)])
def test_grid_selection_matches_oracle_traversal(cell, leaf):
    db = W.CFG1.db(2000)
    og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    pg = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False)
    info = pg.info()
    oi = og.info()
    assert (info[4], info[5], info[0]) == (oi["nodes"], oi["leaves"], oi["points"])
    rects = np.concatenate([W.CFG1.rects(), api.synth_rects(9, 200, 1e-4, 0.5, True)])
    # Rectangles on cell edges (BoxIntersect drops leaves that only touch).
    rng = np.random.default_rng(5)
    k = rng.integers(0, int(1 / cell), size=(100, 2)).astype(np.float64)
    span = rng.integers(1, 4, size=(100, 2))
    edge = np.stack([k[:, 0] * cell, 1 - (k[:, 1] + span[:, 1]) * cell, (k[:, 0] + span[:, 0]) * cell,
                     1 - k[:, 1] * cell], 1).astype(np.float32)
    rects = np.concatenate([rects, edge]).astype(np.float32)
    tasks, cand = pg.select(rects)
    for q in range(len(rects)):
        want = _merged_runs(og.leaves(rects[q]))
        assert _task_runs(tasks, q) == want, q
        assert cand[q] == sum(c for _, c in want)


def test_grid_store_order_matches_oracle():
    """Selection runs index the product store; the oracle's store has the
    reference's order (cells in Morton order, sub-trajectories in id order), so
    identical runs mean identical candidate points."""
    db = W.CFG1.db(500)
    og = O.Grid(W.UNIT_BOX, 0.01, 64).add(db)
    x, y, t, p = og.store()
    # Every cell's block in the oracle store is contiguous and in (tid, pid) order.
    cells = np.array([og.cell_of(a, b) for a, b in zip(x, y)])
    assert (np.diff(cells) >= 0).all()
    same = np.diff(cells) == 0
    key = t.astype(np.int64) << 32 | p
    assert (np.diff(key)[same] > 0).all()


def test_fsg_and_stig_selection_cover_brute_force():
    db = W.CFG1.db(1500)
    rects = np.concatenate([W.CFG1.rects(), api.synth_rects(3, 100, 1e-4, 0.3, True)]).astype(np.float32)
    brute = O.brute_range(db, rects, "digest")[0]
    for ix in (api.FsgIndex(W.UNIT_BOX, 0.01).build(db, upload=False),
               api.StigIndex(64).build(db, upload=False)):
        tasks, cand = ix.select(rects)
        assert (cand >= brute).all()
        assert (tasks[:, 1] <= 1024).all()


def test_stig_leaf_count():
    db = W.CFG1.db(700)
    s = api.StigIndex(100).build(db, upload=False).info()
    assert s[0] == db.npoints
    assert s[5] >= db.npoints // 100


def test_generator_is_deterministic_and_shaped():
    a = W.CFG1.db(300)
    b = W.CFG1.db(300, threads=3)
    assert (a.offsets == b.offsets).all() and (a.lon == b.lon).all() and (a.lat == b.lat).all()
    L = a.lengths
    assert L.min() >= 16 and L.max() <= 256
    assert (a.lon >= 0).all() and (a.lon <= 1).all() and (a.lat >= 0).all() and (a.lat <= 1).all()
    steps = np.hypot(np.diff(a.lon), np.diff(a.lat))
    inside = np.ones(len(steps), bool)
    inside[a.offsets[1:-1] - 1] = False
    assert 0.5e-4 < np.median(steps[inside]) < 3e-4
    L2 = W.CFG2.lengths(100_000)
    assert 8 <= L2.min() and L2.max() <= 2048 and 60 <= np.median(L2) <= 68
    r = W.CFG2.rects(5000)
    side = r[:, 2] - r[:, 0]
    assert 0.9e-4 < side.min() and side.max() < 0.11
    assert abs(np.median(np.log10(side)) - (-2.5)) < 0.1


def test_host_rejects_points_outside_the_grid():
    db = api.synth_walks(1, np.array([3, 3], np.int32))
    db.lon[4] = 2.0
    with pytest.raises(N.GatError):
        api.GridIndex(W.UNIT_BOX, 0.01, 256).build(db, upload=False)
    with pytest.raises(N.GatError):
        api.FsgIndex(W.UNIT_BOX, 0.01).build(db, upload=False)


def _node_selected(n, box):
    """The batched form's node test (gat_tree_node flags), in float32."""
    if n["flags"] & api.NODE_UNTESTED:
        return True
    b = n["box"]
    if n["flags"] & api.NODE_INCLUSIVE:
        return bool(b[0] <= box[2] and b[2] >= box[0] and b[1] <= box[3] and b[3] >= box[1])
    return O.box_intersect(box, b) != 0


def _wave_select(nodes, box, nchild, stack_cap=4096, wave=64):
    """tree_select_kernel restated: the stack holds first-child indices of
    selected inner nodes; pop up to 64 from the top, test each popped node's
    children in lane order, record selected leaves, push the first-child index
    of each selected inner child (lane by lane).  Returns the selected leaf
    runs and the largest stack size reached."""
    sel = []
    root = nodes[0]
    stack = []
    if root["first_child"] < 0:
        if _node_selected(root, box) and root["count"] > 0:
            sel.append((int(root["start"]), int(root["count"])))
    elif _node_selected(root, box):
        stack.append(int(root["first_child"]))
    return sel, _walk(nodes, box, nchild, stack, sel, stack_cap, wave)


def _walk(nodes, box, nchild, stack, sel, stack_cap=4096, wave=64):
    """The walk of tree_select_kernel from the given stack until it is empty;
    appends selected leaf runs to sel and returns the largest stack size."""
    top_max = len(stack)
    while stack:
        m = min(len(stack), wave, (stack_cap - len(stack)) // (nchild - 1))
        assert m > 0
        batch = stack[len(stack) - m:]
        del stack[len(stack) - m:]
        pushes = []
        for fc in batch:
            for c in range(nchild):
                n = nodes[fc + c]
                if not _node_selected(n, box):
                    continue
                if n["first_child"] < 0:
                    if n["count"] > 0:
                        sel.append((int(n["start"]), int(n["count"])))
                else:
                    pushes.append(int(n["first_child"]))
        stack.extend(pushes)
        top_max = max(top_max, len(stack))
    return top_max


def _spread(v):
    """Morton interleave: v's bits to the even positions."""
    out = 0
    for b in range(16):
        out |= ((v >> b) & 1) << (2 * b)
    return out


def _table_select(nodes, table, level, geom, box, stack_cap=4096, wave=64):
    """tree_select_kernel with a start table restated: the candidate blocks
    within one block of the rectangle (float32 arithmetic as on the device),
    64 at a time; each owner taken at its first block inside the range and
    tested once; then the walk from the selected inner owners.  NaN and
    inverted rectangles walk from the root."""
    f32 = np.float32
    box = np.asarray(box, np.float32)
    if not (box[0] <= box[2] and box[1] <= box[3]):
        return _wave_select(nodes, box, 4, stack_cap, wave)
    x0, y1, block = (f32(v) for v in geom)
    nb = 1 << level
    inv = f32(1) / block

    def block_of(v, d):
        t = np.floor(f32(v)) + f32(d)
        return int(min(max(t, f32(0)), f32(nb - 1)))

    c0, c1 = block_of((box[0] - x0) * inv, -1), block_of((box[2] - x0) * inv, 1)
    r0, r1 = block_of((y1 - box[3]) * inv, -1), block_of((y1 - box[1]) * inv, 1)
    bw = c1 - c0 + 1
    ncand = bw * (r1 - r0 + 1)
    sel, top_max = [], 0
    for tb in range(0, ncand, wave):
        stack = []
        for b in range(tb, min(tb + wave, ncand)):
            r, c = r0 + b // bw, c0 + b % bw
            w = int(table[_spread(c) | (_spread(r) << 1)])
            o, lv = w & ((1 << 27) - 1), w >> 27
            span = 1 << (level - lv)
            if (max(r & ~(span - 1), r0), max(c & ~(span - 1), c0)) != (r, c):
                continue
            n = nodes[o]
            if not _node_selected(n, box):
                continue
            if n["first_child"] < 0:
                if n["count"] > 0:
                    sel.append((int(n["start"]), int(n["count"])))
            else:
                stack.append(int(n["first_child"]))
        top_max = max(top_max, _walk(nodes, box, 4, stack, sel, stack_cap, wave))
    return sel, top_max


def _runs(leaves):
    runs = []
    for s, c in sorted(leaves):
        if runs and runs[-1][0] + runs[-1][1] == s:
            runs[-1][1] += c
        else:
            runs.append([s, c])
    return runs


@pytest.mark.parametrize("cell,leaf", [(0.01, 256), (1 / 16, 8), (1 / 1024, 8)])
def test_grid_device_tree_selects_oracle_leaves(cell, leaf):
    """The breadth-first tree the batched form uploads selects exactly the
    oracle's leaves (Grid.cpp:428-446), and the traversal stays inside the
    stack bound the upload checks."""
    db = W.CFG1.db(1500)
    og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    nodes, depth = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False).tree_nodes()
    assert nodes[0]["flags"] == api.NODE_UNTESTED and (nodes[1:]["flags"] == 0).all()
    inner = nodes["first_child"] >= 0
    assert (nodes["first_child"][inner] > np.nonzero(inner)[0]).all()
    rng = np.random.default_rng(11)
    k = rng.integers(0, int(1 / cell), size=(20, 2)).astype(np.float64)
    edge = np.stack([k[:, 0] * cell, 1 - (k[:, 1] + 2) * cell, (k[:, 0] + 2) * cell, 1 - k[:, 1] * cell],
                    1).astype(np.float32)
    rects = np.concatenate([W.CFG1.rects()[:20], api.synth_rects(4, 20, 1e-4, 0.5, True), edge,
                            np.array([[0, 0, 1, 1], [-1, -1, 2, 2]], np.float32)]).astype(np.float32)
    for r in rects:
        sel, top = _wave_select(nodes, r, 4)
        want = [[int(s), int(c)] for s, c in og.leaves(r) if c > 0]
        assert _runs(sel) == _runs(want)
        assert top <= (depth + 1) * 4 * 64


@pytest.mark.parametrize("cell,leaf", [(1 / 1024, 8), (0.001, 16), (0.001, 64)])
def test_grid_start_table_selects_oracle_leaves(cell, leaf):
    """The walk from the start table (the owners of one level's blocks) selects
    exactly the leaves the walk from the root selects, which are the oracle's
    (Grid.cpp:428-446), on rectangles from points to the whole domain, edge-
    and block-aligned ones, and NaN and inverted ones (which start at the
    root); every block has one owner at or above the table's level."""
    db = W.CFG1.db(3000)
    og = O.Grid(W.UNIT_BOX, cell, leaf).add(db)
    g = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False)
    nodes, depth = g.tree_nodes()
    table, level, geom = g.start_table()
    assert level == min(7, depth - 2) >= 2 and len(table) == 4 ** level
    own, lv = table & ((1 << 27) - 1), table >> 27
    assert (lv <= level).all() and (own < len(nodes)).all()
    assert ((lv == level) | (nodes["first_child"][own] < 0)).all()
    rng = np.random.default_rng(29)
    nb = 1 << level
    k = rng.integers(0, nb, size=(30, 2)).astype(np.float32)
    bs, x0, y1 = geom[2], geom[0], geom[1]
    aligned = np.stack([x0 + k[:, 0] * bs, y1 - (k[:, 1] + 3) * bs, x0 + (k[:, 0] + 3) * bs, y1 - k[:, 1] * bs], 1)
    pts = rng.random((20, 2)).astype(np.float32)
    rects = np.concatenate([
        W.CFG1.rects()[:40], api.synth_rects(7, 60, 1e-4, 0.5, True), aligned,
        np.concatenate([pts, pts], 1),
        np.array([[0, 0, 1, 1], [-1, -1, 2, 2], [0.3, 0.3, 0.2, 0.4], [np.nan, 0, 1, 1],
                  [0.5, 0.5, 0.5, 0.5], [-np.inf, -np.inf, np.inf, np.inf]], np.float32)]).astype(np.float32)
    for r in rects:
        sel, top = _table_select(nodes, table, level, geom, r)
        root_sel, _ = _wave_select(nodes, r, 4)
        assert _runs(sel) == _runs(root_sel), r
        if np.isfinite(r).all():
            want = [[int(s), int(c)] for s, c in og.leaves(r) if c > 0]
            assert _runs(sel) == _runs(want), r
        assert top <= (depth + 1) * 4 * 64


def test_stig_device_tree_selects_host_leaves():
    """STIG's breadth-first tree (split half-planes on inner nodes, leaf boxes)
    selects the same points as the host kd search (STIG.cpp:179-252)."""
    db = W.CFG1.db(1200)
    ix = api.StigIndex(32).build(db, upload=False)
    nodes, depth = ix.tree_nodes()
    rects = np.concatenate([W.CFG1.rects()[:30], api.synth_rects(6, 30, 1e-4, 0.5, True),
                            np.array([[0, 0, 1, 1], [0.5, 0.5, 0.5, 0.5]], np.float32)]).astype(np.float32)
    tasks, _ = ix.select(rects)
    for q, r in enumerate(rects):
        sel, top = _wave_select(nodes, r, 2)
        t = tasks[tasks[:, 2] == q]
        assert _runs(sel) == _runs([(int(s), int(c)) for s, c, _ in t])
        assert top <= (depth + 1) * 2 * 64


def test_record_layouts_match_headers(tmp_path):
    """The numpy records the bindings pass through the C ABI have the field
    offsets and sizes the headers give their structs (compiled and printed)."""
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "gat_host.h"
#define F(T, m) printf("%s.%s %zu\n", #T, #m, offsetof(T, m))
int main(void) {
  F(gath_request, kind); F(gath_request, box); F(gath_request, qlen); F(gath_request, qoff);
  F(gath_request, arrival_ms); printf("gath_request %zu\n", sizeof(gath_request));
  F(gath_outcome, admit_ms); F(gath_outcome, done_ms); F(gath_outcome, batch); F(gath_outcome, count);
  F(gath_outcome, digest); printf("gath_outcome %zu\n", sizeof(gath_outcome));
  F(gat_tree_node, box); F(gat_tree_node, first_child); F(gat_tree_node, start); F(gat_tree_node, count);
  F(gat_tree_node, flags); printf("gat_tree_node %zu\n", sizeof(gat_tree_node));
  F(gat_range_run, q); F(gat_range_run, count); F(gat_range_run, first);
  printf("gat_range_run %zu\n", sizeof(gat_range_run));
  return 0;
}
''')
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                               text=True).stdout.split("\n") if line)
    for name, dt in (("gath_request", api.REQUEST), ("gath_outcome", api.OUTCOME), ("gat_tree_node", api.TREE_NODE),
                     ("gat_range_run", api.RUN)):
        assert int(got[name]) == dt.itemsize, name
        for field, (_, off) in dt.fields.items():
            assert int(got[f"{name}.{field}"]) == off, (name, field)


def test_serve_request_stream_shape():
    db = W.CFG1.db(300)
    req, qlon, qlat = W.serve_requests(W.CFG5, db, 0.5, rate=2000)
    assert req.dtype == api.REQUEST and len(req) == 1000
    assert (np.diff(req["arrival_ms"]) >= 0).all()
    frac = (req["kind"] == api.REQ_RANGE).mean()
    assert 0.7 < frac < 0.9
    t = req[req["kind"] == api.REQ_TOPK]
    assert ((t["qlen"] >= 16) & (t["qlen"] <= 512)).all()
    assert (t["qoff"] + t["qlen"] <= len(qlon)).all() and len(qlon) == len(qlat)
    r = req[req["kind"] == api.REQ_RANGE]["box"]
    side = r[:, 2] - r[:, 0]
    assert (side > 0).all() and (side <= 5e-2 * 1.0001).all()
    again, _, _ = W.serve_requests(W.CFG5, db, 0.5, rate=2000)
    assert (again == req).all()


# ------------------------------------------------------------- MortonGrid

def _edge_and_random_rects(cell, n_edge=60, seed=5):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, int(1 / cell), size=(n_edge, 2)).astype(np.float64)
    span = rng.integers(1, 4, size=(n_edge, 2))
    edge = np.stack([k[:, 0] * cell, 1 - (k[:, 1] + span[:, 1]) * cell, (k[:, 0] + span[:, 0]) * cell,
                     1 - k[:, 1] * cell], 1).astype(np.float32)
    return np.concatenate([W.CFG1.rects()[:60], api.synth_rects(seed, 60, 1e-4, 0.5, True), edge,
                           np.array([[0, 0, 1, 1], [-1, -1, 2, 2], [3, 3, 4, 4]], np.float32)]).astype(np.float32)


@pytest.mark.parametrize("cell,leaf", [(0.01, 256), (1 / 16, 8), (0.001, 32), (0.25, 100000)])
def test_mgrid_selection_matches_oracle(cell, leaf):
    """The product MortonGrid's host selection (MortonGrid.cpp:355-393) gives
    the oracle's leaves as store runs, and the same leaves as the product Grid."""
    db = W.CFG1.db(2000)
    om = O.MortonGrid(W.UNIT_BOX, cell, leaf).add(db)
    pm = api.MortonGridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False)
    pg = api.GridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False)
    info, oi = pm.info(), om.info()
    assert (info[0], info[5], info[6]) == (oi["points"], oi["leaves"], oi["total_level"])
    assert info[5] == pg.info()[5]
    rects = _edge_and_random_rects(cell)
    tasks, cand = pm.select(rects)
    gtasks, gcand = pg.select(rects)
    assert (tasks == gtasks).all() and (cand == gcand).all()
    for q in range(len(rects)):
        want = _merged_runs([(s, c) for _, _, s, c in om.leaves(rects[q]).tolist()])
        assert _task_runs(tasks, q) == want, q


def _morton_compact(v):
    v &= 0x55555555
    v = (v | (v >> 1)) & 0x33333333
    v = (v | (v >> 2)) & 0x0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF
    return (v | (v >> 8)) & 0x0000FFFF


def _morton_wave_select(codes, starts, level, xmin, ymax, cs, box, stack_cap=4096, wave=64):
    """morton_select_kernel restated: the stack holds (first cell, level) of
    selected inner nodes; children's boxes come from their first and last
    cells in float32, a child is a leaf when the code at its first cell names
    its level, and its points are the store starts of its first cell and of
    the cell after it."""
    f32 = np.float32
    xmin, ymax, cs = f32(xmin), f32(ymax), f32(cs)

    def block_box(first, span):
        last = first + span - 1
        c0, r0 = _morton_compact(first), _morton_compact(first >> 1)
        c1, r1 = _morton_compact(last), _morton_compact(last >> 1)
        return np.array([xmin + cs * f32(c0), ymax - cs * f32(r1 + 1), xmin + cs * f32(c1 + 1),
                         ymax - cs * f32(r0)], np.float32)

    sel = []
    stack = []
    code = int(codes[0])
    if code & 15 == 0:
        cnt = int(starts[1 << (2 * level)] - starts[0])
        if cnt > 0:
            sel.append((int(starts[0]), cnt))
    else:
        stack.append((0, 0))
    top_max = len(stack)
    while stack:
        m = min(len(stack), wave, (stack_cap - len(stack)) // 3)
        assert m > 0
        batch = stack[len(stack) - m:]
        del stack[len(stack) - m:]
        pushes = []
        for first, lv in batch:
            cl = lv + 1
            span = 1 << (2 * (level - cl))
            for c in range(4):
                fk = first + c * span
                if O.box_intersect(box, block_box(fk, span)) == 0:
                    continue
                code = int(codes[fk])
                cnt = int(starts[fk + span] - starts[fk])
                if code & 15 == cl:
                    if cnt > 0:
                        sel.append((int(starts[fk]), cnt))
                elif cl < level:
                    pushes.append((fk, cl))
        stack.extend(pushes)
        top_max = max(top_max, len(stack))
    return sel, top_max


@pytest.mark.parametrize("cell,leaf", [(0.01, 256), (1 / 16, 8), (1 / 1024, 8), (0.25, 100000)])
def test_mgrid_device_tree_selects_oracle_leaves(cell, leaf):
    """The implicit tree the batched form uploads — leaf records in Morton
    order and one word per cell, (rank << 4) | level at each leaf's first
    cell — walked as morton_select_kernel walks it, selects the oracle's
    leaves, within the stack bound."""
    db = W.CFG1.db(1500)
    om = O.MortonGrid(W.UNIT_BOX, cell, leaf).add(db)
    pm = api.MortonGridIndex(W.UNIT_BOX, cell, leaf).build(db, upload=False)
    recs, depth = pm.tree_nodes()
    L = int(pm.info()[6])
    assert depth == L and len(recs) == om.info()["leaves"]
    assert (recs["first_child"] == -1).all()
    # Leaves in Morton order tile the store.
    assert recs["start"][0] == 0 and (recs["start"][1:] == recs["start"][:-1] + recs["count"][:-1]).all()
    assert int(recs["count"].sum()) == db.npoints
    # The cell codes, rebuilt from the leaf boxes (first cell and width), and
    # the store start of every cell.
    starts = pm.cell_starts()
    assert len(starts) == (1 << (2 * L)) + 1 and starts[-1] == db.npoints
    codes = np.full(1 << (2 * L), 0xFFFFFFFF, np.uint64)
    for k, r in enumerate(recs):
        col = int(round(float(r["box"][0]) / cell))
        row = int(round((1.0 - float(r["box"][3])) / cell))
        width = int(round(float(r["box"][2] - r["box"][0]) / cell))
        lv = L - (width.bit_length() - 1)
        codes[O.morton(col, row)] = (k << 4) | lv
        assert starts[O.morton(col, row)] == r["start"]
    rects = _edge_and_random_rects(cell, 20, 11)
    for r in rects:
        sel, top = _morton_wave_select(codes, starts, L, 0.0, 1.0, cell, r)
        want = [(int(s), int(c)) for _, _, s, c in om.leaves(r).tolist() if c > 0]
        assert _runs(sel) == _runs(want)
        assert top <= (L + 1) * 4 * 64


@pytest.mark.parametrize("g", [1, 7, 64])
def test_candidate_grid_scan_reports_every_kept_pair_once(g):
    """The top-k filter's grid scan (restated in tests/cand_grid.py with the
    kernel's float32 cell arithmetic): every pair the box rule keeps is found
    exactly once, on random, NaN, inverted and wide boxes and on boxes placed
    a float-difference of exactly u apart."""
    from tests import cand_grid as CG
    u = np.float32(api.eps_threshold(5e-4))
    q, c = CG.cases(g, 60, 400)
    # Candidates right at the rule's edge of query 0: c.xmin - q.xmax == u and one float below.
    q[0] = np.array([0.3, 0.3, 0.31, 0.31], np.float32)
    edge = np.float32(q[0, 2] + u)
    while edge - q[0, 2] < u:
        edge = np.nextafter(edge, np.float32(1))
    c[0] = np.array([edge, 0.3, edge + np.float32(0.001), 0.31], np.float32)
    c[1] = np.array([np.nextafter(edge, np.float32(0)), 0.3, edge + np.float32(0.001), 0.31], np.float32)
    want = {(i, j) for i in range(len(q)) for j in np.nonzero(CG.keep(q[i], c, u))[0]}
    got = CG.scan(CG.build(c, g), q, c, u)
    assert len(got) == len(set(got))
    assert set(got) == want
    assert (0, 1) in want and (0, 0) not in want
