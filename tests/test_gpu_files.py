"""The reference's end-to-end flow from files on the GPU (main.cpp:117-142):
raw vehicle records ingested (PreProcess.cpp:48-201), written in the
preprocessed format (PreProcess.cpp:203-223) and read back
(readTraFromFormatedFile, PreProcess.cpp:225-315); then the Grid, MortonGrid,
STIG and FSG indexes are built on the device over the file's bounds and answer
range rectangles and an EDR top-k.

Parity: the range answers equal the oracle's restatements as sets of (trajectory
id, point id) pairs (GT-quadtree and MortonGrid against Grid::rangeQueryBatch,
Grid.cpp:341-405; STIG and FSG against inclusive brute force), through the plain
and batched forms; the top-k ids and distances equal the oracle's brute-force
EDR (Grid.cpp:1222-1318) sorted by (distance, id).  Bar: bit-exact."""
import numpy as np
import pytest

import oracle as O
from gat_amd import api

pytestmark = pytest.mark.gpu

EPS = 5e-4
CELL = 0.005


def _vehicle_csv(rng, nveh, steps):
    """Vehicles driving random walks around (121.2, 31.2), one record every
    10-40 s, with a few long stops that split trajectories (gap > 3600 s)."""
    rows = []
    for v in range(nveh):
        lon, lat = rng.uniform(121.0, 121.4), rng.uniform(31.0, 31.4)
        t = int(rng.integers(0, 20_000))
        for _ in range(steps):
            t += int(rng.integers(10, 40)) if rng.random() > 0.003 else 4000
            lon += rng.normal(0, 1e-4)
            lat += rng.normal(0, 1e-4)
            day, rem = divmod(t, 86400)
            when = f"2014-07-{1 + day:02d} {rem // 3600:02d}:{rem % 3600 // 60:02d}:{rem % 60:02d}"
            rows.append(",".join(["r", "s", f"V{v:04d}", f"{lon:.6f}", f"{lat:.6f}", "a", "b", "c", when, "x", "y",
                                  "z"]))
    # Records arrive in time order, interleaved across vehicles.
    order = np.argsort([r.split(",")[8] for r in rows], kind="stable")
    return "\n".join(rows[i] for i in order) + "\n"


def _sets(per_query):
    return [set(map(tuple, np.asarray(r).tolist())) for r in per_query]


@pytest.fixture(scope="module")
def from_files(gpu, tmp_path_factory):
    d = tmp_path_factory.mktemp("files")
    rng = np.random.default_rng(117)
    csv = d / "vehicles.csv"
    csv.write_text(_vehicle_csv(rng, 300, 400))
    db0, tid, tm, bnd = api.ingest_vehicle_csv(csv, "2014-07-01")
    pre = d / "preprocessed.txt"
    api.write_trajectory_file(pre, db0, tid, tm, bnd)
    db, _, _, bounds = api.read_trajectory_file(pre)
    assert db.ntraj >= 300 and db.npoints > 100_000
    assert db.npoints > 50 * db.ntraj  # time-ordered records make long trajectories
    box = tuple(float(v) for v in bounds)
    lo = np.array(box[:2], np.float64)
    ext = np.array(box[2:], np.float64) - lo
    side = 10 ** rng.uniform(-3, np.log10(0.05), 150)
    ctr = lo + rng.random((150, 2)) * ext
    rects = np.stack([ctr[:, 0] - side / 2, ctr[:, 1] - side / 2, ctr[:, 0] + side / 2, ctr[:, 1] + side / 2],
                     axis=1).astype(np.float32)
    print(f"[files] {db.ntraj} trajectories, {db.npoints:,} points read back, bounds {box}", flush=True)
    return db, box, rects


def _check(ix, want, rects, name):
    assert _sets(ix.range_pairs(rects).per_query()) == want, f"{name}: plain"
    assert _sets(ix.range_batched(rects).per_query()) == want, f"{name}: batched"
    ix.set_batch_params(16, 0)
    assert _sets(ix.range_batched(rects).per_query()) == want, f"{name}: leaf-major"


def test_file_flow_range(from_files):
    db, box, rects = from_files
    want_grid = _sets(O.Grid(box, CELL, 256).add(db).range(rects))
    want_brute = _sets(O.brute_range(db, rects))
    assert sum(map(len, want_brute)) > 10_000
    _check(api.GridIndex(box, CELL, 256).build(db, device=True), want_grid, rects, "GT-quadtree")
    _check(api.MortonGridIndex(box, CELL, 256).build(db, device=True), want_grid, rects, "MortonGrid")
    _check(api.StigIndex(256).build(db, device=True), want_brute, rects, "STIG")
    _check(api.FsgIndex(box, CELL).build(db, device=True), want_brute, rects, "FSG")


def test_file_flow_topk(from_files):
    db, box, _ = from_files
    g = api.GridIndex(box, CELL, 256).build(db, device=True)
    qs = api.synth_queries(7, np.array([64, 200, 17], np.int32), db, copy_mode=2)
    all_ids = np.arange(1, db.ntraj + 1, dtype=np.int32)
    ids_b, dist_b, _ = g.similarity_batched(qs, 10, EPS)
    ids_p, dist_p = g.similarity(qs, 10, EPS)
    for i in range(qs.ntraj):
        wi, wd = O.topk(O.edr_batch(qs.xy(i + 1), db, EPS, nthreads=8), all_ids, 10)
        assert (ids_b[i] == wi).all() and (dist_b[i] == wd).all(), f"batched query {i}"
        assert (ids_p[i] == wi).all() and (dist_p[i] == wd).all(), f"plain query {i}"
