"""GPU tests of the serving loop: range rectangles and EDR top-k queries arrive
over time (Poisson, 2000 requests/s, 80% range) and are batched as they come.
Every result in the mix must equal the same request answered alone, bit for
bit (range: count and order-independent pair digest; top-k: ids and
distances); batches must really form and range and top-k work must overlap."""
import numpy as np
import pytest

import oracle as O
from gat_amd import api, workloads as W

pytestmark = pytest.mark.gpu
K, EPS = 20, 5e-4


def _query(req, i, qlon, qlat):
    b, n = int(req["qoff"][i]), int(req["qlen"][i])
    return api.TrajSet(np.array([0, n], np.int64), qlon[b:b + n].copy(), qlat[b:b + n].copy())


@pytest.fixture(scope="module")
def served(gpu):
    db = W.CFG2.db(60_000)
    g = api.GridIndex(W.UNIT_BOX, 0.004, 256).build(db)
    req, qlon, qlat = W.serve_requests(W.CFG5, db, 0.4, rate=2000)
    out, ids, dists, st = api.serve(g, req, qlon, qlat, K, EPS)
    return db, g, req, qlon, qlat, out, ids, dists, st


def test_each_result_same_alone_and_in_the_mix(served):
    db, g, req, qlon, qlat, out, ids, dists, st = served
    rng = req["kind"] == api.REQ_RANGE
    assert st["range_done"] == rng.sum() and st["topk_done"] == (~rng).sum()
    for i in np.nonzero(rng)[0]:
        alone = g.range_pairs(req["box"][i][None, :])
        assert out["count"][i] == alone.counts()[0] and out["digest"][i] == alone.digests()[0], i
    for i in np.nonzero(~rng)[0]:
        a_ids, a_d, _ = g.similarity_batched(_query(req, i, qlon, qlat), K, EPS)
        assert (ids[i] == a_ids[0]).all() and (dists[i] == a_d[0]).all(), i


def test_range_in_the_mix_matches_oracle(served):
    """Every range result of the mix against the oracle's restatement of
    Grid::rangeQueryBatch (Grid.cpp:341-405): count and pair digest."""
    db, g, req, qlon, qlat, out, ids, dists, st = served
    rng = np.nonzero(req["kind"] == api.REQ_RANGE)[0]
    want_c, want_d = O.Grid(W.UNIT_BOX, 0.004, 256).add(db).range(req["box"][rng], "digest", nthreads=16)
    assert want_c.sum() > 0
    assert (out["count"][rng] == want_c).all() and (out["digest"][rng] == want_d).all()


def test_range_admission_never_grows_buffers(served):
    """Range requests are charged an exact bound of their result and list
    buffers (from the leaves they select), so a range-only stream whose
    requests fit the budget is served with the workspaces reserved up front:
    no device allocation while serving.  A request above the budget on its own
    is still admitted, alone, and its buffers grow."""
    db, g, req, qlon, qlat, out, ids, dists, st = served
    rng = req[req["kind"] == api.REQ_RANGE].copy()
    want = out["count"][req["kind"] == api.REQ_RANGE]
    r_out, _, _, r_st = api.serve(g, rng, qlon, qlat, K, EPS, range_budget=4e9)
    assert r_st["device_allocs"] == 0
    assert r_st["range_done"] == len(rng) and (r_out["count"] == want).all()
    assert 0 < r_st["peak_range_bytes"] <= 4e9
    t_out, _, _, t_st = api.serve(g, rng, qlon, qlat, K, EPS, range_budget=1.0)
    assert t_st["range_batches"] == len(rng) and (t_out["count"] == want).all()


def test_topk_in_the_mix_matches_oracle(served):
    db, g, req, qlon, qlat, out, ids, dists, st = served
    tid = np.arange(1, db.ntraj + 1, dtype=np.int32)
    assert st["peak_topk_survivors"] > 0
    for i in np.nonzero(req["kind"] == api.REQ_TOPK)[0][:12]:
        full = O.edr_batch(_query(req, i, qlon, qlat).xy(1), db, EPS, nthreads=16)
        oi, od = O.topk(full, tid, K)
        assert (ids[i] == oi).all() and (dists[i] == od).all(), i


def test_batches_form_and_overlap(served):
    db, g, req, qlon, qlat, out, ids, dists, st = served
    rng = req["kind"] == api.REQ_RANGE
    assert (out["admit_ms"] >= req["arrival_ms"] - 1e-6).all() and (out["done_ms"] >= out["admit_ms"]).all()
    assert st["range_batches"] < rng.sum() and np.bincount(out["batch"][rng]).max() >= 2
    # Some range batch and some top-k batch were on the device at the same time.
    rb = {b: (out["admit_ms"][rng][out["batch"][rng] == b].min(), out["done_ms"][rng][out["batch"][rng] == b].max())
          for b in np.unique(out["batch"][rng])}
    tb = {b: (out["admit_ms"][~rng][out["batch"][~rng] == b].min(),
              out["done_ms"][~rng][out["batch"][~rng] == b].max()) for b in np.unique(out["batch"][~rng])}
    overlaps = sum(1 for a in rb.values() for b in tb.values() if a[0] < b[1] and b[0] < a[1])
    assert overlaps > 0


def test_plain_path_gives_the_same_answers(served):
    db, g, req, qlon, qlat, out, ids, dists, st = served
    sub = req[:120].copy()
    p_out, p_ids, p_d, p_st = api.serve(g, sub, qlon, qlat, K, EPS, batched=False)
    assert p_st["range_batches"] + p_st["topk_batches"] == len(sub)
    assert (p_out["count"] == out["count"][:120]).all() and (p_out["digest"] == out["digest"][:120]).all()
    t = sub["kind"] == api.REQ_TOPK
    assert (p_ids[t] == ids[:120][t]).all() and (p_d[t] == dists[:120][t]).all()


def test_memory_budget_bounds_batches(served):
    db, g, req, qlon, qlat, out, ids, dists, st = served
    sub = req[:200].copy()
    b_out, _, _, b_st = api.serve(g, sub, qlon, qlat, K, EPS, range_budget=1.0, topk_budget=1.0)
    rng = sub["kind"] == api.REQ_RANGE
    # A budget below any single request admits one request per batch.
    assert b_st["range_batches"] == rng.sum() and b_st["topk_batches"] == (~rng).sum()
    assert (b_out["count"] == out["count"][:200]).all() and (b_out["digest"] == out["digest"][:200]).all()
