#!/usr/bin/env python3
"""Benchmark of the trajectory-query hot path on one MI355X.

Headline (value): range queries/s of the batched GT-quadtree form on the
1M-trajectory workload (88.2M points, 1024x1024 Morton grid, 256-point leaves,
10k squares with log-uniform side in [1e-4, 1e-1]).  One step = the 10k
rectangles through Grid::rangeQueryBatchedGPU: box upload, the captured device
pipeline (tree selection, scans, leaf grouping, leaf-major refine) replayed as
one graph, per-query count and region readback.  The point store and tree are
resident in HBM before timing starts; the result pairs stay there.

Beside it on the same line: the plain drop-in on the same batch (host leaf
selection as Grid::rangeQueryBatchGPU, task upload, the per-task refine
kernel) and the ratio; STIG and FSG; the plain and batched EDR cell-update
rates over one pair set, the EDR kernel's VALU roofline and the batched top-k
rate; the serving loop on a 2B-point index; the mixed batch of 100k rectangles
over 10M trajectories routed across the three indexes; the leaf-major refine
kernel's roofline (algorithmic bytes 12 per point of a selected leaf + 8 per
result pair, over its HIP-event time on its own stream); and the oracle's CPU
restatements timed on this host: Grid::rangeQueryBatch, the EDR brute force
and the reference's frequency-distance top-k loop.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Integer VALU issue of one MI355X: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per
# wave64 instruction x 64 lanes = 3.93e13 lane-ops/s.
VALU_PEAK_TOPS = 39.3
# VALU instructions per DP cell of edr_wave_kernel<N, 16> (AMD ISA of the built
# kernel: 2 subtractions + 2 compares for the match, the row shift, M + s,
# min3, compare, select, the +1 add).
EDR_VALU_PER_CELL = 10
METRIC = "range queries/s + EDR cell-updates/s, batched vs plain; p99 latency"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def load_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc, committed under profiles/."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get(name)
    except (ValueError, OSError):
        return None


def time_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    wall = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        res = fn()
        wall.append(time.perf_counter() - s)
    return res, time.perf_counter() - t0, np.array(wall)


def time_range(index, rects, steps, warmup, label):
    res, total, wall = time_steps(lambda: index.range_pairs(rects), steps, warmup)
    qps = steps * len(rects) / total
    log(f"{label}: {qps:,.0f} queries/s, step {1e3 * total / steps:.2f} ms, kernel {res.kernel_ms:.3f} ms, "
        f"select {res.select_ms:.2f} ms, {res.pairs_total:,} pairs from {res.candidates:,} candidates")
    return res, qps, total, wall


def time_batched(index, rects, steps, warmup, label):
    res, total, wall = time_steps(lambda: index.range_batched(rects, graph=True), steps, warmup)
    qps = steps * len(rects) / total
    log(f"{label}: {qps:,.0f} queries/s, step {1e3 * total / steps:.2f} ms, device pipeline {res.kernel_ms:.3f} ms, "
        f"replayed {res.replayed}, {res.leaf_queries:,} (leaf, query) entries, {res.work_items:,} items")
    return res, qps, total, wall


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 10M-trajectory mixed batch")
    ap.add_argument("--serve-ntraj", type=int, default=None,
                    help="trajectories of the serving index (default: the full 22.7M, about 2B points)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="length of the served request stream")
    ap.add_argument("--mixed-ntraj", type=int, default=None,
                    help="trajectories of the mixed workload (default: the full 10M, about 0.88B points)")
    ap.add_argument("--mixed-nq", type=int, default=None, help="rectangles of the mixed workload (default 100k)")
    ap.add_argument("--only", choices=["all", "range", "batched", "mgrid", "edr", "serve", "mixed", "copy", "hbm"],
                    default="all",
                    help="range / batched: just the timed GT-quadtree batches of the plain / batched form (for "
                         "profiling); mgrid: the batched form over MortonGrid's implicit tree; edr: just the EDR section; copy: just the stream-copy calibration kernel; "
                         "hbm: write / read / packed-run access-pattern ceilings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); runs over several GPUs are "
              "not supported", file=sys.stderr)
        return 2

    import gat_amd as G
    from gat_amd import workloads as W

    G.init(0)
    dev = G.device_name().strip()
    log(f"device {dev}")
    if args.only == "copy":
        gbs, ms = G.stream_copy_bandwidth(1 << 32, max(args.steps, 1))
        log(f"stream copy: {gbs:.1f} GB/s ({ms:.3f} ms per 4 GiB copy)")
        print(json.dumps({"stream_copy_gbs": round(gbs, 1), "kernel_ms": round(ms, 4), "bytes_per_launch": 2 << 32}))
        return 0
    if args.only == "hbm":
        print(json.dumps(bench_hbm(G)))
        return 0
    if args.only == "edr":
        print(json.dumps(bench_edr(G, W)))
        return 0
    if args.only == "mixed":
        print(json.dumps(bench_mixed(G, W, args.mixed_ntraj, args.mixed_nq)))
        return 0
    if args.only == "serve":
        print(json.dumps(bench_serve(G, W, args.serve_ntraj, args.serve_seconds)))
        return 0
    w = W.CFG2
    t = time.perf_counter()
    db = w.db()
    rects = w.rects()
    log(f"workload {w.name}: {db.ntraj:,} trajectories, {db.npoints:,} points, {len(rects):,} rectangles "
        f"({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    kind = G.MortonGridIndex if args.only == "mgrid" else G.GridIndex
    grid = kind(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    build_device_s = time.perf_counter() - t
    log(f"GT-quadtree built on the device and resident ({build_device_s:.2f} s): {grid.info().tolist()}")

    if args.only == "range":
        res, qps, total, _ = time_range(grid, rects, args.steps, args.warmup, "GT-quadtree plain")
        alg = 12 * res.candidates + 8 * res.pairs_total
        print(json.dumps({"value": round(qps, 1), "kernel_ms": round(res.kernel_ms, 4),
                          "achieved_gbs": round(alg / (res.kernel_ms * 1e-3) / 1e9, 1), "algorithmic_bytes": alg}))
        return 0
    if args.only in ("batched", "mgrid"):
        label = "MortonGrid batched" if args.only == "mgrid" else "GT-quadtree batched"
        res, qps, total, _ = time_batched(grid, rects, args.steps, args.warmup, label)
        runs = [grid.range_batched(rects, graph=False) for _ in range(3)]
        refine_ms = float(np.median([x.refine_ms for x in runs]))
        select_ms = float(np.median([x.select_ms for x in runs]))
        print(json.dumps({"value": round(qps, 1), "pipeline_ms": round(res.kernel_ms, 4),
                          "refine_kernel_ms": round(refine_ms, 4), "selection_ms": round(select_ms, 4)}))
        return 0

    tuned_range = grid.tune_batched(rects)
    log(f"autotuned rectangles per work item: {tuned_range}")
    pres, pqps, ptotal, pwall = time_range(grid, rects, args.steps, args.warmup, "GT-quadtree plain")
    bres, bqps, btotal, bwall = time_batched(grid, rects, args.steps, args.warmup, "GT-quadtree batched")
    # Per-kernel events come from step-by-step runs (a graph replay times the whole pipeline).
    refine_ms = []
    for _ in range(5):
        r = grid.range_batched(rects, graph=False)
        refine_ms.append(r.refine_ms)
    refine_ms = float(np.median(refine_ms))
    unique = grid.last_unique_candidates()
    alg_bytes = 12 * unique + 8 * r.pairs_total
    achieved = alg_bytes / (refine_ms * 1e-3) / 1e9
    same = bool((bres.counts() == pres.counts()).all() and (bres.digests() == pres.digests()).all())
    log(f"leaf-major refine {refine_ms:.3f} ms: {achieved:,.0f} GB/s algorithmic ({unique:,} distinct candidate "
        f"points, {r.pairs_total:,} pairs); selection {r.select_ms:.3f} ms; batched == plain: {same}")
    plain_alg = 12 * pres.candidates + 8 * pres.pairs_total
    plain_achieved = plain_alg / (pres.kernel_ms * 1e-3) / 1e9
    traffic = load_traffic("leaf_refine_runs_kernel:cfg2")
    copy_gbs, _ = G.stream_copy_bandwidth(1 << 32, 10)
    log(f"stream copy reference: {copy_gbs:.1f} GB/s")

    t = time.perf_counter()
    host_grid = G.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db)
    build_host_s = time.perf_counter() - t
    same_store = bool((host_grid.cell_starts() == grid.cell_starts()).all())
    del host_grid
    log(f"GT-quadtree host build for comparison: {build_host_s:.2f} s (device {build_device_s:.2f} s); same cell "
        f"offsets: {same_store}")
    t = time.perf_counter()
    stig_host = G.StigIndex(w.leaf).build(db)
    stig_host_s = time.perf_counter() - t
    t = time.perf_counter()
    stig = G.StigIndex(w.leaf).build(db, device=True)
    stig_device_s = time.perf_counter() - t
    fsg = G.FsgIndex(W.UNIT_BOX, w.fsg_cell).build(db, device=True)
    hres = stig_host.range_batched(rects)
    hc, hd = hres.counts(), hres.digests()
    del stig_host
    dres = stig.range_batched(rects)
    stig_same = bool((dres.counts() == hc).all() and (dres.digests() == hd).all())
    log(f"STIG kd-split on the device {stig_device_s:.2f} s (host {stig_host_s:.2f} s), same answers: {stig_same}; "
        f"FSG built on the device")
    few = max(3, args.steps // 2)
    _, fqps, _, _ = time_range(fsg, rects, few, 1, "FSG plain")
    _, sqps, _, _ = time_range(stig, rects, few, 1, "STIG plain")
    _, sbqps, _, _ = time_batched(stig, rects, few, 1, "STIG batched")
    _, fbqps, _, _ = time_batched(fsg, rects, few, 1, "FSG batched")
    del fsg, stig
    # MortonGrid: the same quadtree kept implicitly, walked on the device
    # without node records; its answers must be the GT-quadtree's.
    mgrid = G.MortonGridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    mgrid.set_batch_params(tuned_range["queries_per_item"], 256)
    _, mqps, _, _ = time_range(mgrid, rects, few, 1, "MortonGrid plain")
    mres, mbqps, _, _ = time_batched(mgrid, rects, args.steps, args.warmup, "MortonGrid batched")
    msel = float(np.median([mgrid.range_batched(rects, graph=False).select_ms for _ in range(3)]))
    msame = bool((mres.counts() == bres.counts()).all() and (mres.digests() == bres.digests()).all())
    log(f"MortonGrid selection {msel:.3f} ms (GT-quadtree {r.select_ms:.3f} ms); answers equal the GT-quadtree's: "
        f"{msame}")
    del mgrid

    edr = bench_edr(G, W, cpu_topk=not args.no_cpu_baseline)
    topk_cpu = edr.pop("topk_cpu")
    del grid
    serve = bench_serve(G, W, args.serve_ntraj, args.serve_seconds)
    mixed = {} if args.no_mixed else bench_mixed(G, W, args.mixed_ntraj, args.mixed_nq)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        t = time.perf_counter()
        og = O.Grid(W.UNIT_BOX, w.grid_cell, w.leaf).add(db)
        log(f"oracle GT-quadtree built ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        counts, digests = og.range(rects, "digest", nthreads=1)
        cpu_s = time.perf_counter() - t
        # The GPU answer of the timed batches, checked against the oracle.
        ok = bool((bres.counts() == counts).all() and (bres.digests() == digests).all())
        t = time.perf_counter()
        og.range(rects, "digest", nthreads=16)
        cpu16_s = time.perf_counter() - t
        cpu = {"value": round(len(rects) / cpu_s, 1), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"all {len(rects)} rectangles of the timed batch, Grid::rangeQueryBatch restated "
                         f"(oracle/oracle.cpp), 1 thread; 16 threads as Grid::rangeQueryBatchMultiThread: "
                         f"{len(rects) / cpu16_s:.1f} queries/s",
               "parity": ok}
        log(f"CPU baseline: {cpu['value']} queries/s (1 thread), {len(rects) / cpu16_s:.1f} (16 threads); "
            f"GPU result identical: {ok}")
        del og
        # EDR on the CPU: Grid::SimilarityExecuter restated, one 128-point query
        # against the first 20k candidates of the top-k workload.
        w3 = W.CFG3
        db3 = w3.db(20_000)
        q3 = G.synth_queries(w3.seed + 7, np.array([128], np.int32), db3, copy_mode=2)
        cells3 = int(db3.lengths.astype(np.int64).sum() * 128)
        rates = {}
        for th in (1, 16):
            t = time.perf_counter()
            O.edr_batch(q3.xy(1), db3, w3.eps, nthreads=th)
            rates[th] = cells3 / (time.perf_counter() - t)
        cpu["edr_cells_per_s"] = {"1_thread": float(f"{rates[1]:.4e}"), "16_threads": float(f"{rates[16]:.4e}"),
                                  "sample": f"one 128-point query x 20k candidates ({cells3:,} cells)"}
        log(f"CPU EDR: {rates[1]:.3e} cells/s (1 thread), {rates[16]:.3e} (16 threads)")
        cpu["topk"] = topk_cpu
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        cpu["host"] = {"model": model, "nproc": os.cpu_count(),
                       "threads_note": "pools capped at 16 threads, the one-GPU share of the host's cores"}

    line = {
        "metric": METRIC,
        "value": round(bqps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * btotal / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random walks)",
        "config": {"workload": "1M trajectories, 88.2M points, 1024x1024 GT-quadtree, 256-point leaves, 10k "
                               "squares log-uniform side [1e-4,1e-1]; batched form (device selection, leaf-major "
                               "join, graph replay)",
                   "index": "GT-quadtree", "rectangles": int(len(rects)), "points": int(db.npoints)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "leaf_refine_runs_kernel", "kernel_ms": round(refine_ms, 4),
                     "algorithmic_bytes": int(alg_bytes)},
        "cpu_baseline": cpu,
        "paths": {
            "gt_quadtree_batched_qps": round(bqps, 1),
            "gt_quadtree_plain_qps": round(pqps, 1),
            "batched_over_plain": round(bqps / pqps, 3),
            "batched_equals_plain": same,
            "plain_refine_kernel_ms": round(pres.kernel_ms, 4),
            "plain_refine_frac_hbm": round(plain_achieved / HBM_PEAK_GBS, 4),
            "batched_selection_ms": round(r.select_ms, 4),
            "autotuned_queries_per_item": tuned_range["queries_per_item"],
            "batched_pipeline_ms": round(bres.kernel_ms, 4),
            "stig_batched_qps": round(sbqps, 1),
            "stig_plain_qps": round(sqps, 1),
            "fsg_batched_qps": round(fbqps, 1),
            "mgrid_batched_qps": round(mbqps, 1),
            "mgrid_plain_qps": round(mqps, 1),
            "mgrid_selection_ms": round(msel, 4),
            "mgrid_equals_grid": msame,
            "fsg_plain_qps": round(fqps, 1),
            "range_batched_p50_ms": round(1e3 * float(np.percentile(bwall, 50)), 3),
            "range_batched_p99_ms": round(1e3 * float(np.percentile(bwall, 99)), 3),
            "range_plain_p50_ms": round(1e3 * float(np.percentile(pwall, 50)), 3),
            "range_plain_p99_ms": round(1e3 * float(np.percentile(pwall, 99)), 3),
            "stream_copy_gbs": round(copy_gbs, 1),
            "gt_quadtree_build_device_s": round(build_device_s, 3),
            "stig_build_device_s": round(stig_device_s, 3),
            "stig_build_host_s": round(stig_host_s, 3),
            "stig_device_build_same_answers": stig_same,
            "gt_quadtree_build_host_s": round(build_host_s, 3),
            "device_build_equals_host": same_store,
            **edr,
            **serve,
            **mixed,
        },
        "device": dev,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_mixed(G, W, ntraj, nq) -> dict:
    """The mixed batch: 100k squares (side log-uniform in [1e-4, 1e-1], 70% of
    centres around 16 Gaussian hotspots) over 10M trajectories, each routed by
    area to FSG, the GT-quadtree or STIG, run in groups whose three pipelines
    are one captured graph; per-rectangle counts and digests, 1000 of them
    checked against the oracle's restatement of the index they were routed to.
    The full batch writes ~0.5 TB of pairs, group by group into one reused
    buffer."""
    w = W.CFG4
    t = time.perf_counter()
    db = w.db(ntraj)
    rects = w.rects(nq)
    g = G.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    s = G.StigIndex(w.leaf).build(db, device=True)
    f = G.FsgIndex(W.UNIT_BOX, w.fsg_cell).build(db, device=True)
    log(f"mixed workload {w.name}: {db.ntraj:,} trajectories, {db.npoints:,} points, {len(rects):,} rectangles, "
        f"three indexes resident ({time.perf_counter() - t:.1f} s)")
    m = G.MixedRange(g, s, f)
    m.run(rects[:6144], group_size=2048)  # warm-up: sizes the workspaces
    counts, digests, route, st = m.run(rects, group_size=2048)
    pipeline_s = (st["ms"] - st["digest_ms"]) / 1e3
    res = {"mixed_points": int(db.npoints), "mixed_rectangles": int(len(rects)),
           "mixed_pairs": int(st["pairs"]), "mixed_groups": int(st["groups"]), "mixed_replays": int(st["replays"]),
           "mixed_routed": {k: int((route == i).sum()) for i, k in enumerate(("fsg", "gt_quadtree", "stig"))},
           "mixed_qps": round(len(rects) / pipeline_s, 1), "mixed_pairs_per_s": float(f"{st['pairs'] / pipeline_s:.4e}"),
           "mixed_wall_s": round(st["ms"] / 1e3, 3), "mixed_digest_s": round(st["digest_ms"] / 1e3, 3)}
    log(f"mixed: {res['mixed_qps']:,.0f} rectangles/s ({st['pairs']:,.0f} pairs, {res['mixed_pairs_per_s']:.3e} "
        f"pairs/s, {st['groups']:.0f} groups, {st['replays']:.0f} graph replays; digests {res['mixed_digest_s']} s)")
    import oracle as O
    rng = np.random.default_rng(4)
    sample = np.sort(rng.choice(len(rects), size=min(1000, len(rects)), replace=False))
    t = time.perf_counter()
    ok = True
    # The GT-quadtree and FSG restatements for their rectangles; STIG's answer
    # is inclusive brute force (its corrected search is exact).
    for r in (1, 0, 2):
        idx = sample[route[sample] == r]
        if len(idx) == 0:
            continue
        if r == 2:
            c, d = O.brute_range(db, rects[idx], "digest", nthreads=16)
        else:
            ref = (O.Grid(W.UNIT_BOX, w.grid_cell, w.leaf) if r == 1 else O.FSG(W.UNIT_BOX, w.fsg_cell)).add(db)
            c, d = ref.range(rects[idx], "digest", nthreads=16)
            del ref
        good = bool((counts[idx] == c).all() and (digests[idx] == d).all())
        log(f"  route {r}: {len(idx)} sampled rectangles identical to the oracle: {good}")
        ok = ok and good
    res["mixed_sample_parity"] = ok
    log(f"mixed parity on {len(sample)} sampled rectangles: {ok} ({time.perf_counter() - t:.1f} s)")
    return res


def bench_serve(G, W, ntraj, seconds) -> dict:
    """The serving loop on the resident 2B-point index: Poisson arrivals at
    2000 requests/s (80% range squares, side log-uniform in [1e-4, 5e-2]; 20%
    EDR top-20, query lengths U{16..512}) for `seconds`, batched as they come;
    then the plain per-request path on the first requests of the same stream
    (it cannot keep up: one top-k query over every trajectory takes seconds)."""
    w = W.CFG5
    t = time.perf_counter()
    db = w.db(ntraj)
    g = G.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db, device=True)
    log(f"serving index {w.name}: {db.ntraj:,} trajectories, {db.npoints:,} points resident "
        f"({time.perf_counter() - t:.1f} s)")
    req, qlon, qlat = W.serve_requests(w, db, seconds)
    out, ids, dists, st = G.serve(g, req, qlon, qlat, w.k, w.eps)
    lat = out["done_ms"] - req["arrival_ms"]
    rng = req["kind"] == 0
    res = {"serve_points": int(db.npoints), "serve_requests": int(len(req)),
           "serve_rate_offered": w.extra["rate"],
           "serve_throughput_rps": round(len(req) / (max(out["done_ms"].max(), req["arrival_ms"].max()) / 1e3), 1),
           "serve_range_p50_ms": round(float(np.percentile(lat[rng], 50)), 3),
           "serve_range_p99_ms": round(float(np.percentile(lat[rng], 99)), 3),
           "serve_topk_p50_ms": round(float(np.percentile(lat[~rng], 50)), 3),
           "serve_topk_p99_ms": round(float(np.percentile(lat[~rng], 99)), 3),
           "serve_range_batches": int(st["range_batches"]), "serve_topk_batches": int(st["topk_batches"])}
    log(f"served {len(req):,} requests in {st['wall_ms']:.0f} ms: range p50 {res['serve_range_p50_ms']} ms p99 "
        f"{res['serve_range_p99_ms']} ms; top-{w.k} p50 {res['serve_topk_p50_ms']} ms p99 "
        f"{res['serve_topk_p99_ms']} ms; {res['serve_range_batches']} range / {res['serve_topk_batches']} top-k "
        f"batches")
    sub = req[:40].copy()
    p_out, _, _, _ = G.serve(g, sub, qlon, qlat, w.k, w.eps, batched=False)
    p_lat = p_out["done_ms"] - sub["arrival_ms"]
    p_rng = sub["kind"] == 0
    res.update({"serve_plain_requests": int(len(sub)),
                "serve_plain_range_p50_ms": round(float(np.percentile(p_lat[p_rng], 50)), 3),
                "serve_plain_topk_p50_ms": round(float(np.percentile(p_lat[~p_rng], 50)), 3) if (~p_rng).any()
                else None,
                "serve_plain_p99_ms": round(float(np.percentile(p_lat, 99)), 3),
                "serve_plain_same_answers": bool((p_out["digest"] == out["digest"][:len(sub)]).all())})
    log(f"plain per-request path on the first {len(sub)} requests: range p50 {res['serve_plain_range_p50_ms']} ms, "
        f"top-k p50 {res['serve_plain_topk_p50_ms']} ms, p99 {res['serve_plain_p99_ms']} ms")
    return res


def bench_hbm(G) -> dict:
    """Ceilings of the HBM access patterns behind the range kernels' result
    writes, over a 4 GiB buffer: streams, then runs of 8-byte pairs packed back
    to back (as per-query result regions fill) written from each run's first
    pair or through line-aligned windows, and 256-byte-aligned runs."""
    out = {}
    for kind, name in [(0, "write stream"), (1, "read stream")]:
        gbs, ms = G.hbm_pattern_bandwidth(kind)
        log(f"{name}: {gbs:.1f} GB/s ({ms:.3f} ms)")
        out[name.replace(" ", "_") + "_gbs"] = round(gbs, 1)
    for run in (28, 32, 110, 255, 256, 257, 830):
        for kind, name, key in [(2, "packed", "packed"), (3, "line-aligned windows", "aligned"),
                                (4, "packed, cached stores", "packed_cached"),
                                (5, "line-aligned windows, cached stores", "aligned_cached")]:
            gbs, ms = G.hbm_pattern_bandwidth(kind, run=run)
            log(f"runs of {run} pairs, {name}: {gbs:.1f} GB/s ({ms:.3f} ms)")
            out[f"runs{run}_{key}_gbs"] = round(gbs, 1)
    return out


def _subset(G, ts, idx):
    """Trajectories idx (0-based) of a TrajSet, as a new TrajSet."""
    lens = np.diff(ts.offsets)[idx]
    offs = np.zeros(len(idx) + 1, np.int64)
    offs[1:] = np.cumsum(lens)
    lon = np.concatenate([ts.lon[ts.offsets[i]:ts.offsets[i + 1]] for i in idx])
    lat = np.concatenate([ts.lat[ts.offsets[i]:ts.offsets[i + 1]] for i in idx])
    return G.TrajSet(offs, lon, lat)


def _loop_vs_exact(fids, fdist, ids, dists):
    """The frequency-distance loop's answers against the exact top-k: ids of the
    exact top-k it missed, and ranks whose distance differs (both sorted)."""
    missed = sum(len(set(ids[i].tolist()) - set(fids[i].tolist())) for i in range(len(ids)))
    worse = int((np.sort(fdist, axis=1) != dists).sum())
    return missed, worse


def cpu_reference_topk(G, W, O, queries, qlen, ids, dists, db) -> dict:
    """The reference's own CPU top-k, Grid::SimilarityQueryBatchCPUParallel
    (Grid.cpp:1133-1220): a frequency-distance queue per query (FVTable.cpp:77-207),
    rounds of 80 candidates through the DP until the worst kept distance no
    longer exceeds the next frequency distance, 20 query threads per batch;
    restated in the oracle.  Timed on 20 queries of the 1024-query batch (k =
    50, 200k candidates) and on the 20k-trajectory workload's one 128-point
    query (k = 10), and compared with the exact top-k: the loop is not exact
    (its bound is unproven and its stop rule drops ties)."""
    w3 = W.CFG3
    sample = np.arange(0, queries.ntraj, queries.ntraj // 20)[:20]
    sub = _subset(G, queries, sample)
    fdk = O.FdTopk(W.UNIT_BOX, w3.grid_cell, db)
    t = time.perf_counter()
    fids, fdist, fst = fdk.topk(sub, w3.k, w3.eps)
    s3 = time.perf_counter() - t
    missed, worse = _loop_vs_exact(fids, fdist, ids[sample], dists[sample])
    del fdk
    w1 = W.CFG1
    db1 = w1.db()
    q1 = G.synth_queries(w1.seed, np.array([128, 128], np.int32), db1, copy_mode=2)
    q1 = _subset(G, q1, np.array([1]))  # the noisy copy of a database trajectory
    fdk1 = O.FdTopk(W.UNIT_BOX, w1.grid_cell, db1)
    t = time.perf_counter()
    f1ids, f1dist, f1st = fdk1.topk(q1, w1.k, w1.eps)
    s1 = time.perf_counter() - t
    e1 = O.edr_batch(q1.xy(1), db1, w1.eps, nthreads=16)
    oi, od = O.topk(e1, np.arange(1, db1.ntraj + 1, dtype=np.int32), w1.k)
    missed1, worse1 = _loop_vs_exact(f1ids, f1dist, oi[None, :], od[None, :])
    out = {"value": round(len(sample) / s3, 4), "unit": "queries/s", "cores": 20, "kind": "port",
           "sample": f"queries {sample.tolist()} of the 1024-query batch (lengths {qlen[sample].min()}-"
                     f"{qlen[sample].max()}), top-{w3.k} over {db.ntraj:,} candidates, one batch of 20 threads",
           "rounds_mean": round(float(fst[:, 0].mean()), 2), "candidates_evaluated": int(fst[:, 1].sum()),
           "exact_ids_missed": int(missed), "ranks_with_other_distance": int(worse),
           "of_ids": int(len(sample) * w3.k),
           "workload_20k": {"ms": round(1e3 * s1, 3), "k": w1.k, "query_len": 128,
                            "rounds": int(f1st[0, 0]), "candidates_evaluated": int(f1st[0, 1]),
                            "exact_ids_missed": int(missed1), "ranks_with_other_distance": int(worse1)}}
    log(f"CPU top-k, the reference's frequency-distance loop: {out['value']} queries/s on 20 threads "
        f"({missed} of {out['of_ids']} exact ids missed, {worse} ranks at another distance); 20k-trajectory "
        f"query {1e3 * s1:.1f} ms ({missed1} of {w1.k} exact ids missed, {worse1} ranks at another distance)")
    return out


def bench_edr(G, W, cpu_topk: bool = False) -> dict:
    """EDR cell-update rates over one pair set (4 queries against the 200k
    candidates of the top-k workload, both query kinds): plain = one launch per
    query over every candidate, batched = the persistent kernel over the same
    800k pairs.  Then the exact top-50 of the 1024-query batch; with cpu_topk,
    two of its queries answered on the CPU by brute force (the oracle's EDR
    over all 200k candidates on 16 threads, then the top 50 by (distance, id))
    and compared with the GPU's answers."""
    w = W.CFG3
    t = time.perf_counter()
    db = w.db()
    g = G.GridIndex(W.UNIT_BOX, w.grid_cell, w.leaf).build(db)
    log(f"EDR workload {w.name}: {db.ntraj:,} candidates, {db.npoints:,} points ({time.perf_counter() - t:.1f} s)")
    tune_q = G.synth_queries(w.seed + 8, np.array([256, 900], np.int32), db, copy_mode=2)
    tuned = g.tune_edr(tune_q, w.eps)
    log(f"autotuned batched EDR: {tuned}")
    qlen = np.array([128, 512, 1024, 64], np.int32)
    q = G.synth_queries(w.seed + 7, qlen, db, copy_mode=2)
    cells = int(db.lengths.astype(np.int64).sum() * int(qlen.sum()))
    g.edr_all(q.xy(1), w.eps)
    t = time.perf_counter()
    plain = np.stack([g.edr_all(q.xy(i + 1), w.eps) for i in range(q.ntraj)])
    plain_s = time.perf_counter() - t
    g.edr_batched_all(q, w.eps)
    t = time.perf_counter()
    batched, st_all = g.edr_batched_all(q, w.eps)
    batched_s = time.perf_counter() - t
    same = bool((batched.reshape(plain.shape) == plain).all())
    plain_rate, batched_rate = cells / plain_s, cells / batched_s
    kernel_rate = cells / (st_all["kernel_ms"] * 1e-3)
    log(f"EDR over {q.ntraj * db.ntraj:,} pairs ({cells:.3e} cells): plain {plain_rate:.3e} cells/s, batched "
        f"{batched_rate:.3e} cells/s wall, {kernel_rate:.3e} in-kernel; equal: {same}")
    del plain, batched
    lo, hi = w.extra["qlen_lo"], w.extra["qlen_hi"]
    qlen = G.synth_lengths(w.seed + 100, w.extra["nq"], "uniform", lo, hi, lo, hi)
    queries = G.synth_queries(w.seed, qlen, db, copy_mode=2)
    g.similarity_batched(queries, w.k, w.eps)
    # Seven timed batches: the median batch time and its stage breakdown.
    runs = []
    for _ in range(7):
        t = time.perf_counter()
        ids, dists, st = g.similarity_batched(queries, w.k, w.eps)
        runs.append((time.perf_counter() - t, st))
    runs.sort(key=lambda x: x[0])
    topk_s, st = runs[len(runs) // 2]
    log(f"top-{w.k} batched: {queries.ntraj / topk_s:,.1f} queries/s (median of {len(runs)} batches: "
        f"{1e3 * topk_s:.2f} ms for {queries.ntraj}, fastest {1e3 * runs[0][0]:.2f} ms; {st['pairs']:,} pairs "
        f"evaluated, {st['pruned']:,} resolved by the mismatch table; filter kernel {st['filter_ms']:.3f} ms, DP "
        f"kernel {st['kernel_ms']:.3f} ms, host job building {st['host_ms']:.3f} ms, merge stage "
        f"{st['merge_ms']:.3f} ms, query upload {st['upload_ms']:.3f} ms)")
    cpu = None
    if cpu_topk:
        import oracle as O
        cpu = {"reference_loop": cpu_reference_topk(G, W, O, queries, qlen, ids, dists, db)}
        order = np.argsort(qlen, kind="stable")
        sample = [int(order[len(order) // 4]), int(order[len(order) // 2])]  # lengths near the 25th and 50th percentile
        all_ids = np.arange(1, db.ntraj + 1, dtype=np.int32)
        ok = True
        t = time.perf_counter()
        for qi in sample:
            d = O.edr_batch(queries.xy(qi + 1), db, w.eps, nthreads=16)
            oi, od = O.topk(d, all_ids, w.k)
            ok = ok and bool((oi == ids[qi][:len(oi)]).all() and (od == dists[qi][:len(od)]).all())
        cpu_s = time.perf_counter() - t
        cpu_cells = int(db.lengths.astype(np.int64).sum()) * int(qlen[sample].sum())
        cpu.update({"value": round(len(sample) / cpu_s, 4), "unit": "queries/s", "cores": 16,
                    "sample": f"queries {sample} (lengths {qlen[sample].tolist()}) of the 1024-query batch, exact "
                              f"top-{w.k} by brute force over all {db.ntraj:,} candidates ({cpu_cells:.3e} cells)",
                    "equals_gpu": ok})
        log(f"CPU top-{w.k}: {cpu['value']} queries/s on 16 threads ({cpu_cells / cpu_s:.3e} cells/s), GPU answers "
            f"identical: {ok}")
    # The EDR roofline: integer VALU issue.  Every DP cell of the batched
    # kernel is EDR_VALU_PER_CELL VALU instructions (counted in the AMD ISA of
    # the built kernel), so the algorithmic lane-ops are cells x that count.
    lane_ops = kernel_rate * EDR_VALU_PER_CELL
    roof = {"bound": "valu", "achieved": float(f"{lane_ops / 1e12:.4g}"), "peak": VALU_PEAK_TOPS,
            "unit": "Tlane-ops/s", "frac": round(lane_ops / 1e12 / VALU_PEAK_TOPS, 4),
            "cells_per_s": float(f"{kernel_rate:.4e}"), "valu_per_cell": EDR_VALU_PER_CELL,
            "kernel": "edr_wave_kernel (all length buckets of one batched launch set)",
            "kernel_ms": round(st_all["kernel_ms"], 4), "cells": cells}
    log(f"EDR roofline: {kernel_rate:.3e} cells/s x {EDR_VALU_PER_CELL} VALU = {lane_ops:.3e} lane-ops/s, "
        f"{roof['frac']:.1%} of {VALU_PEAK_TOPS}e12")
    return {"edr_roofline": roof,
            "edr_plain_cells_per_s": float(f"{plain_rate:.4e}"),
            "edr_batched_cells_per_s": float(f"{batched_rate:.4e}"),
            "edr_batched_kernel_cells_per_s": float(f"{kernel_rate:.4e}"),
            "edr_batched_over_plain": round(batched_rate / plain_rate, 3),
            "edr_batched_equals_plain": same,
            "topk_batched_qps": round(queries.ntraj / topk_s, 2),
            "topk_batch_ms": round(1e3 * topk_s, 3),
            "topk_stage_ms": {"query_upload": round(st["upload_ms"], 3), "filter_kernel": round(st["filter_ms"], 3),
                              "dp_kernels": round(st["kernel_ms"], 3), "host_job_building": round(st["host_ms"], 3),
                              "merge_stage": round(st["merge_ms"], 3)},
            "topk_chain": "device: filter, jobs, DP, survivor lists and merge queued with no host step",
            "topk_pairs_evaluated": int(st["pairs"]),
            "edr_autotuned": tuned,
            "topk_cpu": cpu}


if __name__ == "__main__":
    sys.exit(main())
