"""numpy-facing wrappers over the native libraries.

Arrays cross the boundary as plain pointers; every wrapper checks dtype and
contiguity first.  Trajectory ids are 1-based (trajectory t is
``offsets[t-1]:offsets[t]`` of ``lon``/``lat``); a point id is the 0-based
position of the point within its trajectory.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _native as N


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _req(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def default_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


def init(device: int = 0) -> None:
    """Selects the GPU; raises if no device or the native library is missing."""
    N.check_host(N.host().gath_init(device))


def device_name() -> str:
    buf = C.create_string_buffer(256)
    N.check_hip(N.hip().gat_device_name(buf, 256))
    return buf.value.decode()


def eps_threshold(eps: float) -> float:
    return float(N.hip().gat_eps_threshold(eps))


def hbm_pattern_bandwidth(kind: int, nbytes: int = 1 << 32, run: int = 0, reps: int = 10) -> tuple[float, float]:
    """Bandwidth of one HBM access pattern (gat_hbm_pattern: 0 write stream,
    1 read stream, 2 packed runs of `run` pairs, 3 the same runs through
    line-aligned windows, 4 and 5 the same two with cached stores): (GB/s of
    the bytes the pattern covers, mean kernel ms), timed with HIP events on
    its own stream."""
    h = N.hip()
    buf, s, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    N.check_hip(h.gat_malloc(C.byref(buf), nbytes))
    try:
        N.check_hip(h.gat_stream_create(C.byref(s)))
        N.check_hip(h.gat_event_create(C.byref(e0)))
        N.check_hip(h.gat_event_create(C.byref(e1)))
        N.check_hip(h.gat_memset(buf, 0, nbytes, s))
        N.check_hip(h.gat_hbm_pattern(kind, buf, nbytes, run, s))
        N.check_hip(h.gat_event_record(e0, s))
        for _ in range(reps):
            N.check_hip(h.gat_hbm_pattern(kind, buf, nbytes, run, s))
        N.check_hip(h.gat_event_record(e1, s))
        N.check_hip(h.gat_event_sync(e1))
        ms = C.c_float()
        N.check_hip(h.gat_event_elapsed_ms(C.byref(ms), e0, e1))
        per = ms.value / reps
        covered = nbytes
        if kind >= 2:
            nruns = 1
            while nruns * 2 * run * 8 <= nbytes:
                nruns *= 2
            covered = nruns * run * 8
        return covered / (per * 1e-3) / 1e9, per
    finally:
        for e in (e0, e1):
            if e.value:
                h.gat_event_destroy(e)
        if s.value:
            h.gat_stream_destroy(s)
        h.gat_free(buf)


def _table(handle) -> tuple["TrajSet", np.ndarray, np.ndarray, np.ndarray]:
    if not handle:
        raise N.GatError(2, N.host().gath_last_error().decode(errors="replace"))
    try:
        sizes = np.zeros(2, np.int64)
        bounds = np.zeros(4, np.float32)
        N.check_host(N.host().gath_traj_info(handle, _ptr(sizes), _ptr(bounds)))
        n, m = int(sizes[0]), int(sizes[1])
        offsets = np.zeros(n + 1, np.int64)
        lon, lat = np.zeros(max(m, 1), np.float32), np.zeros(max(m, 1), np.float32)
        tid, tim = np.zeros(max(m, 1), np.int32), np.zeros(max(m, 1), np.int32)
        N.check_host(N.host().gath_traj_copy(handle, _ptr(offsets), _ptr(lon), _ptr(lat), _ptr(tid), _ptr(tim)))
        return TrajSet(offsets, lon[:m], lat[:m]), tid[:m], tim[:m], bounds
    finally:
        N.host().gath_traj_free(handle)


def read_trajectory_file(path: str) -> tuple["TrajSet", np.ndarray, np.ndarray, np.ndarray]:
    """The reference's preprocessed trajectory file (PreProcess.cpp:225-315):
    (trajectories, per-point tid field, per-point time, bounds xmin, ymin,
    xmax, ymax)."""
    return _table(N.host().gath_traj_read(str(path).encode()))


def ingest_vehicle_csv(path: str, base_date: str) -> tuple["TrajSet", np.ndarray, np.ndarray, np.ndarray]:
    """Trajectories from raw vehicle records (PreProcess.cpp:48-201), as
    read_trajectory_file returns them; times in seconds from base_date."""
    return _table(N.host().gath_traj_ingest_csv(str(path).encode(), base_date.encode()))


def write_trajectory_file(path: str, db: "TrajSet", tid: np.ndarray | None = None, time: np.ndarray | None = None,
                          bounds: np.ndarray | None = None) -> None:
    """Writes the preprocessed format (PreProcess.cpp:203-223).  Defaults: tid
    field = the trajectory id, time = 0, bounds = the points' box."""
    m = db.npoints
    if tid is None:
        tid = np.repeat(np.arange(1, db.ntraj + 1, dtype=np.int32), np.diff(db.offsets))
    if time is None:
        time = np.zeros(m, np.int32)
    if bounds is None:
        bounds = (np.array([db.lon.min(), db.lat.min(), db.lon.max(), db.lat.max()], np.float32) if m else
                  np.zeros(4, np.float32))
    N.check_host(N.host().gath_traj_write(str(path).encode(), _ptr(_req(db.offsets, np.int64)), db.ntraj,
                                          _ptr(_req(db.lon, np.float32)), _ptr(_req(db.lat, np.float32)),
                                          _ptr(_req(tid, np.int32)), _ptr(_req(time, np.int32)),
                                          _ptr(_req(bounds, np.float32))))


def stream_copy_bandwidth(nbytes: int = 1 << 32, reps: int = 10) -> tuple[float, float]:
    """HBM bandwidth of the plain streaming-copy kernel: (GB/s counting read +
    write bytes, mean kernel ms), timed with HIP events on its own stream."""
    h = N.hip()
    src, dst, s, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    N.check_hip(h.gat_malloc(C.byref(src), nbytes))
    N.check_hip(h.gat_malloc(C.byref(dst), nbytes))
    try:
        N.check_hip(h.gat_stream_create(C.byref(s)))
        N.check_hip(h.gat_event_create(C.byref(e0)))
        N.check_hip(h.gat_event_create(C.byref(e1)))
        N.check_hip(h.gat_memset(src, 1, nbytes, s))
        N.check_hip(h.gat_stream_copy(dst, src, nbytes, s))
        N.check_hip(h.gat_event_record(e0, s))
        for _ in range(reps):
            N.check_hip(h.gat_stream_copy(dst, src, nbytes, s))
        N.check_hip(h.gat_event_record(e1, s))
        N.check_hip(h.gat_event_sync(e1))
        ms = C.c_float()
        N.check_hip(h.gat_event_elapsed_ms(C.byref(ms), e0, e1))
        per = ms.value / reps
        return 2 * nbytes / (per * 1e-3) / 1e9, per
    finally:
        for e in (e0, e1):
            if e.value:
                h.gat_event_destroy(e)
        if s.value:
            h.gat_stream_destroy(s)
        h.gat_free(src)
        h.gat_free(dst)


# ---------------------------------------------------------------- workloads
@dataclass
class TrajSet:
    offsets: np.ndarray  # int64, ntraj + 1
    lon: np.ndarray      # float32
    lat: np.ndarray      # float32

    @property
    def ntraj(self) -> int:
        return len(self.offsets) - 1

    @property
    def npoints(self) -> int:
        return int(self.offsets[-1])

    @property
    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets).astype(np.int32)

    def xy(self, t: int) -> np.ndarray:
        """Interleaved (lon, lat) float32 points of trajectory t (1-based)."""
        b, e = self.offsets[t - 1], self.offsets[t]
        out = np.empty((e - b, 2), np.float32)
        out[:, 0] = self.lon[b:e]
        out[:, 1] = self.lat[b:e]
        return out

    def interleaved(self) -> np.ndarray:
        out = np.empty((self.npoints, 2), np.float32)
        out[:, 0] = self.lon
        out[:, 1] = self.lat
        return out

    @staticmethod
    def from_lengths(lengths: np.ndarray) -> "TrajSet":
        off = np.zeros(len(lengths) + 1, np.int64)
        np.cumsum(lengths, out=off[1:])
        return TrajSet(off, np.empty(int(off[-1]), np.float32), np.empty(int(off[-1]), np.float32))


def synth_lengths(seed: int, n: int, kind: str, a: float, b: float, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    h = N.host()
    if kind == "uniform":
        h.gath_synth_lengths_uniform(seed, n, int(a), int(b), _ptr(out))
    elif kind == "lognormal":
        h.gath_synth_lengths_lognormal(seed, n, float(a), float(b), lo, hi, _ptr(out))
    else:
        raise ValueError(kind)
    return out


def synth_walks(seed: int, lengths: np.ndarray, step_sigma: float = 1e-4, threads: int | None = None) -> TrajSet:
    ts = TrajSet.from_lengths(_req(lengths, np.int32))
    N.host().gath_synth_walks(seed, ts.ntraj, _ptr(ts.offsets), step_sigma, _ptr(ts.lon), _ptr(ts.lat),
                              threads or default_threads())
    return ts


def synth_rects(seed: int, n: int, side_lo: float, side_hi: float, log_uniform: bool = True,
                hot_frac: float = 0.0, nhot: int = 0, hot_sigma: float = 0.0) -> np.ndarray:
    out = np.empty((n, 4), np.float32)
    N.host().gath_synth_rects(seed, n, side_lo, side_hi, int(log_uniform), hot_frac, nhot, hot_sigma, _ptr(out))
    return out


def synth_queries(seed: int, lengths: np.ndarray, db: TrajSet | None, copy_mode: int,
                  noise_sigma: float = 2e-4, step_sigma: float = 1e-4) -> TrajSet:
    q = TrajSet.from_lengths(_req(lengths, np.int32))
    qlen = q.lengths
    if db is None:
        db = TrajSet(np.zeros(1, np.int64), np.zeros(0, np.float32), np.zeros(0, np.float32))
    N.host().gath_synth_queries(seed, q.ntraj, _ptr(qlen), _ptr(q.offsets), _ptr(db.offsets), db.ntraj,
                                _ptr(db.lon), _ptr(db.lat), copy_mode, noise_sigma, step_sigma,
                                _ptr(q.lon), _ptr(q.lat))
    return q


# ---------------------------------------------------------------- indexes
# gath_request / gath_outcome (include/gat_host.h).
REQUEST = np.dtype([("kind", np.int32), ("box", np.float32, 4), ("qlen", np.int32), ("qoff", np.int64),
                    ("arrival_ms", np.float64)], align=True)
OUTCOME = np.dtype([("admit_ms", np.float64), ("done_ms", np.float64), ("batch", np.int32), ("count", np.uint32),
                    ("digest", np.uint64)], align=True)
REQ_RANGE, REQ_TOPK = 0, 1

# gat_tree_node (include/gat_hip.h).
TREE_NODE = np.dtype([("box", np.float32, 4), ("first_child", np.int32), ("start", np.uint32), ("count", np.uint32),
                      ("flags", np.uint32)])
NODE_INCLUSIVE, NODE_UNTESTED = 1, 2
RUN = np.dtype([("q", np.uint32), ("count", np.uint32), ("first", np.uint64)])  # gat_range_run


class RangeResult:
    """A compacted-pair result that stays in HBM until the index's next query."""

    def __init__(self, index: "_Index", nq: int, stats: np.ndarray, ms: np.ndarray):
        self._index = index
        self.nq = nq
        self.pairs_total = int(stats[0])
        self.candidates = int(stats[1])
        self.tasks = int(stats[2])
        self.select_ms, self.upload_ms, self.kernel_ms = (float(x) for x in ms)
        # Batched form: (leaf, query) entries, leaf-major work items, graph replay.
        self.leaf_queries = int(stats[2]) if len(stats) > 3 else 0
        self.work_items = int(stats[3]) if len(stats) > 3 else 0
        self.replayed = bool(stats[4]) if len(stats) > 4 else False

    def counts(self) -> np.ndarray:
        out = np.empty(self.nq, np.int64)
        N.check_host(N.host().gath_index_last_counts(self._index.handle, _ptr(out)))
        return out

    def pairs(self) -> np.ndarray:
        """(pairs_total, 2) int32 (tid, pid), query by query."""
        out = np.empty((max(self.pairs_total, 1), 2), np.int32)
        N.check_host(N.host().gath_index_last_fetch(self._index.handle, _ptr(out), self.pairs_total))
        return out[: self.pairs_total]

    def runs(self) -> np.ndarray:
        """Run-indexed batches: the run records (RUN dtype) of the result's
        compacted pairs; empty for per-query layouts."""
        n = int(N.host().gath_index_last_runs(self._index.handle, None, 0))
        if n < 0:
            N.check_host(n)
        out = np.zeros(max(n, 1), RUN)
        if n > 0:
            got = int(N.host().gath_index_last_runs(self._index.handle, out.ctypes.data, n))
            if got < 0:
                N.check_host(got)
        return out[:n]

    def per_query(self) -> list[np.ndarray]:
        c = self.counts()
        p = self.pairs()
        cut = np.concatenate([[0], np.cumsum(c)])
        return [p[cut[i]:cut[i + 1]] for i in range(self.nq)]

    def digests(self) -> np.ndarray:
        out = np.empty(self.nq, np.uint64)
        N.check_host(N.host().gath_index_last_digests(self._index.handle, _ptr(out)))
        return out


class _Index:
    handle = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            N.host().gath_index_free(h)
            self.handle = None

    def build(self, db: TrajSet, upload: bool = True, device: bool = False) -> "_Index":
        """Builds the index over `db`; upload=False keeps it on the host.
        device=True builds the store on the GPU (Grid, MortonGrid and FSG: the
        same cell-ordered store; STIG: the kd order, same tree shape)."""
        N.check_host(N.host().gath_index_build(self.handle, _ptr(_req(db.offsets, np.int64)), db.ntraj,
                                               _ptr(_req(db.lon, np.float32)), _ptr(_req(db.lat, np.float32)),
                                               (2 if device else 1) if upload else 0))
        self.ntraj = db.ntraj
        return self

    def set_threads(self, n: int) -> None:
        N.check_host(N.host().gath_index_set_threads(self.handle, n))

    def info(self) -> np.ndarray:
        out = np.zeros(8, np.int64)
        N.check_host(N.host().gath_index_info(self.handle, _ptr(out)))
        return out

    def cell_of(self, lon: float, lat: float) -> int:
        return int(N.host().gath_index_cell_of(self.handle, lon, lat))

    def select(self, boxes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Host candidate selection: (tasks (n, 3) uint32 [start, count, qid], candidates per query)."""
        b = _req(boxes, np.float32).reshape(-1, 4)
        cand = np.empty(len(b), np.uint64)
        n = N.host().gath_index_select(self.handle, _ptr(b), len(b), None, 0, _ptr(cand))
        if n < 0:
            raise N.GatError(2, "selection failed")
        tasks = np.empty((max(n, 1), 3), np.uint32)
        n2 = N.host().gath_index_select(self.handle, _ptr(b), len(b), _ptr(tasks), n, _ptr(cand))
        assert n2 == n
        return tasks[:n], cand

    def range_bitmap(self, boxes: np.ndarray) -> np.ndarray:
        b = _req(boxes, np.float32).reshape(-1, 4)
        out = np.empty((len(b), self.ntraj + 1), np.uint8)
        N.check_host(N.host().gath_index_range_bitmap(self.handle, _ptr(b), len(b), _ptr(out)))
        return out

    def range_pairs(self, boxes: np.ndarray) -> RangeResult:
        b = _req(boxes, np.float32).reshape(-1, 4)
        stats = np.zeros(3, np.int64)
        ms = np.zeros(3, np.float64)
        N.check_host(N.host().gath_index_range_pairs(self.handle, _ptr(b), len(b), _ptr(stats), _ptr(ms)))
        return RangeResult(self, len(b), stats, ms)

    def range_batched(self, boxes: np.ndarray, graph: bool = False) -> RangeResult:
        """Batched form (Grid, STIG, FSG): device selection plus the leaf-major join
        for the whole batch; graph=True captures the device pipeline and replays
        it on later calls that fit.  Same result sets as range_pairs."""
        b = _req(boxes, np.float32).reshape(-1, 4)
        stats = np.zeros(5, np.int64)
        ms = np.zeros(4, np.float64)
        N.check_host(N.host().gath_index_range_batched(self.handle, _ptr(b), len(b), int(graph), _ptr(stats),
                                                       _ptr(ms)))
        res = RangeResult(self, len(b), stats, ms[:3])
        res.refine_ms = float(ms[3])
        return res

    def last_unique_candidates(self) -> int:
        """Points of the leaves the last batched run selected at least once."""
        out = np.zeros(1, np.int64)
        N.check_host(N.host().gath_index_last_unique_candidates(self.handle, _ptr(out)))
        return int(out[0])

    def set_result_layout(self, run_indexed: bool = True) -> "_Index":
        """Leaf-major result layout: pairs by work item with a run index
        (default), or per-query regions.  Counts, digests and fetched pairs are
        the same either way."""
        N.check_host(N.host().gath_index_set_result_layout(self.handle, int(run_indexed)))
        return self

    def system_range_test(self, which: str, boxes: np.ndarray) -> tuple[int, int]:
        """The reference's SystemTest range entry point `which` ("grid",
        "stig", "fsg", "mgrid": rangeQueryTest, STIGrangeQueryTest,
        FSGrangeQueryTest, MortonGridRangeQueryTest) on this index: (candidate
        points scanned, result pairs)."""
        kinds = {"grid": 0, "stig": 1, "fsg": 2, "mgrid": 3}
        b = _req(boxes, np.float32).reshape(-1, 4)
        out = np.zeros(2, np.int64)
        N.check_host(N.host().gath_system_range_test(self.handle, kinds[which], _ptr(b), len(b), _ptr(out)))
        self._system_last = (len(b), int(out[0]), int(out[1]))
        return int(out[0]), int(out[1])

    def system_range_result(self) -> RangeResult:
        """The result of the last system_range_test call (its pairs stay in
        HBM until the index's next query): counts, pairs and digests."""
        nq, cand, pairs = self._system_last
        return RangeResult(self, nq, np.array([pairs, cand, 0], np.int64), np.zeros(3))

    def set_batch_params(self, queries_per_item: int = 32, direct_max_queries: int = 256) -> "_Index":
        """Batched form: rectangles per leaf-major work item (1..64), and the
        largest batch that takes the direct form (per-task refine, nothing
        sized by the tree touched)."""
        N.check_host(N.host().gath_index_set_batch_params(self.handle, int(queries_per_item),
                                                          int(direct_max_queries)))
        return self

    def tune_batched(self, boxes: np.ndarray, reps: int = 3) -> dict:
        """Autotunes the batched form on `boxes` (device pipeline time, median
        of reps): rectangles per work item, then leaf-major against direct at
        this batch size; keeps the fastest."""
        b = _req(boxes, np.float32).reshape(-1, 4)
        out = np.zeros(8, np.float64)
        N.check_host(N.host().gath_tune_range(self.handle, _ptr(b), len(b), int(reps), _ptr(out)))
        return {"queries_per_item": int(out[0]),
                "ms": dict(zip([4, 8, 16, 32, 64], [round(float(x), 4) for x in out[1:6]])),
                "direct_ms": round(float(out[6]), 4), "direct_max_queries": int(out[7])}

    def fetch_store(self) -> tuple[np.ndarray, np.ndarray, np.ndarray | None, int]:
        """The resident point store: (xy (n, 2) float32, id uint32, pid uint16 or
        None, pid_bits)."""
        meta = np.zeros(2, np.int64)
        N.check_host(N.host().gath_index_fetch_store(self.handle, None, None, None, _ptr(meta)))
        n, bits = int(meta[0]), int(meta[1])
        xy = np.zeros((max(n, 1), 2), np.float32)
        ids = np.zeros(max(n, 1), np.uint32)
        pid = np.zeros(max(n, 1), np.uint16) if bits == 0 else None
        if n > 0:
            N.check_host(N.host().gath_index_fetch_store(self.handle, xy.ctypes.data, ids.ctypes.data,
                                                         pid.ctypes.data if pid is not None else None, _ptr(meta)))
        return xy[:n], ids[:n], (pid[:n] if pid is not None else None), bits

    def cell_starts(self) -> np.ndarray:
        """Grid / FSG: the store offset of every cell (cells + 1 entries)."""
        n = int(N.host().gath_index_cell_starts(self.handle, None, 0))
        if n < 0:
            N.check_host(n)
        out = np.zeros(max(n, 1), np.int64)
        N.host().gath_index_cell_starts(self.handle, out.ctypes.data, n)
        return out[:n]

    def tree_nodes(self) -> tuple[np.ndarray, int]:
        """The tree the batched form uploads: a structured array of gat_tree_node
        in breadth-first order, and its depth."""
        depth = np.zeros(1, np.int32)
        n = int(N.host().gath_index_tree_nodes(self.handle, None, 0, _ptr(depth)))
        if n < 0:
            N.check_host(n)
        out = np.zeros(max(n, 1), TREE_NODE)
        N.host().gath_index_tree_nodes(self.handle, out.ctypes.data, n, _ptr(depth))
        return out[:n], int(depth[0])

    def start_table(self) -> tuple[np.ndarray, int, np.ndarray]:
        """The start table a GT-quadtree gives its device walk: per block of
        the table's level (Morton order) the owning node as (level << 27) |
        breadth-first index, the level, and (xmin, ymax, block side); an empty
        table (level 0) when the index has none."""
        level = np.zeros(1, np.int32)
        geom = np.zeros(3, np.float32)
        n = int(N.host().gath_index_start_table(self.handle, None, 0, _ptr(level), _ptr(geom)))
        if n < 0:
            N.check_host(n)
        out = np.zeros(max(n, 1), np.uint32)
        N.host().gath_index_start_table(self.handle, out.ctypes.data, n, _ptr(level), _ptr(geom))
        return out[:n], int(level[0]), geom


class GridIndex(_Index):
    """GT-quadtree (Grid.cpp): Morton-ordered cells, quadtree leaves < max_points_in_node."""

    def __init__(self, range_box=(0.0, 0.0, 1.0, 1.0), cell_size: float = 0.01, max_points_in_node: int = 256):
        r = np.asarray(range_box, np.float32)
        self.handle = N.host().gath_grid_new(_ptr(r), cell_size, max_points_in_node)

    def set_edr_params(self, split: int = 512, per_cu_short: int = 0, per_cu_long: int = 0) -> "GridIndex":
        """Batched-EDR length split between the short and long buckets, and
        workgroups per CU of each bucket's launch (0: kernel default)."""
        N.check_host(N.host().gath_grid_set_edr_params(self.handle, int(split), int(per_cu_short),
                                                       int(per_cu_long)))
        return self

    def tune_edr(self, queries: TrajSet, eps: float, reps: int = 2) -> dict:
        """Autotunes batched EDR (bucket split, workgroups per CU of each
        bucket) on every pair of `queries` x the index's trajectories."""
        out = np.zeros(7, np.float64)
        N.check_host(N.host().gath_tune_edr(self.handle, _ptr(_req(queries.offsets, np.int64)), queries.ntraj,
                                            _ptr(_req(queries.lon, np.float32)), _ptr(_req(queries.lat, np.float32)),
                                            eps, int(reps), _ptr(out)))
        return {"split": int(out[0]), "per_cu_short": int(out[1]), "per_cu_long": int(out[2]),
                "kernel_ms": round(float(out[3]), 4), "evaluations": int(out[4]),
                "waves_per_long_pair": int(out[5]), "one_round_ms": round(float(out[6]), 4)}

    def set_edr_waves(self, waves_per_pair: int, every_long_pair: bool = False) -> "GridIndex":
        """Waves per long pair (1, 2, 4, 8) for batches that fit one round:
        the pairs one wave would keep busy past the batch's longest pair on
        that many waves take them, the rest one wave each; every_long_pair
        gives them to every long pair."""
        waves_per_pair = -int(waves_per_pair) if every_long_pair and waves_per_pair > 1 else waves_per_pair
        N.check_host(N.host().gath_grid_set_edr_waves(self.handle, int(waves_per_pair)))
        return self

    def set_topk_device_chain(self, on: bool) -> "GridIndex":
        """Top-k batches chained on the device (filter, jobs, DP, survivor
        lists, merge; the default) or with jobs and lists built on the host."""
        N.check_host(N.host().gath_grid_set_topk_device_chain(self.handle, int(bool(on))))
        return self

    def edr_all(self, q_xy: np.ndarray, eps: float) -> np.ndarray:
        q = _req(q_xy, np.float32).reshape(-1, 2)
        lon, lat = np.ascontiguousarray(q[:, 0]), np.ascontiguousarray(q[:, 1])
        out = np.empty(self.ntraj, np.int32)
        N.check_host(N.host().gath_grid_edr_all(self.handle, _ptr(lon), _ptr(lat), len(q), eps, _ptr(out)))
        return out

    def similarity(self, queries: TrajSet, k: int, eps: float) -> tuple[np.ndarray, np.ndarray]:
        ids = np.empty((queries.ntraj, k), np.int32)
        dists = np.empty((queries.ntraj, k), np.int32)
        N.check_host(N.host().gath_grid_similarity(self.handle, _ptr(_req(queries.offsets, np.int64)), queries.ntraj,
                                                   _ptr(_req(queries.lon, np.float32)),
                                                   _ptr(_req(queries.lat, np.float32)), k, eps, _ptr(ids),
                                                   _ptr(dists)))
        return ids, dists

    def system_similarity_test(self, query_xy: np.ndarray, scale: int, k: int, eps: float, cpu=None) -> dict:
        """The reference's SystemTest::similarityQueryTest (SystemTest.cpp:164-249):
        the query replicated `scale` times, the CPU leg timed (`cpu(queries:
        TrajSet, k) -> ids (nq, k)`, skipped when None: this library has no
        CPU query path), then the GPU leg, Grid::SimilarityQueryBatchOnGPU.
        Returns {"gpu_ids", "cpu_ids", "cpu_ms", "gpu_ms"}."""
        q = _req(query_xy, np.float32).reshape(-1, 2)
        qlon, qlat = np.ascontiguousarray(q[:, 0]), np.ascontiguousarray(q[:, 1])
        gpu_ids = np.zeros((scale, k), np.int32)
        cpu_ids = np.zeros((scale, k), np.int32)
        ms = np.zeros(2, np.float64)
        fn_t = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int64), C.c_int, C.POINTER(C.c_float),
                           C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int32))
        failure = []

        def leg(_ctx, offs, nq, lon, lat, kk, ids):
            try:
                o = np.ctypeslib.as_array(offs, (nq + 1,)).copy()
                m = int(o[-1])
                qs = TrajSet(o, np.ctypeslib.as_array(lon, (m,)).copy(), np.ctypeslib.as_array(lat, (m,)).copy())
                got = np.asarray(cpu(qs, kk), np.int32).reshape(nq, kk)
                np.ctypeslib.as_array(ids, (nq * kk,))[:] = got.reshape(-1)
                return 0
            except Exception as e:  # noqa: BLE001 - reported after the C call returns
                failure.append(e)
                return 1

        cb = fn_t(leg) if cpu is not None else None
        rc = N.host().gath_system_similarity_test(self.handle, _ptr(qlon), _ptr(qlat), len(q), int(scale), int(k),
                                                  eps, C.cast(cb, C.c_void_p) if cb else None, None, _ptr(gpu_ids),
                                                  _ptr(cpu_ids) if cpu is not None else None, _ptr(ms))
        if failure:
            raise failure[0]
        N.check_host(rc)
        return {"gpu_ids": gpu_ids, "cpu_ids": cpu_ids if cpu is not None else None, "cpu_ms": float(ms[0]),
                "gpu_ms": float(ms[1])}

    def edr_batched_all(self, queries: TrajSet, eps: float) -> tuple[np.ndarray, dict]:
        """Distances of every (query, trajectory) pair, (nq, ntraj) int32, from
        one persistent batched launch."""
        dist = np.empty((queries.ntraj, self.ntraj), np.int32)
        stats = np.zeros(5, np.int64)
        ms = np.zeros(5, np.float64)
        N.check_host(N.host().gath_grid_edr_batched_all(
            self.handle, _ptr(_req(queries.offsets, np.int64)), queries.ntraj, _ptr(_req(queries.lon, np.float32)),
            _ptr(_req(queries.lat, np.float32)), eps, _ptr(dist), _ptr(stats), _ptr(ms)))
        return dist, _edr_stats(stats, ms)

    def edr_batched_pairs(self, queries: TrajSet, pairs: np.ndarray, eps: float) -> tuple[np.ndarray, dict]:
        """Distances of given (query index, trajectory index) pairs, both
        0-based, (n, 2) int32, from the batched kernels."""
        p = _req(pairs, np.int32).reshape(-1, 2)
        dist = np.empty(max(len(p), 1), np.int32)
        stats = np.zeros(5, np.int64)
        ms = np.zeros(5, np.float64)
        N.check_host(N.host().gath_grid_edr_batched_pairs(
            self.handle, _ptr(_req(queries.offsets, np.int64)), queries.ntraj, _ptr(_req(queries.lon, np.float32)),
            _ptr(_req(queries.lat, np.float32)), _ptr(p), len(p), eps, _ptr(dist), _ptr(stats), _ptr(ms)))
        return dist[:len(p)], _edr_stats(stats, ms)

    def similarity_batched(self, queries: TrajSet, k: int, eps: float) -> tuple[np.ndarray, np.ndarray, dict]:
        """Exact top-k (ascending distance, then id) of every query: the candidate
        filter resolves most pairs through the all-mismatch table, the rest run
        the batched DP."""
        ids = np.empty((queries.ntraj, k), np.int32)
        dists = np.empty((queries.ntraj, k), np.int32)
        stats = np.zeros(5, np.int64)
        ms = np.zeros(5, np.float64)
        N.check_host(N.host().gath_grid_similarity_batched(
            self.handle, _ptr(_req(queries.offsets, np.int64)), queries.ntraj, _ptr(_req(queries.lon, np.float32)),
            _ptr(_req(queries.lat, np.float32)), k, eps, _ptr(ids), _ptr(dists), _ptr(stats), _ptr(ms)))
        return ids, dists, _edr_stats(stats, ms)

    def mismatch_table(self) -> np.ndarray:
        w = MAX_TRAJ_LEN + 1
        out = np.empty((w, w), np.int32)
        N.check_host(N.host().gath_grid_mismatch_table(self.handle, _ptr(out)))
        return out


MAX_TRAJ_LEN = 2048


def _edr_stats(stats, ms) -> dict:
    return {"pairs": int(stats[0]), "cells": int(stats[1]), "jobs": int(stats[2]), "pruned": int(stats[3]),
            "kernel_ms": float(ms[0]), "filter_ms": float(ms[1]), "host_ms": float(ms[2]),
            "upload_ms": float(ms[3]), "merge_ms": float(ms[4])}


def serve(grid: "GridIndex", requests: np.ndarray, qlon: np.ndarray, qlat: np.ndarray, k: int, eps: float,
          batched: bool = True, range_slots: int = 2, max_range_batch: int = 4096, max_topk_batch: int = 256,
          range_budget: float = 4e9, topk_budget: float = 1e9):
    """Replays a request stream (REQUEST records, arrival times ascending)
    through the serving loop.  Returns (outcomes, topk ids, topk distances,
    stats dict); top-k rows are meaningful for top-k requests only."""
    if requests.dtype != REQUEST:
        raise TypeError("requests must have the REQUEST dtype")
    r = np.ascontiguousarray(requests)
    n = len(r)
    lon = _req(qlon, np.float32)
    lat = _req(qlat, np.float32)
    if len(lon) != len(lat):
        raise ValueError("qlon and qlat must have the same length")
    topk = r["kind"] == 1
    if topk.any() and ((r["qoff"][topk] < 0) | (r["qoff"][topk] + r["qlen"][topk] > len(lon))).any():
        raise ValueError("a top-k request reads past the query point arrays")
    out = np.zeros(max(n, 1), OUTCOME)
    ids = np.zeros((max(n, 1), k), np.int32)
    dists = np.zeros((max(n, 1), k), np.int32)
    cfg = np.array([range_slots, max_range_batch, max_topk_batch, range_budget, topk_budget, k, eps, int(batched)],
                   np.float64)
    st = np.zeros(11, np.float64)
    N.check_host(N.host().gath_serve(grid.handle, r.ctypes.data, n, _ptr(lon), _ptr(lat), len(lon), _ptr(cfg),
                                     out.ctypes.data, _ptr(ids), _ptr(dists), _ptr(st)))
    keys = ["range_batches", "topk_batches", "range_done", "topk_done", "wall_ms", "peak_range_bytes",
            "peak_topk_bytes", "range_budget", "topk_budget", "peak_topk_survivors", "device_allocs"]
    return out[:n], ids[:n], dists[:n], dict(zip(keys, st.tolist()))


class MixedRange:
    """One batch over the three indexes of a data set: each rectangle routed by
    area (FSG below fsg_max_area, GT-quadtree below grid_max_area, STIG above),
    in groups of group_size per index; the three device pipelines of a group run
    as one captured graph.  Results: per-rectangle counts and digests."""

    def __init__(self, grid: "GridIndex", stig: "StigIndex", fsg: "FsgIndex"):
        self._keep = (grid, stig, fsg)
        self.handle = N.host().gath_mixed_new(grid.handle, stig.handle, fsg.handle)
        if not self.handle:
            N.check_host(2)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            N.host().gath_mixed_free(h)
            self.handle = None

    def run(self, boxes: np.ndarray, fsg_max_area: float = 1e-6, grid_max_area: float = 1e-3,
            group_size: int = 2048, graph: bool = True):
        b = _req(boxes, np.float32).reshape(-1, 4)
        n = len(b)
        counts = np.zeros(max(n, 1), np.uint32)
        digests = np.zeros(max(n, 1), np.uint64)
        route = np.zeros(max(n, 1), np.int8)
        st = np.zeros(7, np.float64)
        cfg = np.array([fsg_max_area, grid_max_area, group_size, int(graph)], np.float64)
        N.check_host(N.host().gath_mixed_run(self.handle, _ptr(b), n, _ptr(cfg), _ptr(counts), _ptr(digests),
                                             _ptr(route), _ptr(st)))
        keys = ["pairs", "candidates", "groups", "replays", "ms", "digest_ms", "captures"]
        return counts[:n], digests[:n], route[:n], dict(zip(keys, st.tolist()))


class MortonGridIndex(GridIndex):
    """MortonGrid (MortonGrid.cpp): the GT-quadtree kept implicitly — leaf
    flags per (level, id), node boxes from their cells; the batched form walks
    it on the device without node records.  Same answers as GridIndex."""

    def __init__(self, range_box=(0.0, 0.0, 1.0, 1.0), cell_size: float = 0.01, max_points_in_node: int = 256):
        r = np.asarray(range_box, np.float32)
        self.handle = N.host().gath_morton_grid_new(_ptr(r), cell_size, max_points_in_node)


class FsgIndex(_Index):
    """FSG (FSG.cpp): flat row-major grid."""

    def __init__(self, range_box=(0.0, 0.0, 1.0, 1.0), cell_size: float = 0.01):
        r = np.asarray(range_box, np.float32)
        self.handle = N.host().gath_fsg_new(_ptr(r), cell_size)


class StigIndex(_Index):
    """STIG (STIG.cpp): kd-split leaves of at most block_size points."""

    def __init__(self, block_size: int = 256):
        self.handle = N.host().gath_stig_new(block_size)
