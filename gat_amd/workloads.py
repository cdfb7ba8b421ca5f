"""The seeded synthetic workloads of the benchmark configurations.

Each trajectory is a 2-D Gaussian random walk (step sigma 1e-4 degrees per
axis, accumulated in float32, reflected at the edges of the unit box) starting
at a uniform point of [0, 1]^2.  Trajectory ids are 1-based in generation
order.  Seeds are the configuration numbers.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import api

UNIT_BOX = (0.0, 0.0, 1.0, 1.0)


@dataclass
class Workload:
    name: str
    seed: int
    ntraj: int
    length_kind: str            # "uniform" (a..b) or "lognormal" (median a, sigma b)
    length_a: float
    length_b: float
    length_lo: int = 8
    length_hi: int = 2048
    grid_cell: float = 0.01     # GT-quadtree cell size (range = unit box)
    leaf: int = 256             # GT-quadtree and STIG leaf capacity
    fsg_cell: float = 0.01
    rect_n: int = 0
    rect_side_lo: float = 0.01
    rect_side_hi: float = 0.01
    rect_log_uniform: bool = False
    rect_hot_frac: float = 0.0
    rect_nhot: int = 0
    rect_hot_sigma: float = 0.0
    eps: float = 5e-4
    k: int = 10
    extra: dict = field(default_factory=dict)

    def lengths(self, ntraj: int | None = None) -> np.ndarray:
        n = self.ntraj if ntraj is None else ntraj
        return api.synth_lengths(self.seed, n, self.length_kind, self.length_a, self.length_b, self.length_lo,
                                 self.length_hi)

    def db(self, ntraj: int | None = None, threads: int | None = None) -> api.TrajSet:
        return api.synth_walks(self.seed, self.lengths(ntraj), 1e-4, threads)

    def rects(self, n: int | None = None) -> np.ndarray:
        return api.synth_rects(self.seed, self.rect_n if n is None else n, self.rect_side_lo, self.rect_side_hi,
                               self.rect_log_uniform, self.rect_hot_frac, self.rect_nhot, self.rect_hot_sigma)


# Plain method, CPU reference path: 20k trajectories, lengths U{16..256};
# GT-quadtree with 0.01-degree cells and 256-point leaves; 100 squares of side
# 0.01; EDR top-10 of one 128-point query, eps 5e-4.
CFG1 = Workload("cfg1", 1, 20_000, "uniform", 16, 256, 16, 256, grid_cell=0.01, leaf=256, fsg_cell=0.01,
                rect_n=100, rect_side_lo=0.01, rect_side_hi=0.01, k=10, extra={"query_len": 128})

# Plain drop-in on the GPU: 1M trajectories, lengths lognormal (median 64,
# sigma 0.8) clipped to [8, 2048]; 1024x1024 Morton grid (cell 0.001); STIG and
# FSG with 256-point leaves / 1/1024-degree cells; 10k squares with
# log-uniform side in [1e-4, 1e-1].
CFG2 = Workload("cfg2", 2, 1_000_000, "lognormal", 64, 0.8, 8, 2048, grid_cell=0.001, leaf=256,
                fsg_cell=1.0 / 1024, rect_n=10_000, rect_side_lo=1e-4, rect_side_hi=1e-1, rect_log_uniform=True)

# Batched EDR top-k: 200k candidates (lognormal median 128, sigma 0.8, clipped
# to [8, 2048]); 1024 queries with lengths U{16..1024}, every other one a noisy
# copy of a database trajectory; eps 5e-4, k = 50.
CFG3 = Workload("cfg3", 3, 200_000, "lognormal", 128, 0.8, 8, 2048, grid_cell=0.001, leaf=256, k=50,
                extra={"nq": 1024, "qlen_lo": 16, "qlen_hi": 1024})

# Batched mixed range queries: 10M trajectories (cfg2 lengths), 100k squares,
# side log-uniform in [1e-4, 1e-1], 70% of centres around 16 Gaussian hotspots
# (sigma 0.02).  Grid cell 0.00025 (4096x4096 cells).
CFG4 = Workload("cfg4", 4, 10_000_000, "lognormal", 64, 0.8, 8, 2048, grid_cell=0.00025, leaf=256,
                fsg_cell=1.0 / 4096, rect_n=100_000, rect_side_lo=1e-4, rect_side_hi=1e-1, rect_log_uniform=True,
                rect_hot_frac=0.7, rect_nhot=16, rect_hot_sigma=0.02)

# Serving loop: about 2B resident points (22.7M trajectories with cfg2
# lengths); Poisson arrivals at 2000 req/s, 80% range (side log-uniform in
# [1e-4, 5e-2]) and 20% EDR top-20 (query length U{16..512}, eps 5e-4).
CFG5 = Workload("cfg5", 5, 22_700_000, "lognormal", 64, 0.8, 8, 2048, grid_cell=0.00025, leaf=256,
                rect_side_lo=1e-4, rect_side_hi=5e-2, rect_log_uniform=True, k=20,
                extra={"rate": 2000.0, "range_frac": 0.8, "qlen_lo": 16, "qlen_hi": 512})

CONFIGS = {w.name: w for w in (CFG1, CFG2, CFG3, CFG4, CFG5)}


def serve_requests(w: Workload, db: api.TrajSet, seconds: float, rate: float | None = None, seed: int | None = None):
    """A Poisson request stream for the serving loop: arrivals at `rate`
    requests/s for `seconds`, a fraction extra["range_frac"] of them range
    squares (side log-uniform in [rect_side_lo, rect_side_hi], uniform
    centres), the rest top-k queries with lengths U{qlen_lo..qlen_hi}, every
    other one a noisy copy of a database trajectory.  Returns (requests,
    query lon, query lat)."""
    rate = w.extra["rate"] if rate is None else rate
    rng = np.random.default_rng(w.seed * 1000 + 17 if seed is None else seed)
    n = max(int(rate * seconds), 1)
    arrival = np.cumsum(rng.exponential(1e3 / rate, n))
    kind = np.where(rng.random(n) < w.extra["range_frac"], api.REQ_RANGE, api.REQ_TOPK).astype(np.int32)
    req = np.zeros(n, api.REQUEST)
    req["kind"] = kind
    req["arrival_ms"] = arrival
    nr = int((kind == api.REQ_RANGE).sum())
    if nr:
        req["box"][kind == api.REQ_RANGE] = api.synth_rects(int(rng.integers(1 << 30)), nr, w.rect_side_lo,
                                                            w.rect_side_hi, True)
    nt = n - nr
    qlon = np.zeros(0, np.float32)
    qlat = np.zeros(0, np.float32)
    if nt:
        lo, hi = w.extra["qlen_lo"], w.extra["qlen_hi"]
        qlen = api.synth_lengths(int(rng.integers(1 << 30)), nt, "uniform", lo, hi, lo, hi)
        q = api.synth_queries(int(rng.integers(1 << 30)), qlen, db, copy_mode=2)
        qlon, qlat = q.lon, q.lat
        req["qlen"][kind == api.REQ_TOPK] = qlen
        req["qoff"][kind == api.REQ_TOPK] = q.offsets[:-1]
    return req, qlon, qlat
