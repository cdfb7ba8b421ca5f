"""TEST INFRASTRUCTURE ONLY — ctypes binding of the parity oracle.

liboracle.so (oracle/oracle.cpp, built by oracle/Makefile) restates the
reference's CPU path; libref_mbb.so (oracle/_ref, built only where the reference
sources are present) is the reference's own MBB.cpp.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker or as the timed CPU baseline.  The product (gat_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_mbb.so"

P = C.c_void_p
I32, I64, U64 = C.c_int32, C.c_int64, C.c_uint64
F32, F64, INT = C.c_float, C.c_double, C.c_int

_FUNCS = {
    "orc_version": (INT, []),
    "orc_pinbox": (INT, [P, F32, F32]),
    "orc_box_intersect": (INT, [P, P]),
    "orc_morton": (INT, [INT, INT]),
    "orc_pair_digest": (U64, [I32, I32]),
    "orc_grid_new": (P, [P, F32, INT]),
    "orc_grid_free": (None, [P]),
    "orc_grid_add": (INT, [P, P, INT, P, P]),
    "orc_grid_info": (INT, [P, P]),
    "orc_grid_cell_of": (INT, [P, F32, F32]),
    "orc_grid_cell_box": (INT, [P, INT, P]),
    "orc_grid_store": (INT, [P, P, P, P, P]),
    "orc_grid_leaves": (I64, [P, P, P, I64]),
    "orc_grid_range": (I64, [P, P, INT, INT, P, I64, P, P, P, I64]),
    "orc_mgrid_new": (P, [P, F32, INT]),
    "orc_mgrid_free": (None, [P]),
    "orc_mgrid_add": (INT, [P, P, INT, P, P]),
    "orc_mgrid_bitmap_index": (INT, [INT, INT]),
    "orc_mgrid_info": (INT, [P, P]),
    "orc_mgrid_is_leaf": (INT, [P, INT, INT]),
    "orc_mgrid_node_box": (INT, [P, INT, INT, P]),
    "orc_mgrid_leaves": (I64, [P, P, P, I64]),
    "orc_mgrid_range": (I64, [P, P, INT, INT, P, I64, P, P, P, I64]),
    "orc_fsg_new": (P, [P, F32]),
    "orc_fsg_free": (None, [P]),
    "orc_fsg_add": (INT, [P, P, INT, P, P]),
    "orc_fsg_cell_of": (INT, [P, F32, F32]),
    "orc_fsg_range": (I64, [P, P, INT, INT, P, I64, P, P, P, I64]),
    "orc_stig_new": (P, []),
    "orc_stig_free": (None, [P]),
    "orc_stig_add": (INT, [P, P, INT, P, P, INT]),
    "orc_stig_range": (I64, [P, P, INT, INT, P, I64, P, P, P, I64]),
    "orc_brute_range": (I64, [P, INT, P, P, P, INT, INT, P, I64, P, P, P, I64]),
    "orc_edr": (INT, [P, INT, P, INT, F64]),
    "orc_edr_antidiag": (INT, [P, INT, P, INT, F64]),
    "orc_edr_textbook": (INT, [P, INT, P, INT, F64]),
    "orc_edr_batch": (INT, [P, INT, P, INT, P, F64, P, INT]),
    "orc_ftable": (INT, [INT, P]),
    "orc_topk": (INT, [P, P, INT, INT, P, P]),
    "orc_morton_x86": (INT, [INT, INT]),
    "orc_fv_neighbors": (INT, [INT, INT, P]),
    "orc_fd_value": (INT, [P, INT, P, INT, INT]),
    "orc_fv_new": (P, [P, F32, P, INT, P, P]),
    "orc_fv_free": (None, [P]),
    "orc_fv_fd": (INT, [P, P, INT, P]),
    "orc_fv_topk": (INT, [P, P, INT, P, INT, F64, P, P, P]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        l = C.CDLL(str(ORACLE_LIB))
        for name, (res, args) in _FUNCS.items():
            f = getattr(l, name)
            f.restype, f.argtypes = res, args
        _lib = l
    return _lib


def ref():
    """The reference's own compiled MBB.cpp, or None where it was not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        l = C.CDLL(str(REF_LIB))
        l.ref_pinbox.restype, l.ref_pinbox.argtypes = INT, [P, F32, F32]
        l.ref_box_intersect.restype, l.ref_box_intersect.argtypes = INT, [P, P]
        _ref = l
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def pinbox(box, x, y) -> int:
    return lib().orc_pinbox(_p(_f32(box)), x, y)


def box_intersect(a, b) -> int:
    return lib().orc_box_intersect(_p(_f32(a)), _p(_f32(b)))


def morton(x: int, y: int) -> int:
    return lib().orc_morton(x, y)


def pair_digest(tid: int, pid: int) -> int:
    return int(lib().orc_pair_digest(tid, pid))


class _RangeIndex:
    handle = None
    _free = None

    def __del__(self):
        if self.handle and self._free:
            getattr(lib(), self._free)(self.handle)
            self.handle = None

    def _range(self, fn, boxes, nthreads, mode, ntraj):
        b = _f32(boxes).reshape(-1, 4)
        nq = len(b)
        if mode == "bitmap":
            out = np.zeros((nq, ntraj + 1), np.uint8)
            fn(self.handle, _p(b), nq, nthreads, _p(out), ntraj + 1, None, None, None, 0)
            return out
        counts = np.zeros(nq, np.int64)
        if mode == "digest":
            dig = np.zeros(nq, np.uint64)
            fn(self.handle, _p(b), nq, nthreads, None, 0, _p(counts), _p(dig), None, 0)
            return counts, dig
        # pairs: first call sizes, second call fills
        total = fn(self.handle, _p(b), nq, nthreads, None, 0, _p(counts), None, None, 0)
        pairs = np.zeros((max(total, 1), 2), np.int32)
        got = fn(self.handle, _p(b), nq, nthreads, None, 0, _p(counts), None, _p(pairs), total)
        assert got == total, (got, total)
        cut = np.concatenate([[0], np.cumsum(counts)])
        return [pairs[cut[i]:cut[i + 1]] for i in range(nq)]


class Grid(_RangeIndex):
    """Restatement of Grid (GT-quadtree) — Grid.cpp:25-446."""
    _free = "orc_grid_free"

    def __init__(self, range_box, cell_size, max_points_in_node):
        self.handle = lib().orc_grid_new(_p(_f32(range_box)), cell_size, max_points_in_node)

    def add(self, db):
        lib().orc_grid_add(self.handle, _p(db.offsets), db.ntraj, _p(db.lon), _p(db.lat))
        self.ntraj = db.ntraj
        return self

    def info(self):
        out = np.zeros(6, np.int64)
        lib().orc_grid_info(self.handle, _p(out))
        return dict(zip(["axis", "cellnum", "total_level", "nodes", "leaves", "points"], out.tolist()))

    def cell_of(self, lon, lat):
        return lib().orc_grid_cell_of(self.handle, lon, lat)

    def cell_box(self, cell):
        out = np.zeros(4, np.float32)
        assert lib().orc_grid_cell_box(self.handle, cell, _p(out)) == 0
        return out

    def store(self):
        n = self.info()["points"]
        x, y = np.zeros(n, np.float32), np.zeros(n, np.float32)
        t, p = np.zeros(n, np.int32), np.zeros(n, np.int32)
        lib().orc_grid_store(self.handle, _p(x), _p(y), _p(t), _p(p))
        return x, y, t, p

    def leaves(self, box):
        n = lib().orc_grid_leaves(self.handle, _p(_f32(box)), None, 0)
        out = np.zeros((max(n, 1), 2), np.int64)
        lib().orc_grid_leaves(self.handle, _p(_f32(box)), _p(out), n)
        return out[:n]

    def range(self, boxes, mode="pairs", nthreads=1):
        return self._range(lib().orc_grid_range, boxes, nthreads, mode, self.ntraj)


class MortonGrid(_RangeIndex):
    """Restatement of MortonGrid (implicit GT-quadtree) — MortonGrid.cpp:8-393."""
    _free = "orc_mgrid_free"

    def __init__(self, range_box, cell_size, max_points_in_node):
        self.handle = lib().orc_mgrid_new(_p(_f32(range_box)), cell_size, max_points_in_node)

    @staticmethod
    def bitmap_index(level, nid):
        return lib().orc_mgrid_bitmap_index(level, nid)

    def add(self, db):
        lib().orc_mgrid_add(self.handle, _p(db.offsets), db.ntraj, _p(db.lon), _p(db.lat))
        self.ntraj = db.ntraj
        return self

    def info(self):
        out = np.zeros(5, np.int64)
        lib().orc_mgrid_info(self.handle, _p(out))
        return dict(zip(["axis", "cellnum", "total_level", "leaves", "points"], out.tolist()))

    def is_leaf(self, level, nid):
        return bool(lib().orc_mgrid_is_leaf(self.handle, level, nid))

    def node_box(self, level, nid):
        out = np.zeros(4, np.float32)
        lib().orc_mgrid_node_box(self.handle, level, nid, _p(out))
        return out

    def leaves(self, box):
        """(level, id, store start, points) of the leaves one rectangle
        selects, in traversal order."""
        n = lib().orc_mgrid_leaves(self.handle, _p(_f32(box)), None, 0)
        out = np.zeros((max(n, 1), 4), np.int64)
        lib().orc_mgrid_leaves(self.handle, _p(_f32(box)), _p(out), n)
        return out[:n]

    def range(self, boxes, mode="pairs", nthreads=1):
        return self._range(lib().orc_mgrid_range, boxes, nthreads, mode, self.ntraj)


class FSG(_RangeIndex):
    """Restatement of FSG — FSG.cpp:7-261."""
    _free = "orc_fsg_free"

    def __init__(self, range_box, cell_size):
        self.handle = lib().orc_fsg_new(_p(_f32(range_box)), cell_size)

    def add(self, db):
        lib().orc_fsg_add(self.handle, _p(db.offsets), db.ntraj, _p(db.lon), _p(db.lat))
        self.ntraj = db.ntraj
        return self

    def cell_of(self, lon, lat):
        return lib().orc_fsg_cell_of(self.handle, lon, lat)

    def range(self, boxes, mode="pairs", nthreads=1):
        return self._range(lib().orc_fsg_range, boxes, nthreads, mode, self.ntraj)


class STIG(_RangeIndex):
    """Corrected restatement of STIG — STIG.cpp:35-252."""
    _free = "orc_stig_free"

    def __init__(self):
        self.handle = lib().orc_stig_new()

    def add(self, db, block):
        lib().orc_stig_add(self.handle, _p(db.offsets), db.ntraj, _p(db.lon), _p(db.lat), block)
        self.ntraj = db.ntraj
        return self

    def range(self, boxes, mode="pairs", nthreads=1):
        return self._range(lib().orc_stig_range, boxes, nthreads, mode, self.ntraj)


def brute_range(db, boxes, mode="pairs", nthreads=1):
    b = _f32(boxes).reshape(-1, 4)
    nq = len(b)
    fn = lib().orc_brute_range
    args = (_p(db.offsets), db.ntraj, _p(db.lon), _p(db.lat), _p(b), nq, nthreads)
    if mode == "bitmap":
        out = np.zeros((nq, db.ntraj + 1), np.uint8)
        fn(*args, _p(out), db.ntraj + 1, None, None, None, 0)
        return out
    counts = np.zeros(nq, np.int64)
    if mode == "digest":
        dig = np.zeros(nq, np.uint64)
        fn(*args, None, 0, _p(counts), _p(dig), None, 0)
        return counts, dig
    total = fn(*args, None, 0, _p(counts), None, None, 0)
    pairs = np.zeros((max(total, 1), 2), np.int32)
    fn(*args, None, 0, _p(counts), None, _p(pairs), total)
    cut = np.concatenate([[0], np.cumsum(counts)])
    return [pairs[cut[i]:cut[i + 1]] for i in range(nq)]


def _xy(a):
    return _f32(a).reshape(-1, 2)


def edr(q, c, eps) -> int:
    """Grid::SimilarityExecuter for one pair (Grid.cpp:1222-1318); q, c: (n, 2) float32."""
    q, c = _xy(q), _xy(c)
    return lib().orc_edr(_p(q), len(q), _p(c), len(c), eps)


def edr_antidiag(q, c, eps) -> int:
    q, c = _xy(q), _xy(c)
    return lib().orc_edr_antidiag(_p(q), len(q), _p(c), len(c), eps)


def edr_textbook(q, c, eps) -> int:
    q, c = _xy(q), _xy(c)
    return lib().orc_edr_textbook(_p(q), len(q), _p(c), len(c), eps)


def edr_batch(q, db, eps, nthreads=1) -> np.ndarray:
    """Distances from query q to trajectories 1..db.ntraj."""
    q = _xy(q)
    cxy = db.interleaved()
    out = np.zeros(db.ntraj, np.int32)
    lib().orc_edr_batch(_p(q), len(q), _p(db.offsets), db.ntraj, _p(cxy), eps, _p(out), nthreads)
    return out


def ftable(maxlen) -> np.ndarray:
    out = np.zeros((maxlen + 1, maxlen + 1), np.int32)
    lib().orc_ftable(maxlen, _p(out))
    return out


def topk(dist, ids, k):
    dist = np.ascontiguousarray(dist, np.int32)
    ids = np.ascontiguousarray(ids, np.int32)
    oi, od = np.zeros(k, np.int32), np.zeros(k, np.int32)
    n = lib().orc_topk(_p(dist), _p(ids), len(dist), k, _p(oi), _p(od))
    return oi[:n], od[:n]


# ---- the reference's own CPU top-k: frequency-distance queue + 80-candidate
# rounds (FVTable.cpp:18-207, Grid.cpp:1021-1220).  Not exact; a timed baseline.

def morton_x86(x: int, y: int) -> int:
    """getIdxFromXY as the reference's x86-64 build evaluates it, negative
    coordinates included (FVTable.cpp:67-73 passes them at the grid's edge)."""
    return lib().orc_morton_x86(x, y)


def fv_neighbors(cell: int, cell_num: int) -> np.ndarray:
    out = np.zeros(8, np.int32)
    lib().orc_fv_neighbors(cell, cell_num, _p(out))
    return out


def fd_value(t: dict, q: dict, cell_num: int) -> int:
    """Frequency distance of two {virtual cell: points} vectors (FVTable.cpp:86-198)."""
    a = np.array(sorted(t.items()), np.int32).reshape(-1, 2)
    b = np.array(sorted(q.items()), np.int32).reshape(-1, 2)
    return lib().orc_fd_value(_p(a), len(a), _p(b), len(b), cell_num)


class FdTopk:
    """The reference's SimilarityQueryBatchCPUParallel (Grid.cpp:1133-1220) over
    a database, with the Grid geometry of (range_box, cell_size)."""

    def __init__(self, range_box, cell_size, db):
        self._db = db
        self.handle = lib().orc_fv_new(_p(_f32(range_box)), cell_size, _p(db.offsets), db.ntraj, _p(db.lon),
                                       _p(db.lat))

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_fv_free(self.handle)
            self.handle = None

    def fd(self, q) -> np.ndarray:
        q = _xy(q)
        out = np.zeros(self._db.ntraj, np.int32)
        lib().orc_fv_fd(self.handle, _p(q), len(q), _p(out))
        return out

    def topk(self, queries, k: int, eps: float):
        """(ids (nq, k) worst first as the reference pops them, distances,
        stats (nq, 2): rounds, candidates evaluated)."""
        xy = queries.interleaved()
        nq = queries.ntraj
        ids = np.zeros((max(nq, 1), k), np.int32)
        dist = np.zeros((max(nq, 1), k), np.int32)
        st = np.zeros((max(nq, 1), 2), np.int64)
        lib().orc_fv_topk(self.handle, _p(queries.offsets), nq, _p(xy), k, eps, _p(ids), _p(dist), _p(st))
        return ids[:nq], dist[:nq], st[:nq]
