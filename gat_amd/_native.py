"""ctypes bindings of the two native libraries built in gat_amd/lib.

libgat_hip.so carries every kernel and the C ABI of include/gat_hip.h;
libgat_host.so carries the C++ index/query classes and include/gat_host.h.
There is no pure-Python or CPU fallback: if a library is missing, loading
raises NativeLibraryMissing and nothing runs.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
HIP_LIB = LIB_DIR / "libgat_hip.so"
HOST_LIB = LIB_DIR / "libgat_host.so"


class NativeLibraryMissing(RuntimeError):
    pass


class GatError(RuntimeError):
    def __init__(self, rc: int, message: str):
        super().__init__(f"[status {rc}] {message}")
        self.rc = rc


P = C.c_void_p
I32, I64, U32, U64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
F32, F64 = C.c_float, C.c_double
INT = C.c_int
SZ = C.c_size_t
PP = C.POINTER(C.c_void_p)



class PointStore(C.Structure):
    """gat_point_store of include/gat_hip.h."""
    _fields_ = [("xy_dev", C.c_void_p), ("id_dev", C.c_void_p), ("pid_dev", C.c_void_p), ("npoints", C.c_int64),
                ("pid_bits", C.c_int32)]


# name: (restype, [argtypes])
HIP_FUNCS = {
    "gat_last_error": (C.c_char_p, []),
    "gat_init": (INT, [INT]),
    "gat_device_name": (INT, [C.c_char_p, INT]),
    "gat_device_props": (INT, [P]),
    "gat_mem_info": (INT, [C.POINTER(SZ), C.POINTER(SZ)]),
    "gat_malloc": (INT, [PP, SZ]),
    "gat_free": (INT, [P]),
    "gat_host_alloc": (INT, [PP, SZ]),
    "gat_host_free": (INT, [P]),
    "gat_memcpy_h2d": (INT, [P, P, SZ, P]),
    "gat_memcpy_d2h": (INT, [P, P, SZ, P]),
    "gat_memcpy_d2d": (INT, [P, P, SZ, P]),
    "gat_memset": (INT, [P, INT, SZ, P]),
    "gat_stream_create": (INT, [PP]),
    "gat_stream_destroy": (INT, [P]),
    "gat_stream_sync": (INT, [P]),
    "gat_device_sync": (INT, []),
    "gat_event_create": (INT, [PP]),
    "gat_event_destroy": (INT, [P]),
    "gat_event_record": (INT, [P, P]),
    "gat_event_sync": (INT, [P]),
    "gat_stream_wait_event": (INT, [P, P]),
    "gat_event_elapsed_ms": (INT, [C.POINTER(F32), P, P]),
    "gat_eps_threshold": (F32, [F64]),
    "gat_stream_copy": (INT, [P, P, SZ, P]),
    "gat_hbm_pattern": (INT, [INT, P, SZ, U32, P]),
    "gat_cell_store_build": (INT, [P, P, P, P, I32, I64, I32, P, P, P, P, P, P]),
    "gatRangeQueryTestHandler": (INT, [P, P, INT, P, P, INT, INT, P]),
    "gatRangeQuerySTIGHandler": (INT, [P, P, INT, P, P, INT, INT, P]),
    "gatRangeQueryTestHandlerNoAlloc": (INT, [P, P, INT, P, P, INT, INT, P]),
    "gat_range_refine_bitmap": (INT, [P, P, I64, P, I64, P, I64, P]),
    "gat_range_refine_pairs": (INT, [P, P, I64, P, I64, P, P, P, I64, P]),
    "gat_pairs_digest": (INT, [P, P, P, I64, I64, P, P]),
    "gatEDRDistanceBatchHandler": (INT, [INT, P, P, P, P, P, P, P, F64, P, P]),
    "gat_edr_batched": (INT, [P, I64, P, I64, P, P, P, P, P, I32, P, P, P, I32, I32, F64, I32, I32, P, P]),
    "gat_edr_jobs": (INT, [P, P, I64, P, I32, P, I32, I32, I32, I32, P, P, P, P, P, P, P, P]),
    "gat_edr_all_pairs": (INT, [I32, I32, P, P, P]),
    "gat_edr_unsort": (INT, [P, P, P, I64, P, P]),
    "gat_edr_survivor_lists": (INT, [P, P, P, I64, I32, P, P, P, P, P]),
    "gat_edr_filter": (INT, [P, I32, P, I32, F64, P, I64, P, P]),
    "gat_edr_filter_grid": (INT, [P, I32, P, I32, P, F64, P, I64, P, P]),
    "gat_edr_mismatch_table": (INT, [I32, P, P]),
    "gat_edr_pack_trajectories": (INT, [P, P, P, P, I32, I64, P, P, P]),
    "gat_edr_topk": (INT, [P, I32, P, I32, P, I32, P, P, P, P, P, I32, I32, P, P, P, P]),
    "gat_fill_u32": (INT, [P, U32, I64, P]),
    "gat_fill_u32_spans": (INT, [I32, P, P, U32, P]),
    "gat_scan_tiles": (I64, [I64]),
    "gat_scan_u64": (INT, [P, I64, P, P, P, P]),
    "gat_scan_leaf_lists": (INT, [P, I64, I32, P, P, P, P]),
    "gat_tree_select": (INT, [P, I64, I32, P, P, P, I64, I64, P, P, P, P, I64, P, P, P]),
    "gat_leaf_tasks": (INT, [P, P, I64, P, I64, P, P]),
    "gat_grid_select": (INT, [P, I32, I32, F32, F32, F32, P, P, I64, I64, P, P, P, P, I64, P, P]),
    "gat_morton_select": (INT, [P, I32, I64, F32, F32, F32, P, P, I64, I64, P, P, P, P, I64, P, P, P]),
    "gat_kd_store_build": (INT, [P, P, P, I32, I64, I32, P, P, P, I32, P, P, I64, P, P, P, P, P, P]),
    "gat_leaf_scatter": (INT, [P, P, I64, P, I64, P, P, P, I32, I64, P, I64, P, I64, P]),
    "gat_leaf_refine": (INT, [P, P, P, I64, P, I64, P, P, I64, P, P, I64, P, P]),
    "gat_leaf_item_regions": (INT, [P, P, I64, P, P, P, P]),
    "gat_leaf_refine_runs": (INT, [P, P, P, I64, P, I64, P, I64, P, I64, P, I64, P, P]),
    "gat_run_counts": (INT, [P, P, P, I64, I64, P, I64, I64, P, P, P]),
    "gat_runs_digest": (INT, [P, I64, P, I64, I64, P, P]),
    "gat_graph_begin": (INT, [P]),
    "gat_graph_end": (INT, [P, P]),
    "gat_graph_launch": (INT, [P, P]),
    "gat_graph_destroy": (INT, [P]),
}

HOST_FUNCS = {
    "gath_last_error": (C.c_char_p, []),
    "gath_init": (INT, [INT]),
    "gath_box_intersect": (INT, [P, P]),
    "gath_pinbox": (INT, [P, F32, F32]),
    "gath_morton": (U32, [U32, U32]),
    "gath_synth_lengths_uniform": (None, [U64, I64, INT, INT, P]),
    "gath_synth_lengths_lognormal": (None, [U64, I64, F64, F64, INT, INT, P]),
    "gath_synth_walks": (None, [U64, I64, P, F64, P, P, INT]),
    "gath_synth_rects": (None, [U64, I64, F64, F64, INT, F64, INT, F64, P]),
    "gath_synth_queries": (None, [U64, INT, P, P, P, INT, P, P, INT, F64, F64, P, P]),
    "gath_grid_new": (P, [P, F32, INT]),
    "gath_morton_grid_new": (P, [P, F32, INT]),
    "gath_grid_set_edr_waves": (INT, [P, INT]),
    "gath_grid_set_topk_device_chain": (INT, [P, INT]),
    "gath_system_range_test": (INT, [P, INT, P, INT, P]),
    "gath_system_similarity_test": (INT, [P, P, P, INT, INT, INT, F64, P, P, P, P, P]),
    "gath_fsg_new": (P, [P, F32]),
    "gath_stig_new": (P, [INT]),
    "gath_index_free": (None, [P]),
    "gath_index_build": (INT, [P, P, INT, P, P, INT]),
    "gath_traj_read": (P, [C.c_char_p]),
    "gath_traj_ingest_csv": (P, [C.c_char_p, C.c_char_p]),
    "gath_traj_free": (None, [P]),
    "gath_traj_info": (INT, [P, P, P]),
    "gath_traj_copy": (INT, [P, P, P, P, P, P]),
    "gath_traj_write": (INT, [C.c_char_p, P, INT, P, P, P, P, P]),
    "gath_traj_days_between": (INT, [C.c_char_p, C.c_char_p]),
    "gath_traj_vehicle_distance": (F32, [F32, F32, F32, F32]),
    "gath_index_info": (INT, [P, P]),
    "gath_index_set_threads": (INT, [P, INT]),
    "gath_index_select": (I64, [P, P, INT, P, I64, P]),
    "gath_index_cell_of": (I64, [P, F32, F32]),
    "gath_index_range_bitmap": (INT, [P, P, INT, P]),
    "gath_index_range_pairs": (INT, [P, P, INT, P, P]),
    "gath_index_range_batched": (INT, [P, P, INT, INT, P, P]),
    "gath_index_tree_nodes": (I64, [P, P, I64, P]),
    "gath_index_start_table": (I64, [P, P, I64, P, P]),
    "gath_index_fetch_store": (INT, [P, P, P, P, P]),
    "gath_index_cell_starts": (I64, [P, P, I64]),
    "gath_index_set_batch_params": (INT, [P, INT, INT]),
    "gath_index_last_unique_candidates": (INT, [P, P]),
    "gath_index_last_counts": (INT, [P, P]),
    "gath_serve": (INT, [P, P, INT, P, P, I64, P, P, P, P, P]),
    "gath_mixed_new": (P, [P, P, P]),
    "gath_mixed_free": (None, [P]),
    "gath_mixed_run": (INT, [P, P, INT, P, P, P, P, P]),
    "gath_grid_set_edr_params": (INT, [P, INT, INT, INT]),
    "gath_tune_range": (INT, [P, P, INT, INT, P]),
    "gath_tune_edr": (INT, [P, P, INT, P, P, F64, INT, P]),
    "gath_index_last_fetch": (INT, [P, P, I64]),
    "gath_index_last_runs": (I64, [P, P, I64]),
    "gath_index_set_result_layout": (INT, [P, INT]),
    "gath_index_last_digests": (INT, [P, P]),
    "gath_grid_edr_all": (INT, [P, P, P, INT, F64, P]),
    "gath_grid_similarity": (INT, [P, P, INT, P, P, INT, F64, P, P]),
    "gath_grid_edr_batched_all": (INT, [P, P, INT, P, P, F64, P, P, P]),
    "gath_grid_edr_batched_pairs": (INT, [P, P, INT, P, P, P, I64, F64, P, P, P]),
    "gath_grid_similarity_batched": (INT, [P, P, INT, P, P, INT, F64, P, P, P, P]),
    "gath_grid_mismatch_table": (INT, [P, P]),
}

_hip = None
_host = None


def _bind(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip():
    """The device library (loaded once, RTLD_GLOBAL so the host library binds to it)."""
    global _hip
    if _hip is None:
        if not HIP_LIB.exists():
            raise NativeLibraryMissing(f"{HIP_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(HIP_LIB), mode=C.RTLD_GLOBAL)
        _bind(lib, HIP_FUNCS)
        _hip = lib
    return _hip


def host():
    global _host
    if _host is None:
        hip()
        if not HOST_LIB.exists():
            raise NativeLibraryMissing(f"{HOST_LIB} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(HOST_LIB), mode=C.RTLD_GLOBAL)
        _bind(lib, HOST_FUNCS)
        _host = lib
    return _host


def check_hip(rc: int) -> None:
    if rc != 0:
        raise GatError(rc, hip().gat_last_error().decode(errors="replace"))


def check_host(rc: int) -> None:
    if rc != 0:
        raise GatError(rc, host().gath_last_error().decode(errors="replace"))


def exported_symbols(path: Path) -> set[str]:
    """Dynamic symbols a library defines (for the ABI-completeness test)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}
