"""gat_amd — an MI355X-native implementation of azh18/GAT's trajectory-query
hot path: range queries over the GT-quadtree, FSG and STIG indexes and the EDR
dynamic program behind top-k similarity, as hand-written gfx950 HIP kernels
behind a C ABI (include/gat_hip.h), driven by a C++ mirror of the reference's
host classes (include/gat_host.h).  This Python package only binds the two
native libraries for tests and benchmarks.
"""
from .api import (FsgIndex, GridIndex, MixedRange, MortonGridIndex, RangeResult, StigIndex, TrajSet, default_threads, device_name,
                  eps_threshold, hbm_pattern_bandwidth, ingest_vehicle_csv, init, read_trajectory_file, serve, stream_copy_bandwidth, synth_lengths, synth_queries, synth_rects,
                  synth_walks, write_trajectory_file)
from ._native import GatError, NativeLibraryMissing

__all__ = [
    "FsgIndex", "GridIndex", "MixedRange", "MortonGridIndex", "RangeResult", "StigIndex", "TrajSet", "default_threads", "device_name",
    "eps_threshold", "hbm_pattern_bandwidth", "ingest_vehicle_csv", "init", "read_trajectory_file", "serve", "stream_copy_bandwidth", "synth_lengths", "synth_queries", "synth_rects", "synth_walks", "write_trajectory_file", "GatError",
    "NativeLibraryMissing",
]
