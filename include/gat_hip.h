/*
 * gat_hip.h — the C ABI of libgat_hip.so, the MI355X (gfx950) device library
 * behind the trajectory-query path of azh18/GAT.
 *
 * This is the drop-in boundary: the reference's host classes call the CUDA
 * handlers declared in cudaKernel.h (cudaKernel.h:14-31) and the CUDA runtime
 * directly; the host classes of this project call only the functions below.
 * Signatures carry plain pointers, sizes and an opaque stream handle — no HIP
 * or torch types.  Every function returns GAT_OK (0) or a non-zero status and
 * never aborts or throws; gat_last_error() returns the calling thread's last
 * error message.  Device pointers are marked `_dev`; host pointers `_host`.
 *
 * Point stores are structure-of-arrays in HBM (gat_point_store): (lon, lat)
 * float pairs and compact (trajectory id, point id) arrays, in the index's own
 * point order.  Trajectory ids are 1-based (Grid.cpp:215-221); a point id is
 * the point's 0-based position within its trajectory.
 */
#ifndef GAT_HIP_H
#define GAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

enum {
  GAT_OK = 0,
  GAT_ERR_HIP = 1,      /* a HIP runtime call failed */
  GAT_ERR_ARG = 2,      /* invalid argument (null pointer, bad size or length) */
  GAT_ERR_NOMEM = 3,    /* device or pinned allocation failed */
  GAT_ERR_CAPACITY = 4, /* an output or scratch buffer is too small */
  GAT_ERR_STATE = 5     /* library not initialised or object misused */
};

/* Longest trajectory the EDR kernels accept (the reference caps at 1024,
 * ConstDefine.h:24; the workloads here reach 2048). */
#define GAT_MAX_TRAJ_LEN 2048
/* NaN points a padded EDR store holds before every trajectory, and after its
 * last one (GAT_EDR_TAIL_PAD). */
#define GAT_EDR_PAD 16
#define GAT_EDR_TAIL_PAD (GAT_MAX_TRAJ_LEN + 64)

/* ------------------------------------------------------------------ runtime */
/* Replaces the reference's direct cudaSetDevice/cudaStreamCreate/cudaMalloc/
 * cudaMemcpyAsync/cudaFree calls (Grid.cpp:470-476, 497, 717; STIG.cpp:271-293;
 * FSG.cpp:165-170, 269, 339; SystemTest.cpp:59-78) and the CUDA_CALL abort
 * macro (cudaKernel.h:10). */
const char* gat_last_error(void);
int gat_init(int device);
int gat_device_name(char* buf, int len);
int gat_device_props(int64_t* out); /* [cu_count, lds_per_block, total_mem, clock_khz, warp_size] */
int gat_mem_info(size_t* free_bytes, size_t* total_bytes);
int gat_malloc(void** ptr_dev, size_t bytes);
int gat_free(void* ptr_dev);
int gat_host_alloc(void** ptr_host, size_t bytes); /* pinned host memory */
int gat_host_free(void* ptr_host);
int gat_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes, void* stream);
int gat_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes, void* stream);
int gat_memcpy_d2d(void* dst_dev, const void* src_dev, size_t bytes, void* stream);
int gat_memset(void* dst_dev, int value, size_t bytes, void* stream);
int gat_stream_create(void** stream);
int gat_stream_destroy(void* stream);
int gat_stream_sync(void* stream);
int gat_device_sync(void);
int gat_event_create(void** event);
int gat_event_destroy(void* event);
int gat_event_record(void* event, void* stream);
int gat_event_sync(void* event);
/* Later work on `stream` waits for `event` (also inside a graph capture: the
 * way a captured pipeline forks onto other streams and joins back). */
int gat_stream_wait_event(void* stream, void* event);
int gat_event_elapsed_ms(float* ms, void* start, void* stop);

/* Streaming copy of `bytes` (a multiple of 16) with a plain grid-stride float4
 * kernel: the stream-copy HBM bandwidth reference, and the known-size kernel
 * whose rocprofv3 FETCH_SIZE / WRITE_SIZE readings calibrate those of the other
 * kernels.  No reference counterpart.  Asynchronous on `stream`. */
int gat_stream_copy(void* dst_dev, const void* src_dev, size_t bytes, void* stream);

/* ---------------------------------------------------------- index build */

/* Geometry of a cell grid: cells_x columns and cells_y rows of cell_size,
 * columns counted right from xmin and rows down from ymax.  morton = 1: the
 * GT-quadtree's square Morton-numbered grid (Grid.cpp:97-124; cells_x ==
 * cells_y, a power of two); 0: FSG's row-major grid (FSG.cpp:7-28). */
typedef struct gat_cell_grid {
  int32_t morton;
  int32_t cells_x, cells_y;
  float xmin, ymax, cell_size;
} gat_cell_grid;

/* Builds the cell-ordered point store on the device (replaces the host
 * bucketing of Grid::addDatasetToGrid, Grid.cpp:202-261, and
 * FSG::addDatasetToGrid, FSG.cpp:95-143): trajectories 1..ntraj, trajectory t
 * at points [offsets[t-1], offsets[t]) of lon / lat (trajectory order;
 * offsets_dev nondecreasing from 0 to npoints).  Every point goes to the cell
 * of the reference's float formula (Grid.cpp:202-208, FSG.cpp:88-93); points
 * are ordered by cell and, within a cell, keep trajectory and point order.
 * Outputs, in that order: xy_out (npoints lon, lat pairs), id_out (trajectory
 * id << pid_bits | point id when pid_bits > 0, else the trajectory id with the
 * point id in pid_out) and cell_start_out (cells_x * cells_y + 1 store
 * offsets, the last = npoints).  *bad_point receives the first point that lies
 * off the grid (nothing else is written then), or -1.  Allocates and frees its
 * own scratch (about 20 bytes per point); returns after `stream` has drained.
 * No reference counterpart (the reference builds on the host). */
int gat_cell_store_build(const gat_cell_grid* grid, const float* lon_dev, const float* lat_dev,
                         const int64_t* offsets_dev, int32_t ntraj, int64_t npoints, int32_t pid_bits,
                         float* xy_out_dev, uint32_t* id_out_dev, uint16_t* pid_out_dev, int64_t* cell_start_out_dev,
                         int64_t* bad_point, void* stream);

/* Builds STIG's kd-tree store on the device (STIG::initial / createIndex,
 * STIG.cpp:35-168, with the corrected split: a node [lo, hi] splits at
 * mid = (lo + hi) / 2 into [lo, mid - 1] and [mid, hi], the points of the left
 * part <= the point at mid <= those of the right part on the node's axis, x at
 * even depths, y at odd).  The host passes the tree's shape (it depends only
 * on node sizes): the split nodes of depth d are
 * (split_lo[k], split_hi[k]) for k in [depth_first[d], depth_first[d+1]),
 * ascending and disjoint, d < depths; and the leaves (leaf_lo, leaf_hi)
 * (inclusive ranges).  Skeleton arrays, median_out (one float per split node:
 * the axis coordinate of the point at mid) and leaf_box_out (xmin, ymin, xmax,
 * ymax per leaf, from (999, 999, -999, -999) as STIG.cpp:77-90) are host
 * memory; the trajectories and the store (as gat_cell_store_build's) are
 * device memory.  Each depth is one stable radix sort, so points with equal
 * coordinates keep their previous order and the build is deterministic.
 * Allocates its own scratch (about 32 bytes per point); returns after
 * `stream` has drained.  No reference counterpart. */
int gat_kd_store_build(const float* lon_dev, const float* lat_dev, const int64_t* offsets_dev, int32_t ntraj,
                       int64_t npoints, int32_t pid_bits, const int64_t* split_lo, const int64_t* split_hi,
                       const int64_t* depth_first, int32_t depths, const int64_t* leaf_lo, const int64_t* leaf_hi,
                       int64_t nleaves, float* xy_out_dev, uint32_t* id_out_dev, uint16_t* pid_out_dev,
                       float* median_out, float* leaf_box_out, void* stream);

/* HBM access-pattern ceilings over a buffer of `bytes` (a positive multiple of
 * 16): kind 0 writes the buffer with 16-byte streaming stores, 1 reads it; 2
 * and 3 write runs of `run` (1..4096) 8-byte pairs packed back to back, one
 * run per wave step in scattered order, as result runs from many waves land
 * in per-query regions: 2 with 64-lane stores from each run's first pair, 3
 * through 64-lane windows aligned to 128-byte lines (streaming stores); 4 and
 * 5 the same with cached stores.  No reference counterpart.  Asynchronous on
 * `stream`. */
int gat_hbm_pattern(int kind, void* buf_dev, size_t bytes, uint32_t run, void* stream);

/* ------------------------------------------------------------------ range */
/* A resident point store: the index's points in its own order.  Ids are
 * compact: with pid_bits > 0, id_dev[i] = (tid << pid_bits) | pid and pid_dev
 * is unused; with pid_bits == 0, id_dev[i] = tid and pid_dev[i] = pid.
 * Replaces the reference's AoS SPoint runs (ConstDefine.h:71-75), which carry
 * x, y and tid for every point and no point id. */
typedef struct gat_point_store {
  const float* xy_dev;     /* npoints (lon, lat) float pairs */
  const uint32_t* id_dev;  /* npoints ids, see above */
  const uint16_t* pid_dev; /* npoints point ids when pid_bits == 0 */
  int64_t npoints;
  int32_t pid_bits;
} gat_point_store;

/* One unit of refine work: `count` consecutive store points starting at
 * `start`, tested against rectangle `qid`.  Replaces RangeQueryStateTable
 * (ConstDefine.h:92-101): the rectangle lives once per query in `boxes_dev`
 * (xmin, ymin, xmax, ymax floats, the MBB member order of MBB.h:5) instead of
 * being copied into every entry, and the point run is an offset into the
 * resident store instead of a raw device pointer. */
typedef struct gat_range_task {
  uint32_t start;
  uint32_t count;
  uint32_t qid;
} gat_range_task;

/* Drop-in for cudaRangeQueryTestHandler (cudaKernel.h:21-22, kernel.cu:1282-1312)
 * and cudaRangeQuerySTIGHandler (cudaKernel.h:28-29, kernel.cu:1211-1242): refine
 * every task and set result_host[qid * maxTrajNum + tid] = 1 for each point
 * inside its rectangle (inclusive containment, MBB.cpp:21-26; the reference's GPU
 * kernels test strictly, kernel.cu:1055).  result_host (maxQueryNum * maxTrajNum
 * bytes) is fully overwritten with 0/1.  Synchronous on `stream`, as the
 * reference (which synchronises the device, kernel.cu:1300-1306). */
int gatRangeQueryTestHandler(const gat_point_store* store, const gat_range_task* tasks_dev, int stateTableLength,
                             const float* boxes_dev, uint8_t* result_host, int maxTrajNum, int maxQueryNum,
                             void* stream);
int gatRangeQuerySTIGHandler(const gat_point_store* store, const gat_range_task* tasks_dev, int stateTableLength,
                             const float* boxes_dev, uint8_t* result_host, int maxTrajNum, int maxQueryNum,
                             void* stream);

/* Largest bitmap (maxQueryNum * maxTrajNum bytes) gatRangeQueryTestHandlerNoAlloc
 * takes: 64 MiB, e.g. 100 rectangles x 671,088 trajectory slots. */
#define GAT_NOALLOC_BITMAP_MAX (64u << 20)

/* No counterpart in the reference: gatRangeQueryTestHandler for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * result as gatRangeQueryTestHandler, but the device bitmap is a fixed
 * GAT_NOALLOC_BITMAP_MAX-byte array the library holds on every GPU it runs on
 * until the process ends, so the call makes no hipMalloc, hipFree or other
 * allocating call.  A bitmap larger than that returns GAT_ERR_CAPACITY before
 * any device work.  Every call on one GPU shares its array (each call clears the
 * part it uses first): calls that overlap on one GPU, from several threads or on
 * several streams, race and must be serialised by the caller.  Synchronous on
 * `stream`. */
int gatRangeQueryTestHandlerNoAlloc(const gat_point_store* store, const gat_range_task* tasks_dev,
                                    int stateTableLength, const float* boxes_dev, uint8_t* result_host,
                                    int maxTrajNum, int maxQueryNum, void* stream);

/* The same refine, with the trajectory bitmap left in device memory
 * (bitmap_dev: rows of `traj_stride` bytes, one row per query, zeroed by the
 * caller).  Asynchronous on `stream`.
 * In every range entry point, tasks whose qid >= nqueries (maxQueryNum) or whose
 * run leaves the store's `npoints` points are skipped, and trajectory ids at or
 * beyond the bitmap stride are ignored: malformed input cannot make the kernel
 * read or write out of bounds. */
int gat_range_refine_bitmap(const gat_point_store* store, const gat_range_task* tasks_dev, int64_t ntasks,
                            const float* boxes_dev, int64_t nqueries, uint8_t* bitmap_dev, int64_t traj_stride,
                            void* stream);

/* The same refine, emitting compacted (tid, pid) int32 pairs.  Query q's pairs
 * are written contiguously from pairs_dev[2 * qbase_dev[q]]; qcount_dev[q] (zeroed
 * by the caller) ends as their number.  The caller sizes each query's region by
 * its candidate count (an exact upper bound).  The order of pairs within a query
 * is unspecified.  Pairs that would land at or beyond `pairs_capacity` (in
 * pairs) are dropped while qcount_dev still counts them, so an undersized
 * region shows as qcount > region size.  Asynchronous on `stream`. */
int gat_range_refine_pairs(const gat_point_store* store, const gat_range_task* tasks_dev, int64_t ntasks,
                           const float* boxes_dev, int64_t nqueries, const uint64_t* qbase_dev,
                           uint32_t* qcount_dev, int32_t* pairs_dev, int64_t pairs_capacity, void* stream);

/* Per-query order-independent digest of a compacted-pair result: digest_dev[q]
 * = wrapping sum over query q's pairs of a 64-bit mix (the splitmix64
 * finaliser) of (uint64(tid) << 32 | pid).  Lets two results be compared as
 * sets without copying them.  Reads stop at `pairs_capacity` pairs.
 * Asynchronous on `stream`.  (No reference
 * counterpart: the reference checks results by dumping them to text files,
 * SystemTest.cpp:79-89.) */
int gat_pairs_digest(const int32_t* pairs_dev, const uint64_t* qbase_dev, const uint32_t* qcount_dev, int64_t nq,
                     int64_t pairs_capacity, uint64_t* digest_dev, void* stream);

/* ------------------------------------------------------------------ EDR */
/* One (query, candidate) pair.  Replaces TaskInfoTableForSimilarity
 * (ConstDefine.h:108-111). */
typedef struct gat_edr_task {
  int32_t qid;
  int32_t cand;
} gat_edr_task;

/* Drop-in for EDRDistance_Batch_Handler (cudaKernel.h:24, kernel.cu:503-508):
 * result_dev[t] = the reference's EDR (Grid.cpp:1222-1318) between query
 * tasks_dev[t].qid (query_xy_dev from 2*query_offset_dev[qid], query_length_dev[qid]
 * points) and its candidate (cand_xy_dev from 2*cand_offset_dev[t],
 * cand_length_dev[t] points).  The SPoint** candidate pointers of the reference
 * become point offsets into one candidate store.  Two points match when
 * |dx| < eps and |dy| < eps with the float differences compared as doubles
 * (Grid.cpp:1281).  Lengths must lie in [1, GAT_MAX_TRAJ_LEN].  Asynchronous on
 * `stream`, as the reference's handler. */
int gatEDRDistanceBatchHandler(int queryTaskNum, const gat_edr_task* tasks_dev, const float* query_xy_dev,
                               const int32_t* query_offset_dev, const int32_t* query_length_dev,
                               const float* cand_xy_dev, const int64_t* cand_offset_dev,
                               const int32_t* cand_length_dev, double eps, int32_t* result_dev, void* stream);

/* A batched-EDR work item: `count` (1..4) consecutive entries of the pair
 * list starting at `first`, all with the same strip count ceil(min length / 16). */
typedef struct gat_edr_job {
  int32_t first;
  int32_t count;
} gat_edr_job;

/* Batched EDR (no reference counterpart; the reference launches one block per
 * pair in rounds of 80 candidates per query, Grid.cpp:1457-1638): out_dev[i] =
 * EDR (Grid.cpp:1222-1318) of pair i = (query pairs_dev[2i], candidate
 * pairs_dev[2i+1]), 0-based indices into the query and candidate stores, for
 * every pair covered by jobs_dev.  One persistent launch hands jobs out
 * through counter_dev (one int32 of scratch, reset by the call); jobs should be
 * ordered longest first.  Both stores must be NaN-padded: at least 16 NaN
 * points before every trajectory and GAT_MAX_TRAJ_LEN + 64 after the last.
 * maxlen (the longest trajectory of any pair) selects the kernel's LDS class
 * (row buffers of 256, 512, 1024 or 2048 columns); a pair with an id outside
 * [0, nqueries) / [0, ncands) or a length outside 1..maxlen gets -1, and jobs
 * reaching past npairs are skipped.  per_cu is the number of single-wave
 * workgroups launched per compute unit (0: the class default).
 * lanes_per_pair is 16 (four pairs per wave, jobs of up to four pairs) or 64
 * (one pair per wave and a quarter of the sequential strips, for the longest
 * pairs; a job's entries past its first are not computed, so such jobs hold one
 * pair).  Callers split pairs into length buckets and launch each bucket with
 * its own maxlen, so short pairs are not run at the occupancy of the longest.
 * window_dev (may be null): two int32 on the device, read when the launch
 * runs, that select jobs [window[0], window[0] + window[1]) of the njobs the
 * array holds (gat_edr_jobs writes them).  Asynchronous on `stream`. */
int gat_edr_batched(const int32_t* pairs_dev, int64_t npairs, const gat_edr_job* jobs_dev, int64_t njobs,
                    const int32_t* window_dev, int32_t* counter_dev, const float* query_xy_dev, const int64_t* query_offset_dev,
                    const int32_t* query_length_dev, int32_t nqueries, const float* cand_xy_dev,
                    const int64_t* cand_offset_dev, const int32_t* cand_length_dev, int32_t ncands, int32_t maxlen,
                    double eps, int32_t per_cu, int32_t lanes_per_pair, int32_t* out_dev, void* stream);

/* The job list of a batch of survivor pairs built on the device (no reference
 * counterpart), so that filter, DP and merge run without a host round trip.
 * Reads the *count_dev pairs (query, candidate) of pairs_dev (the filter's
 * output; at most capacity); writes them to sorted_dev grouped by length
 * bucket (longer length up to `split`; up to max(split, 1024), or `split` in a
 * batch of fewer than 64 jobs per CU; longer; and, in such a batch with waves
 * > 1, the long pairs that would outlast the batch's longest several-wave
 * pair, or every long pair when every_long) and within a bucket by strip count
 * and longer length (in sixteen steps of the bucket's range), descending; jobs_dev (capacity + 4 entries) gets jobs of
 * up to four pairs of one strip count (one pair in the two long buckets) and
 * windows_dev[2b], [2b+1] the first job and job count of bucket b (four
 * buckets), for gat_edr_batched's windows; origin_dev (may be null; capacity
 * entries) gets each sorted pair's index in pairs_dev.  scratch_dev: capacity * 2
 * + GAT_EDR_JOBS_SCRATCH_EXTRA int32;
 * *cells_dev = the sum of the pairs' DP cells.  When *count_dev exceeds the
 * capacity nothing is sorted, every window is empty and status bit 1 is set
 * (*status_dev is cleared first).  Asynchronous on `stream`. */
#define GAT_EDR_JOBS_SCRATCH_EXTRA 32768
int gat_edr_jobs(const int32_t* pairs_dev, const uint64_t* count_dev, int64_t capacity, const int32_t* query_length_dev,
                 int32_t nqueries, const int32_t* cand_length_dev, int32_t ncands, int32_t split, int32_t waves,
                 int32_t every_long, int32_t* sorted_dev, int32_t* origin_dev, int32_t* scratch_dev,
                 gat_edr_job* jobs_dev, int32_t* windows_dev, int32_t* status_dev, uint64_t* cells_dev,
                 void* stream);

/* Every (query, candidate) pair of nq x nc: pairs_dev[2i], [2i+1] = (i / nc,
 * i % nc); *count_dev = nq * nc (at most INT32_MAX).  For gat_edr_jobs over a
 * whole query batch.  Asynchronous on `stream`. */
int gat_edr_all_pairs(int32_t nq, int32_t nc, int32_t* pairs_dev, uint64_t* count_dev, void* stream);

/* Distances back in the pairs' first order: out_dev[origin_dev[p]] =
 * sorted_dist_dev[p] for p below *count_dev (nothing when it exceeds the
 * capacity); origin_dev and the order are gat_edr_jobs'.  Asynchronous on
 * `stream`. */
int gat_edr_unsort(const int32_t* origin_dev, const int32_t* sorted_dist_dev, const uint64_t* count_dev,
                   int64_t capacity, int32_t* out_dev, void* stream);

/* Per-query survivor lists for gat_edr_topk from gat_edr_jobs' sorted pairs
 * and their distances (dist_dev in the same order): start_dev (nq + 1) = the
 * lists' offsets, ids_dev / dists_dev = each query's candidates ascending by
 * id with their distances.  nq <= 8192.  A query with more than 2048
 * survivors keeps its list unsorted and sets status bit 2.  Nothing is
 * listed when *count_dev exceeds the capacity.  Asynchronous on `stream`. */
int gat_edr_survivor_lists(const int32_t* sorted_dev, const int32_t* dist_dev, const uint64_t* count_dev,
                           int64_t capacity, int32_t nq, int32_t* start_dev, int32_t* ids_dev, int32_t* dists_dev,
                           int32_t* status_dev, void* stream);

/* Candidate filter for exact top-k pruning (no reference counterpart; the
 * reference orders candidates by an unproven frequency-distance bound,
 * FVTable.cpp:77-207): writes every (query, candidate) index pair whose
 * eps-dilated bounding boxes can hold a matching point pair to pairs_out_dev,
 * each once, in no defined order (up to `capacity` pairs; *count_dev, reset by
 * the call, ends as the full number).  Boxes are (xmin, ymin, xmax, ymax) floats per trajectory.  The test
 * is exactly conservative under float rounding, so every dropped pair has no
 * matching point pair and its distance is gat_edr_mismatch_table's value.
 * nq <= 65535.  Asynchronous on `stream`. */
int gat_edr_filter(const float* query_box_dev, int32_t nq, const float* cand_box_dev, int32_t nc, double eps,
                   int32_t* pairs_out_dev, int64_t capacity, uint64_t* count_dev, void* stream);

/* A grid over the candidates' boxes for gat_edr_filter_grid (no reference
 * counterpart; built by the host index, Grid::uploadTrajectories).  The cell
 * column of a coordinate v is floor((v - x0) * inv_x) clamped to
 * 0 .. cells_per_side - 1, in float (rows likewise with y0 and inv_y).
 * Cells are numbered row by row; entries cell_start_dev[cell] ..
 * cell_start_dev[cell + 1] - 1 list the candidates (entry_id_dev, with their
 * boxes in entry_box_dev) whose box is finite, not inverted, and touches the
 * cell (every cell from the cell of its min corner to the cell of its max
 * corner).  Every other candidate, and any the builder chose not to list, is
 * in always_dev.  The kernel trusts the layout: cell_start_dev must be
 * non-decreasing from 0, and entry_id_dev / entry_box_dev must hold
 * cell_start_dev[cells_per_side^2] entries (ids outside [0, nc) are skipped). */
typedef struct gat_edr_cand_grid {
  int32_t cells_per_side;
  float x0, y0, inv_x, inv_y;
  const uint32_t* cell_start_dev; /* cells_per_side^2 + 1 */
  const int32_t* entry_id_dev;
  const float* entry_box_dev; /* 4 floats per entry */
  const int32_t* always_dev;
  int32_t nalways;
} gat_edr_cand_grid;

/* gat_edr_filter's output through the candidate grid: the same set of pairs
 * (each once, in no defined order), found by testing only the candidates of
 * the cells each query's box (widened by eps and one cell) touches, plus the
 * always-tested ones; a query whose box is not finite or inverted tests every
 * candidate.  cand_box_dev holds every candidate's box as for gat_edr_filter.
 * *count_dev is reset by the call.  Asynchronous on `stream`. */
int gat_edr_filter_grid(const float* query_box_dev, int32_t nq, const float* cand_box_dev, int32_t nc,
                        const gat_edr_cand_grid* grid, double eps, int32_t* pairs_out_dev, int64_t capacity,
                        uint64_t* count_dev, void* stream);

/* Exact top-k merge on the device (no reference counterpart: the reference
 * keeps a k-entry priority queue per query on the host, Grid.cpp:1600-1729,
 * fed in its frequency-distance order).  For
 * query q (length qlen_dev[q]), every candidate c in 0..nc-1 with a length
 * clen_dev[c] in 1..GAT_MAX_TRAJ_LEN has distance surv_dist_dev[j] when c =
 * surv_id_dev[j] for some j in [surv_start_dev[q], surv_start_dev[q + 1])
 * (each query's survivors ascending by candidate, at most once), else
 * ftable_dev[min * w + max] of the two lengths (gat_edr_mismatch_table,
 * w = GAT_MAX_TRAJ_LEN + 1).  len_ids_dev holds the candidates grouped by
 * length, ascending inside each length, length L at [len_off_dev[L],
 * len_off_dev[L + 1]) for L in 0..GAT_MAX_TRAJ_LEN.  Writes the kk smallest
 * by (distance, candidate) to out_*_dev[q * k + i], i < kk, ids 1-based;
 * kk <= min(k, candidates), k <= 1024.  *status_dev (reset by the call) ends
 * non-zero when the inputs break these rules (a survivor distance outside
 * 0..2 * GAT_MAX_TRAJ_LEN, a survivor listed twice).  One workgroup per query;
 * asynchronous on `stream`. */
int gat_edr_topk(const int32_t* ftable_dev, int32_t w, const int32_t* qlen_dev, int32_t nq, const int32_t* clen_dev,
                 int32_t nc, const int32_t* len_ids_dev, const int32_t* len_off_dev, const int32_t* surv_start_dev,
                 const int32_t* surv_id_dev, const int32_t* surv_dist_dev, int32_t k, int32_t kk, int32_t* out_ids_dev,
                 int32_t* out_dist_dev, int32_t* status_dev, void* stream);

/* Builds the padded store the batched EDR kernels read (no reference
 * counterpart: the reference packs its queries into SPoint arrays on the host and
 * copies them one by one, Grid.cpp:1388-1412): fills
 * xy_dev (total_points (lon, lat) float pairs) with NaN, then writes trajectory
 * t's points lon_dev / lat_dev[src_offsets_dev[t] .. src_offsets_dev[t + 1]) as
 * pairs from point dst_offsets_dev[t], and its (xmin, ymin, xmax, ymax) to
 * box_dev[4t..4t+3].  Points that would land outside the store are dropped.
 * Asynchronous on `stream`. */
int gat_edr_pack_trajectories(const float* lon_dev, const float* lat_dev, const int64_t* src_offsets_dev,
                              const int64_t* dst_offsets_dev, int32_t ntraj, int64_t total_points, float* xy_dev,
                              float* box_dev, void* stream);

/* out_dev[r * (maxlen + 1) + c] = the recurrence of Grid.cpp:1264-1306 with
 * every substitution costing 1, for 0 <= r, c <= maxlen: the distance of a pair
 * of lengths r <= c with no matching point.  Asynchronous on `stream`. */
int gat_edr_mismatch_table(int32_t maxlen, int32_t* out_dev, void* stream);

/* The float threshold u with (double)|d| < eps  <=>  |d| < u for every float d:
 * the smallest float >= eps. */
float gat_eps_threshold(double eps);

/* ---------------------------------------------------------------------------
 * Batched range form: candidate selection on the device and a leaf-major join.
 *
 * No counterpart in the reference, which selects leaves on the host per query
 * (Grid::findMatchNodeInQuadTreeGPU, Grid.cpp:703-761; STIG::searchTree,
 * STIG.cpp:179-252) and refines one (query, leaf chunk) per block.  Here a batch
 * of rectangles is selected against a device copy of the tree, the selected
 * (leaf, query) pairs are grouped by leaf, and each leaf's points are read once
 * and tested against every rectangle that selected it.  Results are the same
 * sets as the plain path: the points of each selected leaf inside the
 * rectangle (MBB::pInBox, MBB.cpp:21-26).
 * ------------------------------------------------------------------------- */

/* A tree node in breadth-first order, so a node's children are contiguous at
 * first_child .. first_child + nchild - 1.  A node is visited when its parent
 * was and its test passes:
 *   flags & GAT_NODE_UNTESTED  -> always (the GT-quadtree root, Grid.cpp:428-446);
 *   flags & GAT_NODE_INCLUSIVE -> box overlaps the rectangle, edges included
 *                                 (STIG's intersectBlock and leaf test, STIG.cpp:340-347);
 *   otherwise                  -> BoxIntersect(rectangle, box) != 0 (MBB.cpp:28-73).
 * Leaves (first_child < 0) name the store run [start, start + count). */
typedef struct gat_tree_node {
  float box[4]; /* xmin, ymin, xmax, ymax */
  int32_t first_child;
  uint32_t start;
  uint32_t count;
  uint32_t flags;
} gat_tree_node;

#define GAT_NODE_INCLUSIVE 1u
#define GAT_NODE_UNTESTED 2u
/* Entries of the per-wave traversal stack; a tree of depth D with nchild
 * (2 or 4) children per inner node needs (D + 1) * nchild * 64 of them. */
#define GAT_TREE_STACK 4096
/* Bit set in *status by gat_tree_select when a traversal would exceed the stack. */
#define GAT_STATUS_STACK 1u

/* dst[0 .. n) = value, by a kernel (safe inside a captured graph). */
int gat_fill_u32(uint32_t* dst_dev, uint32_t value, int64_t n, void* stream);

/* The same for up to four spans in one launch: dst_dev[k][0 .. n[k]) = value
 * for k < nspans (dst_dev and n are host arrays; a span of length 0 may have
 * a null pointer). */
int gat_fill_u32_spans(int32_t nspans, uint32_t* const* dst_dev, const int64_t* n, uint32_t value, void* stream);

/* Number of tile partials the scans below need for n elements. */
int64_t gat_scan_tiles(int64_t n);

/* Exclusive prefix sum: out[i] = in[0] + ... + in[i-1]; *total = sum of all.
 * tiles_dev holds gat_scan_tiles(n) entries of scratch.  out may alias in.
 * With out_dev null only *total is written (tiles_dev unused). */
int gat_scan_u64(const uint64_t* in_dev, int64_t n, uint64_t* out_dev, uint64_t* total_dev, uint64_t* tiles_dev,
                 void* stream);

/* Per node, the offsets of its query list and of its work items, a work item
 * being up to qb queries of one leaf: node_off[2i] = sum over j < i of
 * count[j], node_off[2i+1] = sum over j < i of ceil(count[j] / qb); totals[0..1]
 * are the two sums.  tiles_dev holds 2 * gat_scan_tiles(n) uint32 entries. */
int gat_scan_leaf_lists(const uint32_t* count_dev, int64_t nnodes, int32_t qb, uint32_t* node_off_dev,
                        uint32_t* totals_dev, uint32_t* tiles_dev, void* stream);

/* Start table of a quadtree walk (no reference counterpart): owner_dev holds,
 * for each of the 4^level blocks of one level of a quadtree over a 2^L x 2^L
 * Morton grid (numbered in Morton order: column bits even, row bits odd, rows
 * counted down from ymax), the node that owns it as (node level << 27) |
 * node index: the block's own node at that level, or the leaf above it that
 * covers it.  A rectangle's walk then starts at the owners of the blocks
 * within one block of its extent (blocks of side `block` from (xmin, ymax)),
 * each owner taken once and tested, instead of at the root.  Only for trees
 * where every box holds its children's boxes and is wider than zero in x
 * (BoxIntersect is then monotone along every root path, so a node is reached
 * exactly when its own box intersects the rectangle); rectangles with NaN or
 * inverted extents start at the root.
 * Optional walk tasks for the first rectangles of order_dev (tasks_dev null:
 * none): ntasks_dev[0] tasks (at most tasks_capacity) stand for the first
 * ntasks_dev[1] entries of order_dev; task i = (query tasks_dev[2i], chunk
 * tasks_dev[2i+1] & 0x7fffffff, bit 31 set on the query's last task) walks
 * the query's candidate blocks [64 chunk, 64 chunk + 64) in row-major order,
 * its last task every block from 64 chunk on; every other entry of order_dev
 * is one task that walks all its blocks.  cand[q] is then accumulated over a
 * query's tasks (it must start at zero).  A NaN or inverted rectangle's walk
 * runs in its chunk-0 task. */
typedef struct gat_tree_start {
  const uint32_t* owner_dev;
  int32_t level; /* 1..15 */
  float xmin, ymax, block;
  const uint32_t* tasks_dev;
  const uint32_t* ntasks_dev; /* 2 entries */
  int64_t tasks_capacity;
} gat_tree_start;

/* Selects, for rectangles q0 .. q0+nq-1 (or, when order_dev is not null, for
 * rectangles order_dev[0 .. nq-1], taken in that order — largest first
 * shortens the tail), the leaves they visit (one wave per rectangle).  Appends each (leaf, query) to leaf_query (2
 * uint32 per entry, order unspecified) while *lq_total < lq_capacity, always
 * advancing *lq_total; adds 1 to node_qcount[leaf] per entry (unless
 * node_qcount_dev is null); writes cand[q] = the points of q's selected leaves.
 * When node_qcount_dev and lq_rank_dev (lq_capacity entries) are both given,
 * lq_rank[i] = entry i's rank among its leaf's entries (the value
 * node_qcount[leaf] had before the entry's increment), which gat_leaf_scatter
 * uses as the entry's place in the leaf's list.  *lq_total and node_qcount
 * must start at zero; leaves with no points are skipped.  start (a host
 * pointer, may be null; nchild 4 only) gives the walk a start table. */
int gat_tree_select(const gat_tree_node* nodes_dev, int64_t nnodes, int32_t nchild, const gat_tree_start* start,
                    const float* boxes_dev,
                    const uint32_t* order_dev, int64_t q0, int64_t nq, uint32_t* node_qcount_dev, uint64_t* cand_dev,
                    uint32_t* leaf_query_dev, uint32_t* lq_rank_dev, int64_t lq_capacity, uint64_t* lq_total_dev,
                    uint32_t* status_dev, void* stream);

/* The same selection for the flat grid (FSG::rangeQueryBatchGPU, FSG.cpp:183-312):
 * rectangle q visits every cell of its column range g1..g2 and row range
 * g3..g4 (the float formula of FSG.cpp:183-186, clamped to the grid), cells
 * numbered row * nx + col from (xmin, ymax).  cells_dev holds one leaf record
 * per cell (first_child -1, the cell's store run); empty cells are skipped.
 * Outputs as gat_tree_select. */
int gat_grid_select(const gat_tree_node* cells_dev, int32_t nx, int32_t ny, float xmin, float ymax, float cell_size,
                    const float* boxes_dev, const uint32_t* order_dev, int64_t q0, int64_t nq,
                    uint32_t* node_qcount_dev, uint64_t* cand_dev, uint32_t* leaf_query_dev, uint32_t* lq_rank_dev,
                    int64_t lq_capacity, uint64_t* lq_total_dev, void* stream);

/* The same selection over an implicit quadtree (MortonGrid, MortonGrid.cpp:8-13,
 * 238-268, 355-393, which replaces the reference's findMatchNodeInQuadTree
 * walk over its isLeaf bitset; no reference GPU counterpart): the GT-quadtree
 * of a 2^L x 2^L Morton grid (L = total_level, 1..14) kept without node
 * records.  Node (level l, id) covers cells [id * 4^(L-l), (id+1) * 4^(L-l))
 * and its box is computed from its first and last cells (cell (row i, col j)
 * spans [xmin + cs*j, xmin + cs*(j+1)] x [ymax - cs*(i+1), ymax - cs*i]).
 * cell_words_dev holds 4^L + 1 pairs of uint32: for cell c, (code, store
 * start of c), where code = (rank << 4) | level at the first cell of each leaf
 * (rank: the leaf's position among the leaves in Morton order) and a value
 * whose low 4 bits name no level elsewhere (0xFFFFFFFF); the last pair's
 * second word is the total point count.  The root is taken as it is, every
 * other node when BoxIntersect(rectangle, box) != 0.  Outputs as
 * gat_tree_select with leaf = rank (node_qcount has nleaves entries); the
 * leaf-major kernels then read leaf records in rank order. */
int gat_morton_select(const uint32_t* cell_words_dev, int32_t total_level, int64_t nleaves, float xmin, float ymax,
                      float cell_size, const float* boxes_dev, const uint32_t* order_dev, int64_t q0, int64_t nq,
                      uint32_t* node_qcount_dev, uint64_t* cand_dev, uint32_t* leaf_query_dev, uint32_t* lq_rank_dev,
                      int64_t lq_capacity, uint64_t* lq_total_dev, uint32_t* status_dev, void* stream);

/* The direct form for small batches: entry i of the (leaf, query) list becomes
 * tasks_dev[i] = {leaf start, leaf count, query} for gat_range_refine_pairs, so
 * no per-node array is touched.  tasks_dev holds lq_capacity tasks. */
int gat_leaf_tasks(const uint32_t* leaf_query_dev, const uint64_t* lq_total_dev, int64_t lq_capacity,
                   const gat_tree_node* nodes_dev, int64_t nnodes, gat_range_task* tasks_dev, void* stream);

/* Groups the selected (leaf, query) entries by leaf.  Entry i, of leaf L and
 * rank r = lq_rank[i] (from the selection), puts the rectangle's id (one
 * uint32) at qlist[node_off[2*L] + r]; every qb-th rank starts a work item
 * (4 uint32: the leaf's store start and count, the item's first list entry,
 * its number of entries <= qb) at items[4 * (node_off[2*L+1] + r/qb)].
 * Entries beyond lq_capacity and ids >= nqueries are ignored; qb is 1..64. */
int gat_leaf_scatter(const uint32_t* leaf_query_dev, const uint64_t* lq_total_dev, int64_t lq_capacity,
                     const gat_tree_node* nodes_dev, int64_t nnodes, const uint32_t* node_off_dev,
                     const uint32_t* node_qcount_dev, const uint32_t* lq_rank_dev, int32_t qb, int64_t nqueries,
                     uint32_t* qlist_dev, int64_t qlist_capacity, uint32_t* items_dev, int64_t items_capacity,
                     void* stream);

/* A run of result pairs of the run-indexed refine: `count` pairs of query q
 * at pairs[first ...]. */
typedef struct gat_range_run {
  uint32_t q;
  uint32_t count;
  uint64_t first;
} gat_range_run;

/* The leaf-major refine with the output laid out by work item (no reference
 * counterpart), in three calls on one stream:
 *
 * gat_leaf_item_regions: work item t's region of `pairs` is
 * [item_base[t], item_base[t+1]), an exclusive scan of its bound (rectangles x
 * leaf points, rounded up to 16 pairs) written to item_base_dev
 * (items_capacity entries; tiles_dev holds gat_scan_tiles(items_capacity)
 * scratch entries); *bound_total_dev receives the sum, which must not exceed
 * the pair capacity of the refine.
 *
 * gat_leaf_refine_runs: an item's hits fill its region from the start,
 * rectangle after rectangle; its list entry e records the rectangle's run
 * (query, hits, first pair) in runs_dev[e] (count 0 when none; runs_capacity
 * must cover the list total); an item it skips as malformed gets empty
 * records.  boxes_dev: the batch's rectangles (4 floats each) that the list's
 * ids index.  No atomics.
 *
 * gat_run_counts: qcount[q] = the pairs of query q summed from the run
 * records of list entries [0, L), L = the last work item's list end (items
 * cover the list in order); qcount must start at zero; when pair_total_dev is
 * not null, the sum of qcount is added to it. */
int gat_leaf_item_regions(const uint32_t* items_dev, const uint32_t* item_total_dev, int64_t items_capacity,
                          uint64_t* item_base_dev, uint64_t* tiles_dev, uint64_t* bound_total_dev, void* stream);
int gat_leaf_refine_runs(const gat_point_store* store, const uint32_t* items_dev, const uint32_t* item_total_dev,
                         int64_t items_capacity, const uint32_t* qlist_dev, int64_t qlist_capacity,
                         const float* boxes_dev, int64_t nqueries, int32_t* pairs_dev, int64_t capacity,
                         gat_range_run* runs_dev, int64_t runs_capacity, const uint64_t* item_base_dev, void* stream);
int gat_run_counts(const gat_point_store* store, const uint32_t* items_dev, const uint32_t* item_total_dev,
                   int64_t items_capacity, int64_t qlist_capacity, const gat_range_run* runs_dev,
                   int64_t runs_capacity, int64_t nqueries, uint32_t* qcount_dev, uint64_t* pair_total_dev,
                   void* stream);

/* Per-query digests of run-indexed pairs (the same value gat_pairs_digest gives
 * for query-major regions): out_dev[q] = wrapping sum of the pair mix over
 * every pair of query q's runs (runs of count 0 contribute nothing).  Clears
 * out_dev first.  Asynchronous. */
int gat_runs_digest(const int32_t* pairs_dev, int64_t npairs, const gat_range_run* runs_dev, int64_t nruns,
                    int64_t nqueries, uint64_t* out_dev, void* stream);

/* The leaf-major refine: per work item the leaf's points and ids are loaded
 * once and tested against each of the item's rectangles; hits go to query q's
 * region pairs[base ...] as (tid, pid), qcount[q] counting them (qcount must
 * start at zero).  *item_total_dev (totals[1] of gat_scan_leaf_lists) gives the
 * number of items; the list's ids index boxes_dev (4 floats per rectangle)
 * and qbase_dev (the per-query region bases, gat_scan_u64 of the candidate
 * counts).  When pair_total_dev is not null, the sum of qcount is added to it
 * afterwards. */
int gat_leaf_refine(const gat_point_store* store, const uint32_t* items_dev, const uint32_t* item_total_dev,
                    int64_t items_capacity, const uint32_t* qlist_dev, int64_t qlist_capacity,
                    const float* boxes_dev, const uint64_t* qbase_dev, int64_t nqueries,
                    uint32_t* qcount_dev, int32_t* pairs_dev, int64_t pairs_capacity, uint64_t* pair_total_dev,
                    void* stream);

/* Stream capture into a replayable graph (hipStreamBeginCapture /
 * hipGraphInstantiate).  Only kernels are recorded by the batched pipelines. */
int gat_graph_begin(void* stream);
int gat_graph_end(void* stream, void** graph_exec);
int gat_graph_launch(void* graph_exec, void* stream);
int gat_graph_destroy(void* graph_exec);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* GAT_HIP_H */
