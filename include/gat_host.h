/*
 * gat_host.h — C ABI of libgat_host.so, the C++ host mirror of the reference's
 * index and query classes (Grid, FSG, STIG, SystemTest; the headers in gat_amd/host).
 *
 * C++ callers use those classes directly, as the reference's main.cpp uses
 * its own (main.cpp:130-164).  This flat ABI exists for callers that cannot
 * link C++ — the Python test harness and bench.py bind it with ctypes.  It has
 * no counterpart in the reference.  Every function returning int returns 0 on
 * success and a gat_hip.h status otherwise; gath_last_error() describes the
 * calling thread's last failure.
 */
#ifndef GAT_HOST_H
#define GAT_HOST_H

#include <stdint.h>

#include "gat_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

const char* gath_last_error(void);
int gath_init(int device);

/* Host geometry: BoxIntersect (MBB.cpp:28-73), MBB::pInBox (MBB.cpp:21-26),
 * getIdxFromXY (Grid.cpp:25-54). */
int gath_box_intersect(const float* a, const float* b);
int gath_pinbox(const float* box, float x, float y);
uint32_t gath_morton(uint32_t x, uint32_t y);

/* Synthetic workloads (gat_amd/host/Synth.h). */
void gath_synth_lengths_uniform(uint64_t seed, int64_t n, int lo, int hi, int32_t* out);
void gath_synth_lengths_lognormal(uint64_t seed, int64_t n, double median, double sigma, int lo, int hi,
                                  int32_t* out);
void gath_synth_walks(uint64_t seed, int64_t ntraj, const int64_t* offsets, double step_sigma, float* lon,
                      float* lat, int threads);
void gath_synth_rects(uint64_t seed, int64_t n, double side_lo, double side_hi, int log_uniform, double hot_frac,
                      int nhot, double hot_sigma, float* out4);
void gath_synth_queries(uint64_t seed, int nq, const int32_t* qlen, const int64_t* q_offsets,
                        const int64_t* db_offsets, int db_ntraj, const float* db_lon, const float* db_lat,
                        int copy_mode, double noise_sigma, double step_sigma, float* qlon, float* qlat);

/* On-disk formats (PreProcess.cpp).  gath_traj_read parses the preprocessed
 * text file (readTraFromFormatedFile, PreProcess.cpp:225-315: trajectory
 * count, "xmin,ymin,xmax,ymax", lengths, then "lat,lon,tid,time;" per point,
 * latitude first); gath_traj_ingest_csv builds trajectories from raw vehicle
 * records (the PreProcess constructor, PreProcess.cpp:48-201: field 2 vehicle
 * id, 3 longitude, 4 latitude, 8 "yyyy-mm-dd hh:mm:ss", times in seconds
 * from base_date; a vehicle's trajectory ends at 1024 points or a gap over
 * 3600 s, points at 50 m/s or more are dropped, Trajectory.cpp:22-48).  Both
 * return a table handle (NULL on error, see gath_last_error) holding
 * trajectories 1..n: gath_traj_info gives sizes[0] = n, sizes[1] = points and
 * the bounds; gath_traj_copy copies offsets (n + 1), lon, lat, tid and time.
 * gath_traj_write writes the preprocessed format (writeTraDataToFile,
 * PreProcess.cpp:203-223).  gath_traj_days_between and
 * gath_traj_vehicle_distance expose the ingest's date arithmetic
 * (DateTime.cpp:44-95) and speed estimate (CalcInAxis.h:9-21). */
void* gath_traj_read(const char* path);
void* gath_traj_ingest_csv(const char* path, const char* base_date);
void gath_traj_free(void* table);
int gath_traj_info(void* table, int64_t* sizes, float* bounds);
int gath_traj_copy(void* table, int64_t* offsets, float* lon, float* lat, int32_t* tid, int32_t* time);
int gath_traj_write(const char* path, const int64_t* offsets, int ntraj, const float* lon, const float* lat,
                    const int32_t* tid, const int32_t* time, const float* bounds);
int gath_traj_days_between(const char* date_a, const char* date_b);
float gath_traj_vehicle_distance(float lat_a, float lon_a, float lat_b, float lon_b);

/* Indexes.  range = (xmin, ymin, xmax, ymax). */
void* gath_grid_new(const float* range, float cell_size, int max_points_in_node);
/* MortonGrid (MortonGrid.h:27-86): the GT-quadtree kept implicitly; same
 * answers as gath_grid_new's index, at most 2^14 cells per axis. */
void* gath_morton_grid_new(const float* range, float cell_size, int max_points_in_node);
void* gath_fsg_new(const float* range, float cell_size);
void* gath_stig_new(int block_size);
void gath_index_free(void* index);
/* Builds the index over trajectories 1..ntraj (offsets: ntraj+1 entries); with
 * upload = 0 only the host structures are built (no GPU needed); 1 buckets the
 * points on the host and uploads the store; 2 uploads the trajectories and
 * builds the store on the device: the same store for Grid, MortonGrid and
 * FSG; for STIG the same tree shape and medians, with points of equal
 * coordinates possibly on the other side of a median (same answers). */
int gath_index_build(void* index, const int64_t* offsets, int ntraj, const float* lon, const float* lat,
                     int upload);
/* out[0..7]: points, trajectories, cells (or leaves for STIG), cells per axis
 * (Grid) / columns (FSG), nodes, leaves, total levels (Grid) / rows (FSG), 0. */
int gath_index_info(void* index, int64_t* out);
int gath_index_set_threads(void* index, int threads);
/* Host candidate selection only: returns the number of tasks; when it is not
 * above `cap`, tasks (start, count, qid triples) are written to tasks_out.
 * candidates_out[q] (nq entries) receives each query's candidate points. */
int64_t gath_index_select(void* index, const float* boxes, int nq, uint32_t* tasks_out, int64_t cap,
                          uint64_t* candidates_out);
/* Cell of a point (Grid: Morton id; FSG: row-major id; -1 outside). */
int64_t gath_index_cell_of(void* index, float lon, float lat);
/* The reference-signature path: trajectory bitmap (nq * (ntraj + 1) bytes). */
int gath_index_range_bitmap(void* index, const float* boxes, int nq, uint8_t* out);
/* Compacted-pair path.  stats[0..2] = pairs, candidates, tasks; ms[0..2] =
 * host selection, task upload, refine kernel.  The result stays on the device
 * until the next call. */
int gath_index_range_pairs(void* index, const float* boxes, int nq, int64_t* stats, double* ms);
/* Batched form (Grid, STIG and FSG): device selection and the leaf-major join, the
 * whole batch at once; graph != 0 captures the device pipeline and replays it
 * on later calls that fit.  stats[0..4] = pairs, candidates, (leaf, query)
 * entries, work items, replayed (0/1); ms[0..3] = device selection, wall time,
 * join kernels (the whole pipeline when replayed), the refine kernel alone
 * (0 when replayed).  Results are read with the gath_index_last_* calls. */
int gath_index_range_batched(void* index, const float* boxes, int nq, int graph, int64_t* stats, double* ms);
/* After a batched run: points of the leaves at least one rectangle selected. */
int gath_index_last_unique_candidates(void* index, int64_t* points);
/* Batched-EDR tunables of a Grid: the length split between the short and long
 * buckets (1..GAT_MAX_TRAJ_LEN) and workgroups per CU of each bucket's launch
 * (0..64, 0 = the kernel's default).  Pairs longer than max(split, 1024)
 * points run in a third launch at the kernel's default. */
int gath_grid_set_edr_params(void* grid, int split, int per_cu_short, int per_cu_long);
/* Waves per pair (1, 2, 4 or 8) for the long pairs of batches that fit one
 * round of the device's waves (top-k survivors, serving batches).  Only the
 * pairs that one wave would keep busy longer than the batch's longest pair
 * takes over that many waves get them; the others run one pair per wave.
 * -2, -4 or -8 give every long pair of such a batch 2, 4 or 8 waves. */
int gath_grid_set_edr_waves(void* grid, int waves_per_pair);
/* Top-k batches (k <= 1024, at most 8192 queries) chained on the device —
 * filter, jobs, DP, survivor lists and merge with one readback (on, the
 * default) — or with the jobs and survivor lists built on the host (off). */
int gath_grid_set_topk_device_chain(void* grid, int on);
/* Tunables of the batched form: rectangles per leaf-major work item (1..64),
 * and the largest batch that takes the direct form (one refine task per
 * selected leaf and rectangle, nothing sized by the tree touched). */
int gath_index_set_batch_params(void* index, int queries_per_item, int direct_max_queries);
/* The reference's SystemTest range entry points (SystemTest.cpp:24-102,
 * 251-384, 386-445) over the rectangles `boxes`: which = 0 rangeQueryTest
 * (a Grid or MortonGrid index), 1 STIGrangeQueryTest, 2 FSGrangeQueryTest,
 * 3 MortonGridRangeQueryTest.  out[0] = candidate points scanned, out[1] =
 * result pairs; the pairs are then the index's last result (gath_index_last_*). */
int gath_system_range_test(void* index, int which, const float* boxes, int nq, int64_t* out);
/* The CPU top-k a caller hands to gath_system_similarity_test: nq queries
 * (q_offsets, nq + 1 entries, into qlon / qlat), k; writes nq * k trajectory
 * ids to `ids` and returns 0, or non-zero on failure. */
typedef int (*gath_cpu_similarity_fn)(void* ctx, const int64_t* q_offsets, int nq, const float* qlon,
                                      const float* qlat, int k, int32_t* ids);
/* The reference's SystemTest::similarityQueryTest (SystemTest.cpp:164-249) on
 * a Grid: the query (qlen points) replicated `scale` times; the CPU leg
 * (SystemTest.cpp:207: cpu_fn, skipped when NULL — this library has no CPU
 * query path) timed, then the GPU leg, Grid::SimilarityQueryBatchOnGPU
 * (SystemTest.cpp:226) with match threshold eps.  gpu_ids and cpu_ids (may be
 * NULL) receive scale * k ids each; ms[0] = CPU leg, ms[1] = GPU leg. */
int gath_system_similarity_test(void* grid, const float* qlon, const float* qlat, int qlen, int scale, int k,
                                double eps, gath_cpu_similarity_fn cpu_fn, void* cpu_ctx, int32_t* gpu_ids,
                                int32_t* cpu_ids, double* ms);
/* The index's tree as the batched form uploads it (breadth-first
 * gat_tree_node; a MortonGrid's implicit tree gives its leaves in Morton
 * order); returns the node count, copies up to cap nodes. */
int64_t gath_index_tree_nodes(void* index, gat_tree_node* out, int64_t cap, int32_t* depth);
/* The start table a GT-quadtree index gives its device walk (gat_tree_start):
 * writes up to cap owners, *level and geometry[0..2] = xmin, ymax, block side;
 * returns the number of owners (4^level), 0 when the index has none. */
int64_t gath_index_start_table(void* index, uint32_t* out, int64_t cap, int32_t* level, float* geometry);
/* The resident point store: meta[0] = points, meta[1] = pid_bits (0: point ids
 * in `pid`); with xy non-null, copies xy (2 floats per point), id and, when
 * pid_bits is 0, pid to the host. */
int gath_index_fetch_store(void* index, float* xy, uint32_t* id, uint16_t* pid, int64_t* meta);
/* Grid / FSG: the cell offsets (cells + 1 entries; returns the count, copies
 * up to cap). */
int64_t gath_index_cell_starts(void* index, int64_t* out, int64_t cap);
int gath_index_last_counts(void* index, int64_t* counts);
/* Pairs of the last result, query by query (query q's pairs follow those of
 * queries 0..q-1); cap in pairs. */
int gath_index_last_fetch(void* index, int32_t* pairs, int64_t cap);
/* Run-indexed batches (leaf-major form, the default): the run records of the
 * last batch (query, count, first pair into its compacted pairs); returns
 * their number (0 for other layouts) and copies them when cap holds them all. */
int64_t gath_index_last_runs(void* index, gat_range_run* out, int64_t cap);
/* Leaf-major result layout of an index's batched form: 1 (default) pairs by
 * work item with a run index, 0 per-query regions.  Counts, digests and
 * gath_index_last_fetch are the same either way. */
int gath_index_set_result_layout(void* index, int run_indexed);
int gath_index_last_digests(void* index, uint64_t* digests);

/* Similarity on a Grid. */
int gath_grid_edr_all(void* grid, const float* qlon, const float* qlat, int qlen, double eps, int32_t* out);
int gath_grid_similarity(void* grid, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat, int k,
                         double eps, int32_t* ids, int32_t* dists);
/* Batched forms.  stats[0..4] = pairs run, DP cells, jobs, pairs resolved by
 * the all-mismatch table, 0; ms[0..4] = DP kernel, filter kernel, host work
 * (job building; the host merge for k > 1024), query store packing and upload,
 * the device merge stage (survivor lists, upload, merge kernel, readback). */
int gath_grid_edr_batched_all(void* grid, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                              double eps, int32_t* dist, int64_t* stats, double* ms);
/* The batched EDR of given (query index, trajectory index) pairs, both
 * 0-based: dist[i] for pairs[2i], pairs[2i + 1]; stats and ms as
 * gath_grid_edr_batched_all. */
int gath_grid_edr_batched_pairs(void* grid, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                                const int32_t* pairs, int64_t npairs, double eps, int32_t* dist, int64_t* stats,
                                double* ms);
int gath_grid_similarity_batched(void* grid, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                                 int k, double eps, int32_t* ids, int32_t* dists, int64_t* stats, double* ms);
/* The all-mismatch table, (GAT_MAX_TRAJ_LEN + 1)^2 int32. */
int gath_grid_mismatch_table(void* grid, int32_t* out);


/* ---------------------------------------------------------------- autotuner
 * Times candidate launch shapes of the batched forms on the caller's sample and
 * leaves the fastest set on the index (no reference counterpart).
 * gath_tune_range: rectangles per leaf-major work item over {4, 8, 16, 32, 64},
 * then the leaf-major against the direct form at this batch size; out[0] = the
 * choice, out[1..5] = median device pipeline ms of each, out[6] = the direct
 * form's ms, out[7] = the resulting direct-form threshold (rectangles).
 * gath_tune_edr (Grid): bucket split and workgroups per CU of each bucket, by
 * coordinate descent, then waves per long pair on a one-round batch (the
 * longest query against 8192 trajectories spread evenly over the database);
 * out[0..6] = split,
 * per-CU short, per-CU long, best kernel ms, evaluations, waves per pair,
 * that batch's kernel ms. */
int gath_tune_range(void* index, const float* boxes, int nq, int reps, double* out);
int gath_tune_edr(void* grid, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat, double eps,
                  int reps, double* out);

/* ---------------------------------------------------------------- mixed
 * One batch of rectangles over the three indexes of a data set: each goes to
 * FSG (area < cfg[0]), the GT-quadtree (area < cfg[1]) or STIG, in groups of
 * cfg[2] rectangles per index, the three device pipelines of a group captured
 * as one graph and replayed (cfg[3] != 0).  Per rectangle: counts, digests
 * (as gat_pairs_digest), route (0 FSG, 1 GT-quadtree, 2 STIG); stats[0..6] =
 * pairs, candidates, groups, graph replays, wall ms, of which digest ms, and
 * the graphs this object has captured so far (a group whose engines' buffers
 * moved since the capture, e.g. grown by an index's own batched call, is run
 * step by step and recaptured instead of replayed). */
void* gath_mixed_new(void* grid, void* stig, void* fsg);
void gath_mixed_free(void* mixed);
int gath_mixed_run(void* mixed, const float* boxes, int nq, const double* cfg, uint32_t* counts, uint64_t* digests,
                   int8_t* route, double* stats);

/* ---------------------------------------------------------------- serving
 * The serving loop over a Grid (no reference counterpart): requests arrive at
 * their arrival times, range slots and the top-k slot batch every waiting
 * request that fits their device-memory budget, and each request's admission
 * and completion times are recorded. */
enum { GATH_REQ_RANGE = 0, GATH_REQ_TOPK = 1 };

typedef struct gath_request {
  int32_t kind;      /* GATH_REQ_RANGE or GATH_REQ_TOPK */
  float box[4];      /* range: xmin, ymin, xmax, ymax */
  int32_t qlen;      /* top-k: query length (1..GAT_MAX_TRAJ_LEN) */
  int64_t qoff;      /* top-k: first query point in the request point arrays */
  double arrival_ms; /* from the start of the run; non-decreasing */
} gath_request;

typedef struct gath_outcome {
  double admit_ms; /* the batch carrying the request was launched */
  double done_ms;  /* the request's result was on the host */
  int32_t batch;   /* batch number (range and top-k batches are numbered separately) */
  uint32_t count;  /* range: result pairs */
  uint64_t digest; /* range: order-independent digest of the pairs (gat_pairs_digest) */
} gath_outcome;

/* cfg[0..7] = range slots, max rectangles per range batch, max queries per
 * top-k batch, range batch budget (bytes), top-k batch budget (bytes), k, eps,
 * batched (0: the plain per-request path).  topk_ids / topk_dists receive k
 * entries per request (used for top-k requests; ids 0 / distance -1 pad a
 * short answer).  stats[0..10] = range batches, top-k batches, range done,
 * top-k done, wall ms, largest range batch charge, largest top-k batch
 * charge (bytes), range budget and top-k budget after the free-memory check,
 * most filter survivors of one top-k batch, device allocations made while
 * serving (buffers that had to grow). */
int gath_serve(void* grid, const gath_request* reqs, int n, const float* qlon, const float* qlat, int64_t nqpoints,
               const double* cfg, gath_outcome* out, int32_t* topk_ids, int32_t* topk_dists, double* stats);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* GAT_HOST_H */
