// The flat C ABI of libgat_host.so (include/gat_host.h) over the C++ classes.
#include <algorithm>
#include <memory>
#include <new>
#include <vector>

#include "FSG.h"
#include "Grid.h"
#include "Autotune.h"
#include "MixedRange.h"
#include "MortonGrid.h"
#include "STIG.h"
#include "TrajFile.h"
#include "Server.h"
#include "SystemTest.h"
#include "Synth.h"
#include "gat_host.h"

namespace {
struct Index {
  enum Kind { kGrid, kFsg, kStig } kind;
  std::unique_ptr<Grid> grid;
  std::unique_ptr<FSG> fsg;
  std::unique_ptr<STIG> stig;
  RangeBatch last;
  int trajNum() const { return kind == kGrid ? grid->trajNum : kind == kFsg ? fsg->trajNum : stig->maxTid; }
};

Index* as_index(void* p) { return static_cast<Index*>(p); }
const MBB* as_boxes(const float* b) { return reinterpret_cast<const MBB*>(b); }

int bad(const char* what) {
  gath_set_error("%s", what);
  return GAT_ERR_ARG;
}
}  // namespace

extern "C" {

int gath_init(int device) {
  GATH_TRY(gat_init(device), "gat_init");
  return GAT_OK;
}

int gath_box_intersect(const float* a, const float* b) {
  return BoxIntersect(MBB(a[0], a[1], a[2], a[3]), MBB(b[0], b[1], b[2], b[3]));
}

int gath_pinbox(const float* box, float x, float y) { return MBB(box[0], box[1], box[2], box[3]).pInBox(x, y); }

uint32_t gath_morton(uint32_t x, uint32_t y) { return getIdxFromXY(x, y); }

void gath_synth_lengths_uniform(uint64_t seed, int64_t n, int lo, int hi, int32_t* out) {
  synth_lengths_uniform(seed, n, lo, hi, out);
}
void gath_synth_lengths_lognormal(uint64_t seed, int64_t n, double median, double sigma, int lo, int hi,
                                  int32_t* out) {
  synth_lengths_lognormal(seed, n, median, sigma, lo, hi, out);
}
void gath_synth_walks(uint64_t seed, int64_t ntraj, const int64_t* offsets, double step_sigma, float* lon,
                      float* lat, int threads) {
  synth_walks(seed, ntraj, offsets, step_sigma, lon, lat, threads);
}
void gath_synth_rects(uint64_t seed, int64_t n, double side_lo, double side_hi, int log_uniform, double hot_frac,
                      int nhot, double hot_sigma, float* out4) {
  synth_rects(seed, n, side_lo, side_hi, log_uniform != 0, hot_frac, nhot, hot_sigma, out4);
}
void gath_synth_queries(uint64_t seed, int nq, const int32_t* qlen, const int64_t* q_offsets,
                        const int64_t* db_offsets, int db_ntraj, const float* db_lon, const float* db_lat,
                        int copy_mode, double noise_sigma, double step_sigma, float* qlon, float* qlat) {
  TrajectoryView db{db_offsets, db_ntraj, db_lon, db_lat};
  synth_queries(seed, nq, qlen, q_offsets, db, copy_mode, noise_sigma, step_sigma, qlon, qlat);
}

void* gath_grid_new(const float* r, float cell_size, int max_points_in_node) {
  auto* ix = new (std::nothrow) Index{Index::kGrid, nullptr, nullptr, nullptr, {}};
  if (!ix) return nullptr;
  ix->grid.reset(new Grid(MBB(r[0], r[1], r[2], r[3]), cell_size, 2, max_points_in_node));
  return ix;
}

void* gath_morton_grid_new(const float* r, float cell_size, int max_points_in_node) {
  auto* ix = new (std::nothrow) Index{Index::kGrid, nullptr, nullptr, nullptr, {}};
  if (!ix) return nullptr;
  ix->grid.reset(new MortonGrid(MBB(r[0], r[1], r[2], r[3]), cell_size, 2, max_points_in_node));
  return ix;
}

void* gath_fsg_new(const float* r, float cell_size) {
  auto* ix = new (std::nothrow) Index{Index::kFsg, nullptr, nullptr, nullptr, {}};
  if (!ix) return nullptr;
  ix->fsg.reset(new FSG(MBB(r[0], r[1], r[2], r[3]), cell_size));
  return ix;
}

void* gath_stig_new(int block_size) {
  auto* ix = new (std::nothrow) Index{Index::kStig, nullptr, nullptr, nullptr, {}};
  if (!ix) return nullptr;
  ix->stig.reset(new STIG());
  ix->stig->blockSize = block_size;
  return ix;
}

void gath_index_free(void* p) { delete as_index(p); }

int gath_index_build(void* p, const int64_t* offsets, int ntraj, const float* lon, const float* lat, int upload) {
  Index* ix = as_index(p);
  if (!ix || !offsets || ntraj < 0) return bad("gath_index_build: bad arguments");
  if (upload < 0 || upload > 2) return bad("gath_index_build: upload must be 0, 1 or 2");
  TrajectoryView db{offsets, ntraj, lon, lat};
  switch (ix->kind) {
    case Index::kGrid:
      ix->grid->deviceBuild = upload == 2;
      return ix->grid->addDatasetToGrid(db, upload != 0);
    case Index::kFsg:
      ix->fsg->deviceBuild = upload == 2;
      return ix->fsg->addDatasetToGrid(db, upload != 0);
    default:
      ix->stig->deviceBuild = upload == 2;
      return ix->stig->initial(ix->stig->blockSize, db, upload != 0);
  }
}

int gath_index_set_threads(void* p, int threads) {
  Index* ix = as_index(p);
  if (!ix) return bad("gath_index_set_threads: null index");
  if (ix->grid) ix->grid->hostThreads = threads;
  if (ix->fsg) ix->fsg->hostThreads = threads;
  if (ix->stig) ix->stig->hostThreads = threads;
  return GAT_OK;
}

int gath_index_info(void* p, int64_t* out) {
  Index* ix = as_index(p);
  if (!ix || !out) return bad("gath_index_info: bad arguments");
  for (int i = 0; i < 8; ++i) out[i] = 0;
  if (ix->grid) {
    Grid& g = *ix->grid;
    out[0] = g.totalPointNum; out[1] = g.trajNum; out[2] = g.cellnum; out[3] = g.cellNum_axis;
    out[4] = int64_t(g.nodes.size()); out[5] = g.leafNum(); out[6] = g.totalLevel;
  } else if (ix->fsg) {
    FSG& f = *ix->fsg;
    out[0] = f.totalPointNum; out[1] = f.trajNum; out[2] = f.cellnum; out[3] = f.cell_num_x; out[6] = f.cell_num_y;
  } else {
    STIG& s = *ix->stig;
    int64_t leaves = 0;
    for (auto& n : s.nodes) leaves += n.leaf;
    out[0] = s.totalPointsNum; out[1] = s.maxTid; out[2] = leaves; out[4] = int64_t(s.nodes.size()); out[5] = leaves;
  }
  return GAT_OK;
}

int64_t gath_index_cell_of(void* p, float lon, float lat) {
  Index* ix = as_index(p);
  if (!ix) return -1;
  if (ix->grid) {
    uint32_t c = ix->grid->cellOf(lon, lat);
    return c == UINT32_MAX ? -1 : int64_t(c);
  }
  if (ix->fsg) return ix->fsg->cellOf(lon, lat);
  return -1;
}

int64_t gath_index_select(void* p, const float* boxes, int nq, uint32_t* tasks_out, int64_t cap,
                          uint64_t* candidates_out) {
  Index* ix = as_index(p);
  if (!ix || nq < 0 || (nq > 0 && !boxes)) return -1;
  std::vector<gat_range_task> tasks;
  std::vector<uint64_t> cand;
  if (ix->grid) ix->grid->selectCandidates(as_boxes(boxes), nq, tasks, cand);
  else if (ix->fsg) ix->fsg->selectCandidates(as_boxes(boxes), nq, tasks, cand);
  else ix->stig->selectCandidates(as_boxes(boxes), nq, tasks, cand);
  const int64_t n = int64_t(tasks.size());
  if (tasks_out && n <= cap)
    for (int64_t i = 0; i < n; ++i) {
      tasks_out[3 * i] = tasks[i].start;
      tasks_out[3 * i + 1] = tasks[i].count;
      tasks_out[3 * i + 2] = tasks[i].qid;
    }
  if (candidates_out)
    for (int q = 0; q < nq; ++q) candidates_out[q] = cand[q];
  return n;
}

int gath_index_range_bitmap(void* p, const float* boxes, int nq, uint8_t* out) {
  Index* ix = as_index(p);
  if (!ix || nq < 0 || (nq > 0 && (!boxes || !out))) return bad("gath_index_range_bitmap: bad arguments");
  std::vector<CPURangeQueryResult> res(static_cast<size_t>(nq));
  int rc;
  if (ix->grid) rc = ix->grid->rangeQueryBatchGPU(as_boxes(boxes), nq, res.data(), nullptr);
  else if (ix->fsg) rc = ix->fsg->rangeQueryBatchGPU(as_boxes(boxes), nq, res.data(), nullptr);
  else rc = ix->stig->rangeQueryGPU(as_boxes(boxes), nq, res.data(), nullptr);
  if (rc) return rc;
  const size_t stride = size_t(ix->trajNum()) + 1;
  for (int q = 0; q < nq; ++q)
    for (size_t t = 0; t < stride; ++t) out[size_t(q) * stride + t] = res[q][t] ? 1 : 0;
  return GAT_OK;
}

int gath_index_range_pairs(void* p, const float* boxes, int nq, int64_t* stats, double* ms) {
  Index* ix = as_index(p);
  if (!ix || nq < 0 || (nq > 0 && !boxes)) return bad("gath_index_range_pairs: bad arguments");
  int rc;
  if (ix->grid) rc = ix->grid->rangeQueryBatchGPUPairs(as_boxes(boxes), nq, ix->last);
  else if (ix->fsg) rc = ix->fsg->rangeQueryBatchGPUPairs(as_boxes(boxes), nq, ix->last);
  else rc = ix->stig->rangeQueryGPUPairs(as_boxes(boxes), nq, ix->last);
  if (rc) return rc;
  if (stats) {
    stats[0] = ix->last.pairTotal;
    stats[1] = ix->last.candidates;
    stats[2] = ix->last.taskNum;
  }
  if (ms) {
    ms[0] = ix->last.selectMs;
    ms[1] = ix->last.uploadMs;
    ms[2] = ix->last.kernelMs;
  }
  return GAT_OK;
}

int gath_index_range_batched(void* p, const float* boxes, int nq, int graph, int64_t* stats, double* ms) {
  Index* ix = as_index(p);
  if (!ix || nq < 0 || (nq > 0 && !boxes)) return bad("gath_index_range_batched: bad arguments");
  int rc;
  if (ix->grid) rc = ix->grid->rangeQueryBatchedGPU(as_boxes(boxes), nq, ix->last, graph != 0);
  else if (ix->stig) rc = ix->stig->rangeQueryBatchedGPU(as_boxes(boxes), nq, ix->last, graph != 0);
  else rc = ix->fsg->rangeQueryBatchedGPU(as_boxes(boxes), nq, ix->last, graph != 0);
  if (rc) return rc;
  if (stats) {
    stats[0] = ix->last.pairTotal;
    stats[1] = ix->last.candidates;
    stats[2] = ix->last.leafQueries;
    stats[3] = ix->last.workItems;
    stats[4] = ix->last.replayed ? 1 : 0;
  }
  if (ms) {
    ms[0] = ix->last.selectMs;
    ms[1] = ix->last.uploadMs;
    ms[2] = ix->last.kernelMs;
    ms[3] = ix->last.refineMs;
  }
  return GAT_OK;
}

int gath_index_last_unique_candidates(void* p, int64_t* points) {
  Index* ix = as_index(p);
  if (!ix || !points) return bad("gath_index_last_unique_candidates: bad arguments");
  return ix->grid ? ix->grid->uniqueCandidates(points)
                  : ix->fsg ? ix->fsg->uniqueCandidates(points) : ix->stig->uniqueCandidates(points);
}

int gath_index_set_batch_params(void* p, int queries_per_item, int direct_max_queries) {
  Index* ix = as_index(p);
  if (!ix || queries_per_item < 1 || queries_per_item > 64 || direct_max_queries < 0)
    return bad("gath_index_set_batch_params: bad arguments");
  RangeEngine& e = ix->grid ? ix->grid->engine() : ix->fsg ? ix->fsg->engine() : ix->stig->engine();
  e.queriesPerItem = queries_per_item;
  e.directMaxQueries = direct_max_queries;
  return GAT_OK;
}

int gath_system_range_test(void* p, int which, const float* boxes, int nq, int64_t* out) {
  Index* ix = as_index(p);
  if (!ix || nq < 0 || (nq > 0 && !boxes) || !out) return bad("gath_system_range_test: bad arguments");
  SystemTest t;
  t.verbose = false;
  int rc;
  const MBB* b = as_boxes(boxes);
  switch (which) {
    case 0:
      if (!ix->grid) return bad("gath_system_range_test: rangeQueryTest needs a Grid");
      t.g = ix->grid.get();
      rc = t.rangeQueryTest(b, nq);
      break;
    case 1:
      if (!ix->stig) return bad("gath_system_range_test: STIGrangeQueryTest needs a STIG");
      t.stig = ix->stig.get();
      rc = t.STIGrangeQueryTest(b, nq);
      break;
    case 2:
      if (!ix->fsg) return bad("gath_system_range_test: FSGrangeQueryTest needs an FSG");
      t.fsg = ix->fsg.get();
      rc = t.FSGrangeQueryTest(b, nq);
      break;
    case 3:
      t.mgrid = dynamic_cast<MortonGrid*>(ix->grid.get());
      if (!t.mgrid) return bad("gath_system_range_test: MortonGridRangeQueryTest needs a MortonGrid");
      rc = t.MortonGridRangeQueryTest(b, nq);
      break;
    default: return bad("gath_system_range_test: which must be 0..3");
  }
  if (rc) return rc;
  out[0] = t.last.candidates;
  out[1] = t.last.pairs;
  ix->last = std::move(t.lastBatch);  // readable through gath_index_last_*
  return GAT_OK;
}

int gath_system_similarity_test(void* p, const float* qlon, const float* qlat, int qlen, int scale, int k,
                                double eps, gath_cpu_similarity_fn cpu_fn, void* cpu_ctx, int32_t* gpu_ids,
                                int32_t* cpu_ids, double* ms) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || qlen <= 0 || !qlon || !qlat || scale < 0 || k < 0 || !gpu_ids || !ms)
    return bad("gath_system_similarity_test: bad arguments");
  if (cpu_fn && !cpu_ids) return bad("gath_system_similarity_test: a CPU leg needs cpu_ids");
  Trajectory t(1);
  for (int i = 0; i < qlen; ++i) t.addSamplePoint(qlon[i], qlat[i]);
  SystemTest st;
  st.verbose = false;
  st.g = ix->grid.get();
  st.g->epsilon = eps;
  if (cpu_fn)
    st.cpuSimilarity = [&](Trajectory* q, int nq, int kk, int* ids) {
      TrajectoryDB db;
      for (int i = 0; i < nq; ++i) {
        for (const SamplePoint& sp : q[i].points) {
          db.lon.push_back(sp.lon);
          db.lat.push_back(sp.lat);
        }
        db.offsets.push_back(int64_t(db.lon.size()));
      }
      if (cpu_fn(cpu_ctx, db.offsets.data(), nq, db.lon.data(), db.lat.data(), kk, ids) != 0) {
        gath_set_error("gath_system_similarity_test: the CPU leg failed");
        return int(GAT_ERR_STATE);
      }
      return int(GAT_OK);
    };
  if (int rc = st.similarityQueryTest(t, scale, k)) return rc;
  std::copy(st.gpuIds.begin(), st.gpuIds.end(), gpu_ids);
  if (cpu_fn) std::copy(st.cpuIds.begin(), st.cpuIds.end(), cpu_ids);
  ms[0] = st.cpuMs;
  ms[1] = st.last.wallMs;
  return GAT_OK;
}

int64_t gath_index_tree_nodes(void* p, gat_tree_node* out, int64_t cap, int32_t* depth) {
  Index* ix = as_index(p);
  if (!ix || !depth || (cap > 0 && !out)) return bad("gath_index_tree_nodes: bad arguments");
  std::vector<gat_tree_node> t;
  int d = 0;
  if (ix->grid) t = ix->grid->treeNodes(&d);
  else if (ix->stig) t = ix->stig->treeNodes(&d);
  else return bad("gath_index_tree_nodes: the FSG index has no tree");
  *depth = d;
  for (int64_t i = 0; i < std::min<int64_t>(cap, int64_t(t.size())); ++i) out[i] = t[size_t(i)];
  return int64_t(t.size());
}

int64_t gath_index_start_table(void* p, uint32_t* out, int64_t cap, int32_t* level, float* geometry) {
  Index* ix = as_index(p);
  if (!ix || !level || !geometry || (cap > 0 && !out)) return bad("gath_index_start_table: bad arguments");
  *level = 0;
  geometry[0] = geometry[1] = geometry[2] = 0.0f;
  if (!ix->grid || dynamic_cast<MortonGrid*>(ix->grid.get())) return 0;  // explicit quadtree walks only
  int depth = 0;
  (void)ix->grid->treeNodes(&depth);
  int k = 0;
  float block = 0;
  const std::vector<uint32_t> t = ix->grid->startTable(depth, &k, &block);
  if (t.empty()) return 0;
  *level = k;
  geometry[0] = ix->grid->range.xmin;
  geometry[1] = ix->grid->range.ymax;
  geometry[2] = block;
  for (int64_t i = 0; i < std::min<int64_t>(cap, int64_t(t.size())); ++i) out[i] = t[size_t(i)];
  return int64_t(t.size());
}

// ------------------------------------------------------------ file formats

namespace {
TrajectoryTable* as_table(void* p) { return static_cast<TrajectoryTable*>(p); }
}  // namespace

void* gath_traj_read(const char* path) {
  if (!path) {
    gath_set_error("gath_traj_read: null path");
    return nullptr;
  }
  auto* t = new (std::nothrow) TrajectoryTable();
  if (!t) return nullptr;
  if (readTrajectoryFile(path, t) != GAT_OK) {
    delete t;
    return nullptr;
  }
  return t;
}

void* gath_traj_ingest_csv(const char* path, const char* base_date) {
  if (!path || !base_date) {
    gath_set_error("gath_traj_ingest_csv: null argument");
    return nullptr;
  }
  auto* t = new (std::nothrow) TrajectoryTable();
  if (!t) return nullptr;
  if (ingestVehicleCsv(path, base_date, t) != GAT_OK) {
    delete t;
    return nullptr;
  }
  return t;
}

void gath_traj_free(void* t) { delete as_table(t); }

int gath_traj_info(void* p, int64_t* sizes, float* bounds) {
  TrajectoryTable* t = as_table(p);
  if (!t || !sizes || !bounds) return bad("gath_traj_info: bad arguments");
  sizes[0] = t->trajNum();
  sizes[1] = t->offsets.back();
  bounds[0] = t->xmin;
  bounds[1] = t->ymin;
  bounds[2] = t->xmax;
  bounds[3] = t->ymax;
  return GAT_OK;
}

int gath_traj_copy(void* p, int64_t* offsets, float* lon, float* lat, int32_t* tid, int32_t* time) {
  TrajectoryTable* t = as_table(p);
  if (!t || !offsets) return bad("gath_traj_copy: bad arguments");
  const size_t np = t->lon.size();
  if (np > 0 && (!lon || !lat || !tid || !time)) return bad("gath_traj_copy: null point array");
  std::copy(t->offsets.begin(), t->offsets.end(), offsets);
  std::copy(t->lon.begin(), t->lon.end(), lon);
  std::copy(t->lat.begin(), t->lat.end(), lat);
  std::copy(t->tid.begin(), t->tid.end(), tid);
  std::copy(t->time.begin(), t->time.end(), time);
  return GAT_OK;
}

int gath_traj_write(const char* path, const int64_t* offsets, int ntraj, const float* lon, const float* lat,
                    const int32_t* tid, const int32_t* time, const float* bounds) {
  if (!path || !offsets || ntraj < 0 || !bounds) return bad("gath_traj_write: bad arguments");
  TrajectoryTable t;
  t.offsets.assign(offsets, offsets + ntraj + 1);
  const int64_t np = t.offsets.back();
  if (t.offsets[0] != 0 || np < 0) return bad("gath_traj_write: offsets must run from 0");
  if (np > 0 && (!lon || !lat || !tid || !time)) return bad("gath_traj_write: null point array");
  t.lon.assign(lon, lon + np);
  t.lat.assign(lat, lat + np);
  t.tid.assign(tid, tid + np);
  t.time.assign(time, time + np);
  t.xmin = bounds[0];
  t.ymin = bounds[1];
  t.xmax = bounds[2];
  t.ymax = bounds[3];
  return writeTrajectoryFile(path, t);
}

int gath_traj_days_between(const char* a, const char* b) {
  if (!a || !b) return -1;
  return daysBetweenDates(a, b);
}

float gath_traj_vehicle_distance(float lat_a, float lon_a, float lat_b, float lon_b) {
  return vehicleDistance(lat_a, lon_a, lat_b, lon_b);
}

int gath_index_fetch_store(void* p, float* xy, uint32_t* id, uint16_t* pid, int64_t* meta) {
  Index* ix = as_index(p);
  if (!ix || !meta) return bad("gath_index_fetch_store: bad arguments");
  RangeEngine& e = ix->grid ? ix->grid->engine() : ix->fsg ? ix->fsg->engine() : ix->stig->engine();
  const gat_point_store& st = e.store();
  meta[0] = st.npoints;
  meta[1] = st.pid_bits;
  if (!xy || st.npoints == 0) return GAT_OK;
  if (!id || (st.pid_bits == 0 && !pid)) return bad("gath_index_fetch_store: null id array");
  const size_t n = size_t(st.npoints);
  GATH_TRY(gat_memcpy_d2h(xy, st.xy_dev, n * 8, e.stream()), "store readback");
  GATH_TRY(gat_memcpy_d2h(id, st.id_dev, n * 4, e.stream()), "store readback");
  if (st.pid_bits == 0) GATH_TRY(gat_memcpy_d2h(pid, st.pid_dev, n * 2, e.stream()), "store readback");
  GATH_TRY(gat_stream_sync(e.stream()), "store readback");
  return GAT_OK;
}

int64_t gath_index_cell_starts(void* p, int64_t* out, int64_t cap) {
  Index* ix = as_index(p);
  if (!ix || (cap > 0 && !out)) return bad("gath_index_cell_starts: bad arguments");
  const std::vector<int64_t>* cs = ix->grid ? &ix->grid->cellStart : ix->fsg ? &ix->fsg->cellStart : nullptr;
  if (!cs) return bad("gath_index_cell_starts: STIG has no cells");
  for (int64_t i = 0; i < std::min<int64_t>(cap, int64_t(cs->size())); ++i) out[i] = (*cs)[size_t(i)];
  return int64_t(cs->size());
}

int gath_index_last_counts(void* p, int64_t* counts) {
  Index* ix = as_index(p);
  if (!ix || !counts) return bad("gath_index_last_counts: bad arguments");
  if (int rc = ix->last.syncHost()) return rc;
  for (int q = 0; q < ix->last.queryNum; ++q) counts[q] = ix->last.count[q];
  return GAT_OK;
}

int gath_index_last_fetch(void* p, int32_t* pairs, int64_t cap) {
  Index* ix = as_index(p);
  if (!ix || (cap > 0 && !pairs)) return bad("gath_index_last_fetch: bad arguments");
  return ix->last.fetchCompact(pairs, cap);
}

int64_t gath_index_last_runs(void* p, gat_range_run* out, int64_t cap) {
  Index* ix = as_index(p);
  if (!ix || (cap > 0 && !out)) return bad("gath_index_last_runs: bad arguments");
  const RangeBatch& b = ix->last;
  if (!b.runIndexed) return 0;
  if (cap >= b.runTotal && b.runTotal > 0)
    if (int rc = ix->last.fetchRuns(out, cap)) return rc;
  return b.runTotal;
}

int gath_index_set_result_layout(void* p, int run_indexed) {
  Index* ix = as_index(p);
  if (!ix) return bad("gath_index_set_result_layout: null index");
  RangeEngine& e = ix->grid ? ix->grid->engine() : ix->fsg ? ix->fsg->engine() : ix->stig->engine();
  e.runIndexed = run_indexed != 0;
  return GAT_OK;
}

int gath_index_last_digests(void* p, uint64_t* digests) {
  Index* ix = as_index(p);
  if (!ix || !digests) return bad("gath_index_last_digests: bad arguments");
  return ix->last.digests(digests);
}

int gath_grid_edr_all(void* p, const float* qlon, const float* qlat, int qlen, double eps, int32_t* out) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || !qlon || !qlat || !out) return bad("gath_grid_edr_all: bad arguments");
  ix->grid->epsilon = eps;
  return ix->grid->edrAll(qlon, qlat, qlen, out);
}

int gath_grid_similarity(void* p, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat, int k,
                         double eps, int32_t* ids, int32_t* dists) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || nq < 0 || k < 0 || !q_offsets) return bad("gath_grid_similarity: bad arguments");
  ix->grid->epsilon = eps;
  TrajectoryView q{q_offsets, nq, qlon, qlat};
  std::vector<TopK> res;
  if (int rc = ix->grid->similarityPlain(q, k, res)) return rc;
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < k; ++j) {
      const bool have = j < int(res[i].ids.size());
      ids[int64_t(i) * k + j] = have ? res[i].ids[j] : 0;
      dists[int64_t(i) * k + j] = have ? res[i].dist[j] : -1;
    }
  return GAT_OK;
}

namespace {
void put_stats(const EdrBatchStats& st, int64_t* stats, double* ms) {
  if (stats) {
    stats[0] = st.pairs;
    stats[1] = st.cells;
    stats[2] = st.jobs;
    stats[3] = st.pruned;
    stats[4] = 0;
  }
  if (ms) {
    ms[0] = st.kernelMs;
    ms[1] = st.filterMs;
    ms[2] = st.hostMs;
    ms[3] = st.uploadMs;
    ms[4] = st.mergeMs;
  }
}
}  // namespace

int gath_grid_edr_batched_all(void* p, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                              double eps, int32_t* dist, int64_t* stats, double* ms) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || nq < 0 || !q_offsets || !dist) return bad("gath_grid_edr_batched_all: bad arguments");
  ix->grid->epsilon = eps;
  TrajectoryView q{q_offsets, nq, qlon, qlat};
  std::vector<int32_t> d;
  EdrBatchStats st;
  if (int rc = ix->grid->edrBatchedAll(q, d, &st)) return rc;
  std::copy(d.begin(), d.end(), dist);
  put_stats(st, stats, ms);
  return GAT_OK;
}

int gath_grid_edr_batched_pairs(void* p, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                                const int32_t* pairs, int64_t npairs, double eps, int32_t* dist, int64_t* stats,
                                double* ms) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || nq < 0 || !q_offsets || npairs < 0 || (npairs > 0 && (!pairs || !dist)))
    return bad("gath_grid_edr_batched_pairs: bad arguments");
  ix->grid->epsilon = eps;
  TrajectoryView q{q_offsets, nq, qlon, qlat};
  std::vector<int32_t> pv(pairs, pairs + 2 * npairs), d;
  EdrBatchStats st;
  if (int rc = ix->grid->edrBatchedPairs(q, pv, d, &st)) return rc;
  std::copy(d.begin(), d.end(), dist);
  put_stats(st, stats, ms);
  return GAT_OK;
}

int gath_grid_similarity_batched(void* p, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat,
                                 int k, double eps, int32_t* ids, int32_t* dists, int64_t* stats, double* ms) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || nq < 0 || k < 0 || !q_offsets) return bad("gath_grid_similarity_batched: bad arguments");
  ix->grid->epsilon = eps;
  TrajectoryView q{q_offsets, nq, qlon, qlat};
  std::vector<TopK> res;
  EdrBatchStats st;
  if (int rc = ix->grid->similarityBatched(q, k, res, &st)) return rc;
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < k; ++j) {
      const bool have = j < int(res[i].ids.size());
      ids[int64_t(i) * k + j] = have ? res[i].ids[j] : 0;
      dists[int64_t(i) * k + j] = have ? res[i].dist[j] : -1;
    }
  put_stats(st, stats, ms);
  return GAT_OK;
}

int gath_grid_mismatch_table(void* p, int32_t* out) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || !out) return bad("gath_grid_mismatch_table: bad arguments");
  const std::vector<int32_t>& F = ix->grid->mismatchTable();
  if (F.empty()) return GAT_ERR_HIP;
  std::copy(F.begin(), F.end(), out);
  return GAT_OK;
}

int gath_grid_set_edr_params(void* p, int split, int per_cu_short, int per_cu_long) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || split < 1 || split > GAT_MAX_TRAJ_LEN || per_cu_short < 0 || per_cu_short > 64 ||
      per_cu_long < 0 || per_cu_long > 64)
    return bad("gath_grid_set_edr_params: bad arguments");
  ix->grid->edrSplit = split;
  ix->grid->edrPerCuShort = per_cu_short;
  ix->grid->edrPerCuLong = per_cu_long;
  return GAT_OK;
}

int gath_grid_set_edr_waves(void* p, int waves_per_pair) {
  Index* ix = as_index(p);
  const int w = waves_per_pair < 0 ? -waves_per_pair : waves_per_pair;
  if (!ix || !ix->grid || (w != 1 && w != 2 && w != 4 && w != 8))
    return bad("gath_grid_set_edr_waves: waves per pair must be 1, 2, 4 or 8 (or -2, -4, -8)");
  ix->grid->edrWavesPerPair = w;
  ix->grid->edrWavesEveryLongPair = waves_per_pair < 0;
  return GAT_OK;
}

int gath_grid_set_topk_device_chain(void* p, int on) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid) return bad("gath_grid_set_topk_device_chain: not a grid index");
  ix->grid->deviceChain = on != 0;
  return GAT_OK;
}

int gath_tune_range(void* p, const float* boxes, int nq, int reps, double* out) {
  Index* ix = as_index(p);
  if (!ix || !out) return bad("gath_tune_range: bad arguments");
  TuneRangeResult r;
  int rc;
  if (ix->grid) rc = tuneRangeBatched(*ix->grid, as_boxes(boxes), nq, reps, &r);
  else if (ix->stig) rc = tuneRangeBatched(*ix->stig, as_boxes(boxes), nq, reps, &r);
  else rc = tuneRangeBatched(*ix->fsg, as_boxes(boxes), nq, reps, &r);
  if (rc) return rc;
  out[0] = r.best;
  for (size_t i = 0; i < r.ms.size(); ++i) out[1 + i] = r.ms[i];
  out[6] = r.directMs;
  out[7] = r.directMaxQueries;
  return GAT_OK;
}

int gath_tune_edr(void* p, const int64_t* q_offsets, int nq, const float* qlon, const float* qlat, double eps,
                  int reps, double* out) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || !q_offsets || !qlon || !qlat || nq < 1 || !out) return bad("gath_tune_edr: bad arguments");
  ix->grid->epsilon = eps;
  TuneEdrResult r;
  if (int rc = tuneEdrBatched(*ix->grid, TrajectoryView{q_offsets, nq, qlon, qlat}, reps, &r)) return rc;
  out[0] = r.split;
  out[1] = r.perCuShort;
  out[2] = r.perCuLong;
  out[3] = r.ms;
  out[4] = r.evaluations;
  out[5] = r.wavesPerPair;
  out[6] = r.smallMs;
  return GAT_OK;
}

void* gath_mixed_new(void* grid, void* stig, void* fsg) {
  Index* g = as_index(grid);
  Index* s = as_index(stig);
  Index* f = as_index(fsg);
  if (!g || !g->grid || !s || !s->stig || !f || !f->fsg) {
    gath_set_error("gath_mixed_new: needs a Grid, a STIG and an FSG index");
    return nullptr;
  }
  return new (std::nothrow) MixedRange(*g->grid, *s->stig, *f->fsg);
}

void gath_mixed_free(void* m) { delete static_cast<MixedRange*>(m); }

int gath_mixed_run(void* m, const float* boxes, int nq, const double* cfg, uint32_t* counts, uint64_t* digests,
                   int8_t* route, double* stats) {
  auto* mr = static_cast<MixedRange*>(m);
  if (!mr || !cfg || nq < 0 || (nq > 0 && (!boxes || !counts || !digests || !route)))
    return bad("gath_mixed_run: bad arguments");
  MixedConfig c;
  c.fsgMaxArea = cfg[0];
  c.gridMaxArea = cfg[1];
  c.groupSize = int(cfg[2]);
  c.graph = cfg[3] != 0;
  MixedResult r;
  if (int rc = mr->run(as_boxes(boxes), nq, c, &r)) return rc;
  std::copy(r.count.begin(), r.count.end(), counts);
  std::copy(r.digest.begin(), r.digest.end(), digests);
  std::copy(r.route.begin(), r.route.end(), route);
  if (stats) {
    const double v[7] = {double(r.pairs), double(r.candidates), double(r.groups), double(r.replays),
                         r.ms,           r.digestMs,           double(mr->captures)};
    for (int i = 0; i < 7; ++i) stats[i] = v[i];
  }
  return GAT_OK;
}

int gath_serve(void* p, const gath_request* reqs, int n, const float* qlon, const float* qlat, int64_t nqpoints,
               const double* cfg, gath_outcome* out, int32_t* topk_ids, int32_t* topk_dists, double* stats) {
  Index* ix = as_index(p);
  if (!ix || !ix->grid || !cfg) return bad("gath_serve: bad arguments (a Grid index and a config are required)");
  ServeConfig c;
  c.rangeSlots = int(cfg[0]);
  c.maxRangeBatch = int(cfg[1]);
  c.maxTopkBatch = int(cfg[2]);
  c.rangeBudgetBytes = cfg[3];
  c.topkBudgetBytes = cfg[4];
  c.k = int(cfg[5]);
  c.eps = cfg[6];
  c.batched = cfg[7] != 0;
  ServeStats st;
  Server server(*ix->grid, c);
  if (int rc = server.run(reqs, n, qlon, qlat, nqpoints, out, topk_ids, topk_dists, &st)) return rc;
  if (stats) {
    const double v[11] = {double(st.rangeBatches), double(st.topkBatches), double(st.rangeDone),
                          double(st.topkDone),     st.wallMs,           st.peakRangeBytes,
                          st.peakTopkBytes,        st.rangeBudgetBytes, st.topkBudgetBytes,
                          double(st.peakTopkSurvivors), double(st.deviceAllocs)};
    for (int i = 0; i < 11; ++i) stats[i] = v[i];
  }
  return GAT_OK;
}

}  // extern "C"
