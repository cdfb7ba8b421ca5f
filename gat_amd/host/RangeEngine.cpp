#include "RangeEngine.h"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <thread>

namespace {
double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}
size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
}  // namespace

void TaskList::addRun(uint64_t start, uint64_t count, uint32_t qid) {
  if (count == 0) return;
  ++runsAdded;
  pointsAdded += count;
  if (open_ && qid == runQid_ && runStart_ + runCount_ == start) {
    runCount_ += count;
    return;
  }
  flush();
  open_ = true;
  runStart_ = start;
  runCount_ = count;
  runQid_ = qid;
}

void TaskList::flush() {
  if (!open_) return;
  for (uint64_t off = 0; off < runCount_; off += kTaskChunk) {
    const uint64_t c = std::min<uint64_t>(kTaskChunk, runCount_ - off);
    tasks.push_back(gat_range_task{uint32_t(runStart_ + off), uint32_t(c), runQid_});
  }
  candidates[runQid_ - firstQuery] += runCount_;
  open_ = false;
}

void RangeEngine::select(int nq, const Selector& selector, int threads, std::vector<gat_range_task>& tasks,
                         std::vector<uint64_t>& candidates) {
  threads = std::max(1, std::min(threads, nq));
  std::vector<TaskList> lists(size_t(std::max(threads, 1)));
  auto work = [&](int part) {
    const int q0 = int(int64_t(nq) * part / threads), q1 = int(int64_t(nq) * (part + 1) / threads);
    TaskList& l = lists[part];
    l.firstQuery = uint32_t(q0);
    l.candidates.assign(size_t(q1 - q0), 0);
    for (int q = q0; q < q1; ++q) {
      selector(q, l);
      l.flush();
    }
  };
  if (threads <= 1) {
    if (nq > 0) work(0);
  } else {
    std::vector<std::thread> pool;
    for (int p = 0; p < threads; ++p) pool.emplace_back(work, p);
    for (auto& t : pool) t.join();
  }
  size_t total = 0;
  for (auto& l : lists) total += l.tasks.size();
  tasks.clear();
  tasks.reserve(total);
  candidates.assign(size_t(std::max(nq, 0)), 0);
  for (auto& l : lists) {
    tasks.insert(tasks.end(), l.tasks.begin(), l.tasks.end());
    for (size_t i = 0; i < l.candidates.size(); ++i) candidates[l.firstQuery + i] = l.candidates[i];
  }
}

namespace {
int bit_length(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }
}  // namespace

int RangeEngine::upload(const float* xy, const int32_t* tp, int64_t n) {
  if (int rc = stream_.create()) return rc;
  n_ = n;
  uint32_t max_tid = 0, max_pid = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (tp[2 * i] < 0 || tp[2 * i + 1] < 0) {
      gath_set_error("store upload: negative id at point %lld", (long long)i);
      return GAT_ERR_ARG;
    }
    max_tid = std::max(max_tid, uint32_t(tp[2 * i]));
    max_pid = std::max(max_pid, uint32_t(tp[2 * i + 1]));
  }
  const int pid_bits = std::max(bit_length(max_pid), 1);
  const bool packed = bit_length(max_tid) + pid_bits <= 32;
  if (!packed && max_pid > 0xFFFFu) {
    gath_set_error("store upload: point ids up to %u do not fit 16 bits", max_pid);
    return GAT_ERR_ARG;
  }
  const size_t cnt = size_t(std::max<int64_t>(n, 1));
  if (int rc = xy_.reserve(cnt * 8)) return rc;
  if (int rc = id_.reserve(cnt * 4)) return rc;
  std::vector<uint32_t> id(static_cast<size_t>(n));
  std::vector<uint16_t> pid;
  if (packed) {
    pid_.release();
    for (int64_t i = 0; i < n; ++i) id[i] = (uint32_t(tp[2 * i]) << pid_bits) | uint32_t(tp[2 * i + 1]);
  } else {
    if (int rc = pid_.reserve(cnt * 2)) return rc;
    pid.resize(size_t(n));
    for (int64_t i = 0; i < n; ++i) {
      id[i] = uint32_t(tp[2 * i]);
      pid[i] = uint16_t(tp[2 * i + 1]);
    }
  }
  GATH_TRY(gat_memcpy_h2d(xy_.get(), xy, size_t(n) * 8, stream_.get()), "store upload");
  GATH_TRY(gat_memcpy_h2d(id_.get(), id.data(), size_t(n) * 4, stream_.get()), "store upload");
  if (!packed) GATH_TRY(gat_memcpy_h2d(pid_.get(), pid.data(), size_t(n) * 2, stream_.get()), "store upload");
  store_ = gat_point_store{xy_.as<float>(), id_.as<uint32_t>(), packed ? nullptr : pid_.as<uint16_t>(), n,
                           packed ? pid_bits : 0};
  return stream_.sync();
}

// The id packing of a store over `db` (as upload() chooses it): pid_bits > 0
// packs (tid, pid) into 32 bits, 0 keeps a separate 16-bit pid array.
static int store_packing(const TrajectoryView& db, int* pidBitsOut) {
  if (db.trajNum < 0 || !db.offsets || db.offsets[0] != 0) {
    gath_set_error("store build: offsets must start at 0");
    return GAT_ERR_ARG;
  }
  uint32_t max_pid = 0;
  for (int t = 1; t <= db.trajNum; ++t) {
    const int64_t len = db.offsets[t] - db.offsets[t - 1];
    if (len < 0) {
      gath_set_error("store build: offsets decrease at trajectory %d", t);
      return GAT_ERR_ARG;
    }
    if (len > 0) max_pid = std::max(max_pid, uint32_t(std::min<int64_t>(len - 1, UINT32_MAX)));
  }
  const int pid_bits = std::max(bit_length(max_pid), 1);
  const bool packed = bit_length(uint32_t(db.trajNum)) + pid_bits <= 32;
  if (!packed && max_pid > 0xFFFFu) {
    gath_set_error("store build: point ids up to %u do not fit 16 bits", max_pid);
    return GAT_ERR_ARG;
  }
  *pidBitsOut = packed ? pid_bits : 0;
  return GAT_OK;
}

// Uploads the trajectories for a device build and sizes the store arrays.
int RangeEngine::stageBuild(const TrajectoryView& db, int pidBits, DeviceBuffer& lon, DeviceBuffer& lat,
                            DeviceBuffer& off) {
  const int64_t n = db.pointNum();
  const size_t cnt = size_t(std::max<int64_t>(n, 1));
  void* s = stream_.get();
  if (int rc = lon.reserve(cnt * 4)) return rc;
  if (int rc = lat.reserve(cnt * 4)) return rc;
  if (int rc = off.reserve((size_t(db.trajNum) + 1) * 8)) return rc;
  if (int rc = xy_.reserve(cnt * 8)) return rc;
  if (int rc = id_.reserve(cnt * 4)) return rc;
  if (pidBits > 0)
    pid_.release();
  else if (int rc = pid_.reserve(cnt * 2))
    return rc;
  if (n > 0) {
    GATH_TRY(gat_memcpy_h2d(lon.get(), db.lon, size_t(n) * 4, s), "trajectory upload");
    GATH_TRY(gat_memcpy_h2d(lat.get(), db.lat, size_t(n) * 4, s), "trajectory upload");
  }
  GATH_TRY(gat_memcpy_h2d(off.get(), db.offsets, (size_t(db.trajNum) + 1) * 8, s), "trajectory upload");
  n_ = 0;
  store_ = gat_point_store{};
  return GAT_OK;
}

int RangeEngine::buildStore(const gat_cell_grid& grid, const TrajectoryView& db, std::vector<int64_t>* cellStart,
                            int64_t* badPoint) {
  if (int rc = stream_.create()) return rc;
  *badPoint = -1;
  int pid_bits = 0;
  if (int rc = store_packing(db, &pid_bits)) return rc;
  const int64_t n = db.pointNum();
  const uint64_t cellnum = uint64_t(grid.cells_x) * uint64_t(grid.cells_y);
  void* s = stream_.get();
  DeviceBuffer lon, lat, off, starts;
  if (int rc = starts.reserve((cellnum + 1) * 8)) return rc;
  if (int rc = stageBuild(db, pid_bits, lon, lat, off)) return rc;
  GATH_TRY(gat_cell_store_build(&grid, lon.as<float>(), lat.as<float>(), off.as<int64_t>(), db.trajNum, n, pid_bits,
                                xy_.as<float>(), id_.as<uint32_t>(), pid_bits > 0 ? nullptr : pid_.as<uint16_t>(),
                                starts.as<int64_t>(), badPoint, s),
           "store build");
  if (*badPoint >= 0) return GAT_OK;
  cellStart->assign(size_t(cellnum) + 1, 0);
  GATH_TRY(gat_memcpy_d2h(cellStart->data(), starts.get(), (cellnum + 1) * 8, s), "cell offsets");
  if (int rc = stream_.sync()) return rc;
  n_ = n;
  store_ = gat_point_store{xy_.as<float>(), id_.as<uint32_t>(), pid_bits > 0 ? nullptr : pid_.as<uint16_t>(), n,
                           pid_bits};
  return GAT_OK;
}

int RangeEngine::buildKdStore(const TrajectoryView& db, const KdSkeleton& k, std::vector<float>* medians,
                              std::vector<float>* leafBoxes) {
  if (int rc = stream_.create()) return rc;
  int pid_bits = 0;
  if (int rc = store_packing(db, &pid_bits)) return rc;
  const int64_t n = db.pointNum();
  if (n < 1 || k.leafLo.empty() || k.depthFirst.empty()) {
    gath_set_error("kd store build: no points or no leaves");
    return GAT_ERR_ARG;
  }
  DeviceBuffer lon, lat, off;
  if (int rc = stageBuild(db, pid_bits, lon, lat, off)) return rc;
  medians->assign(k.splitLo.size(), 0.0f);
  leafBoxes->assign(k.leafLo.size() * 4, 0.0f);
  GATH_TRY(gat_kd_store_build(lon.as<float>(), lat.as<float>(), off.as<int64_t>(), db.trajNum, n, pid_bits,
                              k.splitLo.data(), k.splitHi.data(), k.depthFirst.data(),
                              int32_t(k.depthFirst.size()) - 1, k.leafLo.data(), k.leafHi.data(),
                              int64_t(k.leafLo.size()), xy_.as<float>(), id_.as<uint32_t>(),
                              pid_bits > 0 ? nullptr : pid_.as<uint16_t>(), medians->data(), leafBoxes->data(),
                              stream_.get()),
           "kd store build");
  n_ = n;
  store_ = gat_point_store{xy_.as<float>(), id_.as<uint32_t>(), pid_bits > 0 ? nullptr : pid_.as<uint16_t>(), n,
                           pid_bits};
  return GAT_OK;
}

// Copies boxes, tasks (and qbase) into one pinned staging block and then to the
// device in as few transfers as the layout allows.
int RangeEngine::stage(const MBB* bounds, int nq, const std::vector<gat_range_task>& tasks, const uint64_t* qbase) {
  static_assert(sizeof(MBB) == 16, "MBB must be four packed floats");
  const size_t box_bytes = align_up(size_t(nq) * 16, 256);
  const size_t base_bytes = qbase ? align_up(size_t(nq) * 8, 256) : 0;
  const size_t task_bytes = align_up(tasks.size() * sizeof(gat_range_task), 256);
  const size_t total = box_bytes + base_bytes + task_bytes;
  if (int rc = hostStage_.reserve(total)) return rc;
  if (int rc = devBoxes_.reserve(box_bytes + base_bytes)) return rc;
  if (int rc = devTasks_.reserve(std::max<size_t>(task_bytes, 256))) return rc;
  char* h = hostStage_.as<char>();
  std::memcpy(h, bounds, size_t(nq) * 16);
  if (qbase) std::memcpy(h + box_bytes, qbase, size_t(nq) * 8);
  std::memcpy(h + box_bytes + base_bytes, tasks.data(), tasks.size() * sizeof(gat_range_task));
  GATH_TRY(gat_memcpy_h2d(devBoxes_.get(), h, box_bytes + base_bytes, stream_.get()), "task upload");
  GATH_TRY(gat_memcpy_h2d(devTasks_.get(), h + box_bytes + base_bytes, tasks.size() * sizeof(gat_range_task),
                          stream_.get()),
           "task upload");
  return GAT_OK;
}

int RangeEngine::runPairs(const MBB* bounds, int nq, const Selector& selector, int threads, RangeBatch& out) {
  if (int rc = stream_.create()) return rc;
  out.queryNum = nq;
  out.hostStale = false;
  out.stream = stream_.get();
  // Per-query regions: nothing of an earlier batched run's layout applies.
  out.runIndexed = false;
  out.runTotal = out.pairSpan = 0;
  out.replayed = false;
  out.leafQueries = out.workItems = 0;
  out.refineMs = 0;
  double t0 = now_ms();
  std::vector<gat_range_task> tasks;
  std::vector<uint64_t> cand;
  select(nq, selector, threads, tasks, cand);
  double t1 = now_ms();
  out.selectMs = t1 - t0;
  out.taskNum = int64_t(tasks.size());
  out.base.assign(size_t(nq), 0);
  uint64_t total = 0;
  for (int q = 0; q < nq; ++q) {
    out.base[q] = total;
    total += cand[q];
  }
  out.candidates = int64_t(total);
  if (int rc = out.pairs.reserve(size_t(std::max<uint64_t>(total, 1)) * 8)) return rc;
  if (int rc = out.qcount.reserve(size_t(std::max(nq, 1)) * 4)) return rc;
  if (int rc = out.qbase.reserve(size_t(std::max(nq, 1)) * 8)) return rc;
  if (int rc = stage(bounds, nq, tasks, out.base.data())) return rc;
  // qbase was staged right after the boxes; copy it to the batch's own buffer.
  GATH_TRY(gat_memcpy_d2d(out.qbase.get(), devBoxes_.as<char>() + align_up(size_t(nq) * 16, 256), size_t(nq) * 8,
                          stream_.get()),
           "qbase copy");
  GATH_TRY(gat_memset(out.qcount.get(), 0, size_t(nq) * 4, stream_.get()), "count reset");
  if (int rc = stream_.sync()) return rc;
  double t2 = now_ms();
  out.uploadMs = t2 - t1;
  if (int rc = timer_.start(stream_.get())) return rc;
  GATH_TRY(gat_range_refine_pairs(&store_, devTasks_.as<gat_range_task>(), int64_t(tasks.size()),
                                  devBoxes_.as<float>(), nq, out.qbase.as<uint64_t>(), out.qcount.as<uint32_t>(),
                                  out.pairs.as<int32_t>(), int64_t(total), stream_.get()),
           "range refine");
  if (int rc = timer_.stop(stream_.get())) return rc;
  out.count.assign(size_t(nq), 0);
  GATH_TRY(gat_memcpy_d2h(out.count.data(), out.qcount.get(), size_t(nq) * 4, stream_.get()), "count readback");
  if (int rc = stream_.sync()) return rc;
  if (int rc = timer_.elapsed(&out.kernelMs)) return rc;
  out.pairTotal = 0;
  for (int q = 0; q < nq; ++q) {
    if (out.count[q] > cand[q]) {
      gath_set_error("range refine: query %d produced %u pairs for %llu candidates", q, out.count[q],
                     (unsigned long long)cand[q]);
      return GAT_ERR_CAPACITY;
    }
    out.pairTotal += out.count[q];
  }
  return GAT_OK;
}

int RangeEngine::runBitmap(const MBB* bounds, int nq, const Selector& selector, int threads, int maxTrajNum,
                           uint8_t* resultHost, bool stigHandler) {
  if (int rc = stream_.create()) return rc;
  std::vector<gat_range_task> tasks;
  std::vector<uint64_t> cand;
  select(nq, selector, threads, tasks, cand);
  if (int rc = stage(bounds, nq, tasks, nullptr)) return rc;
  auto handler = stigHandler ? gatRangeQuerySTIGHandler : gatRangeQueryTestHandler;
  GATH_TRY(handler(&store_, devTasks_.as<gat_range_task>(), int(tasks.size()), devBoxes_.as<float>(), resultHost,
                   maxTrajNum, nq, stream_.get()),
           "range handler");
  return GAT_OK;
}

int RangeBatch::syncHost() {
  if (!hostStale) return GAT_OK;
  count.resize(size_t(queryNum));
  base.resize(size_t(queryNum));
  GATH_TRY(gat_memcpy_d2h(count.data(), qcount.get(), size_t(queryNum) * 4, stream), "count readback");
  // A run-indexed batch has no per-query regions (its runs place the pairs).
  if (runIndexed)
    std::fill(base.begin(), base.end(), uint64_t(0));
  else
    GATH_TRY(gat_memcpy_d2h(base.data(), qbase.get(), size_t(queryNum) * 8, stream), "region readback");
  GATH_TRY(gat_stream_sync(stream), "count readback");
  hostStale = false;
  return GAT_OK;
}

// Run-indexed batches: the runs and the used pairs, copied once.
static int download_runs(const RangeBatch& b, std::vector<gat_range_run>& runs, std::vector<int32_t>& pairs) {
  runs.resize(size_t(b.runTotal));
  pairs.resize(size_t(b.pairSpan) * 2);
  if (b.runTotal > 0)
    GATH_TRY(gat_memcpy_d2h(runs.data(), b.runs.get(), size_t(b.runTotal) * sizeof(gat_range_run), b.stream),
             "run readback");
  if (b.pairSpan > 0)
    GATH_TRY(gat_memcpy_d2h(pairs.data(), b.pairs.get(), size_t(b.pairSpan) * 8, b.stream), "pair readback");
  GATH_TRY(gat_stream_sync(b.stream), "pair readback");
  for (const gat_range_run& r : runs)
    if (r.count > 0 && (r.q >= uint32_t(b.queryNum) || r.first + r.count > uint64_t(b.pairSpan))) {
      gath_set_error("run index: a run lies outside the batch");
      return GAT_ERR_STATE;
    }
  return GAT_OK;
}

int RangeBatch::fetchRuns(gat_range_run* out, int64_t capacity) {
  if (!runIndexed) return GAT_OK;
  if (runTotal > capacity) {
    gath_set_error("fetchRuns: %lld runs, room for %lld", (long long)runTotal, (long long)capacity);
    return GAT_ERR_CAPACITY;
  }
  if (runTotal > 0)
    GATH_TRY(gat_memcpy_d2h(out, runs.get(), size_t(runTotal) * sizeof(gat_range_run), stream), "run readback");
  GATH_TRY(gat_stream_sync(stream), "run readback");
  return GAT_OK;
}

int RangeBatch::fetchCompact(int32_t* out, int64_t capacityPairs) {
  if (int rc = syncHost()) return rc;
  if (pairTotal > capacityPairs) {
    gath_set_error("fetch: %lld pairs, room for %lld", (long long)pairTotal, (long long)capacityPairs);
    return GAT_ERR_CAPACITY;
  }
  std::vector<int64_t> at(size_t(queryNum) + 1, 0);
  for (int q = 0; q < queryNum; ++q) at[size_t(q) + 1] = at[size_t(q)] + count[size_t(q)];
  if (runIndexed) {
    std::vector<gat_range_run> rs;
    std::vector<int32_t> ps;
    if (int rc = download_runs(*this, rs, ps)) return rc;
    for (const gat_range_run& r : rs) {
      if (r.count == 0) continue;
      std::memcpy(out + 2 * at[r.q], ps.data() + 2 * r.first, size_t(r.count) * 8);
      at[r.q] += r.count;
    }
    return GAT_OK;
  }
  for (int q = 0; q < queryNum; ++q) {
    if (count[size_t(q)] == 0) continue;
    GATH_TRY(gat_memcpy_d2h(out + 2 * at[size_t(q)], pairs.as<int32_t>() + 2 * base[size_t(q)],
                            size_t(count[size_t(q)]) * 8, stream),
             "pair readback");
  }
  GATH_TRY(gat_stream_sync(stream), "pair readback");
  return GAT_OK;
}

int RangeBatch::fetch(int32_t* out, int64_t capacityPairs) {
  if (int rc = syncHost()) return rc;
  // Query regions are laid out back to back by candidate count; copy each
  // query's filled prefix to its slot in `out`.
  int64_t need = queryNum ? int64_t(base.back()) + int64_t(count.back()) : 0;
  if (need > capacityPairs) {
    gath_set_error("fetch: %lld pair slots needed, %lld available", (long long)need, (long long)capacityPairs);
    return GAT_ERR_CAPACITY;
  }
  if (runIndexed) {
    std::vector<gat_range_run> rs;
    std::vector<int32_t> ps;
    if (int rc = download_runs(*this, rs, ps)) return rc;
    std::vector<uint64_t> at(base);
    for (const gat_range_run& r : rs) {
      if (r.count == 0) continue;
      std::memcpy(out + 2 * at[r.q], ps.data() + 2 * r.first, size_t(r.count) * 8);
      at[r.q] += r.count;
    }
    return GAT_OK;
  }
  for (int q = 0; q < queryNum; ++q) {
    if (count[q] == 0) continue;
    GATH_TRY(gat_memcpy_d2h(out + 2 * base[q], pairs.as<int32_t>() + 2 * base[q], size_t(count[q]) * 8, stream),
             "pair readback");
  }
  GATH_TRY(gat_stream_sync(stream), "pair readback");
  return GAT_OK;
}

int RangeBatch::digests(uint64_t* out) const {
  if (queryNum == 0) return GAT_OK;
  DeviceBuffer d;
  if (int rc = d.reserve(size_t(queryNum) * 8)) return rc;
  return digestsInto(d.as<uint64_t>(), out);
}

int RangeBatch::digestsInto(uint64_t* dev, uint64_t* out) const {
  if (queryNum == 0) return GAT_OK;
  if (runIndexed) {
    GATH_TRY(gat_runs_digest(pairs.as<int32_t>(), int64_t(pairs.size() / 8), runs.as<gat_range_run>(), runTotal,
                             queryNum, dev, stream),
             "run digest");
  } else {
    GATH_TRY(gat_pairs_digest(pairs.as<int32_t>(), qbase.as<uint64_t>(), qcount.as<uint32_t>(), queryNum,
                              int64_t(pairs.size() / 8), dev, stream),
             "pair digest");
  }
  if (out) {
    GATH_TRY(gat_memcpy_d2h(out, dev, size_t(queryNum) * 8, stream), "digest readback");
    GATH_TRY(gat_stream_sync(stream), "digest readback");
  }
  return GAT_OK;
}

// ---------------------------------------------------------------------------
// Batched form.

int DeviceTree::upload(const std::vector<gat_tree_node>& host, int nc, int d, void* stream) {
  if (host.empty() || nc < 2) {
    gath_set_error("device tree: empty tree or fewer than two children per node");
    return GAT_ERR_ARG;
  }
  // A traversal keeps at most nchild * 64 stack entries per depth level.
  if (int64_t(d + 1) * nc * 64 > GAT_TREE_STACK) {
    gath_set_error("device tree: depth %d with %d children exceeds the traversal stack", d, nc);
    return GAT_ERR_ARG;
  }
  if (int rc = nodes.reserve(host.size() * sizeof(gat_tree_node))) return rc;
  GATH_TRY(gat_memcpy_h2d(nodes.get(), host.data(), host.size() * sizeof(gat_tree_node), stream), "tree upload");
  GATH_TRY(gat_stream_sync(stream), "tree upload");
  nnodes = int64_t(host.size());
  nchild = nc;
  depth = d;
  startLevel = 0;  // a start table belongs to one tree
  ++generation;
  return GAT_OK;
}

int DeviceTree::uploadStart(const std::vector<uint32_t>& owner, int level, float x0, float y1, float block,
                            void* stream) {
  if (nnodes == 0 || nchild != 4 || level < 1 || level > 15 || owner.size() != (size_t(1) << (2 * level)) ||
      !(block > 0.0f)) {
    gath_set_error("device tree: bad start table (level %d, %zu owners)", level, owner.size());
    return GAT_ERR_ARG;
  }
  for (uint32_t o : owner)
    if ((o & ((1u << 27) - 1)) >= uint64_t(nnodes) || int(o >> 27) > level) {
      gath_set_error("device tree: start table names a node outside the tree");
      return GAT_ERR_ARG;
    }
  if (int rc = startOwner.reserve(owner.size() * 4)) return rc;
  GATH_TRY(gat_memcpy_h2d(startOwner.get(), owner.data(), owner.size() * 4, stream), "start table upload");
  GATH_TRY(gat_stream_sync(stream), "start table upload");
  startLevel = level;
  xmin = x0;
  ymax = y1;
  startBlock = block;
  ++generation;
  return GAT_OK;
}

int DeviceTree::uploadGrid(const std::vector<gat_tree_node>& cells, int gx, int gy, float x0, float y1, float cs,
                           void* stream) {
  if (gx <= 0 || gy <= 0 || cells.size() != size_t(gx) * size_t(gy) || !(cs > 0.0f)) {
    gath_set_error("device grid: %d x %d cells with %zu records", gx, gy, cells.size());
    return GAT_ERR_ARG;
  }
  if (int rc = nodes.reserve(cells.size() * sizeof(gat_tree_node))) return rc;
  GATH_TRY(gat_memcpy_h2d(nodes.get(), cells.data(), cells.size() * sizeof(gat_tree_node), stream), "grid upload");
  GATH_TRY(gat_stream_sync(stream), "grid upload");
  nnodes = int64_t(cells.size());
  nchild = 0;
  depth = 0;
  nx = gx;
  ny = gy;
  xmin = x0;
  ymax = y1;
  cellSize = cs;
  ++generation;
  return GAT_OK;
}

int DeviceTree::uploadImplicit(const std::vector<gat_tree_node>& leaves, const std::vector<uint32_t>& words,
                               int level, float x0, float y1, float cs, void* stream) {
  if (level < 1 || level > 14 || leaves.empty() || words.size() != 2 * ((size_t(1) << (2 * level)) + 1) ||
      2 * leaves.size() > words.size() || !(cs > 0.0f)) {
    gath_set_error("implicit quadtree: level %d with %zu leaves and %zu cell words", level, leaves.size(),
                   words.size());
    return GAT_ERR_ARG;
  }
  if (int64_t(level + 1) * 4 * 64 > GAT_TREE_STACK) {
    gath_set_error("implicit quadtree: depth %d exceeds the traversal stack", level);
    return GAT_ERR_ARG;
  }
  if (int rc = nodes.reserve(leaves.size() * sizeof(gat_tree_node))) return rc;
  if (int rc = cellWords.reserve(words.size() * sizeof(uint32_t))) return rc;
  GATH_TRY(gat_memcpy_h2d(nodes.get(), leaves.data(), leaves.size() * sizeof(gat_tree_node), stream),
           "implicit tree upload");
  GATH_TRY(gat_memcpy_h2d(cellWords.get(), words.data(), words.size() * sizeof(uint32_t), stream),
           "implicit tree upload");
  GATH_TRY(gat_stream_sync(stream), "implicit tree upload");
  nnodes = int64_t(leaves.size());
  nchild = 4;
  depth = level;
  implicitLevel = level;
  xmin = x0;
  ymax = y1;
  cellSize = cs;
  ++generation;
  return GAT_OK;
}

RangeEngine::~RangeEngine() { dropGraph(); }

int RangeEngine::attachStore(const RangeEngine& src) {
  if (int rc = stream_.create()) return rc;
  store_ = src.store_;
  n_ = src.n_;
  return GAT_OK;
}

int RangeEngine::reserveBatched(const DeviceTree& tree, int maxQueries, int64_t maxCandidates, RangeBatch& out) {
  if (int rc = stream_.create()) return rc;
  const int64_t N = tree.nnodes;
  const int64_t nq = std::max(maxQueries, 1);
  dropGraph();
  if (int rc = devBoxes_.reserve(boxLayout(int(nq)).bytes)) return rc;
  if (int rc = hostStage_.reserve(boxLayout(int(nq)).bytes)) return rc;
  if (int rc = nodeCount_.reserve(size_t(N) * 4)) return rc;
  if (int rc = nodeOff_.reserve(size_t(N) * 8)) return rc;
  if (int rc = tiles_.reserve(size_t(std::max(2 * gat_scan_tiles(N) * 4, gat_scan_tiles(nq) * 8)))) return rc;
  if (int rc = cand_.reserve(size_t(nq) * 8)) return rc;
  if (int rc = out.qbase.reserve(size_t(nq) * 8)) return rc;
  if (int rc = out.qcount.reserve(size_t(nq) * 4)) return rc;
  if (int rc = counters_.reserve(256)) return rc;
  if (int rc = countersHost_.reserve(256)) return rc;
  // Leaves average well over 16 points, so candidates / 16 entries is ample.
  const int64_t entries = std::max<int64_t>(maxCandidates / 16 + nq * 64, 1 << 16);
  lqCap_ = std::max(lqCap_, entries);
  qlistCap_ = std::max(qlistCap_, entries);
  itemsCap_ = std::max(itemsCap_, entries);
  pairsCap_ = std::max(pairsCap_, std::max<int64_t>(maxCandidates, 1024));
  if (int rc = lq_.reserve(size_t(lqCap_) * 8)) return rc;
  if (int rc = lqRank_.reserve(size_t(lqCap_) * 4)) return rc;
  if (int rc = qlist_.reserve(size_t(qlistCap_) * 4)) return rc;
  if (int rc = items_.reserve(size_t(itemsCap_) * 16)) return rc;
  if (int rc = directTasks_.reserve(size_t(lqCap_) * sizeof(gat_range_task))) return rc;
  if (int rc = out.runs.reserve(size_t(qlistCap_) * sizeof(gat_range_run))) return rc;
  if (int rc = itemBase_.reserve(size_t(itemsCap_) * 8)) return rc;
  if (int rc = itemTiles_.reserve(size_t(gat_scan_tiles(itemsCap_)) * 8)) return rc;
  return out.pairs.reserve(size_t(pairsCap_) * 8);
}

// Direct form: selection without per-node counts, then one refine task per
// selected (leaf, rectangle).  Boxes are already on the device.
int RangeEngine::runDirect(const DeviceTree& tree, int nq, RangeBatch& out) {
  void* s = stream_.get();
  const int64_t N = tree.nnodes;
  uint32_t* ctr = counters_.as<uint32_t>();
  const uint32_t* ch = countersHost_.as<uint32_t>();
  auto u64 = [&](int w) { return uint64_t(ch[w]) | (uint64_t(ch[w + 1]) << 32); };
  for (int attempt = 0;; ++attempt) {
    if (int rc = selTimer_.start(s)) return rc;
    GATH_TRY(gat_fill_u32(out.qcount.as<uint32_t>(), 0, nq, s), "batch reset");
    GATH_TRY(gat_fill_u32(ctr, 0, 16, s), "batch reset");
    if (int rc = enqueueTreeSelect(tree, nq, nullptr)) return rc;
    GATH_TRY(gat_scan_u64(cand_.as<uint64_t>(), nq, out.qbase.as<uint64_t>(), reinterpret_cast<uint64_t*>(ctr + 2),
                          tiles_.as<uint64_t>(), s),
             "region scan");
    if (int rc = selTimer_.stop(s)) return rc;
    GATH_TRY(gat_memcpy_d2h(countersHost_.get(), counters_.get(), 64, s), "counter readback");
    if (int rc = stream_.sync()) return rc;
    if (ch[4] & GAT_STATUS_STACK) {
      gath_set_error("batched range: traversal stack exceeded");
      return GAT_ERR_STATE;
    }
    if (u64(0) <= uint64_t(lqCap_)) break;
    if (attempt > 0 || u64(0) >= uint64_t(UINT32_MAX)) {
      gath_set_error("batched range: %llu (leaf, query) entries do not fit", (unsigned long long)u64(0));
      return GAT_ERR_CAPACITY;
    }
    lqCap_ = int64_t(u64(0) + u64(0) / 4 + 1024);
    if (int rc = lq_.reserve(size_t(lqCap_) * 8)) return rc;
  }
  float selMs = 0;
  if (int rc = selTimer_.elapsed(&selMs)) return rc;
  const int64_t ntasks = int64_t(u64(0));
  const int64_t cand = int64_t(u64(2));
  if (pairsCap_ < cand || pairsCap_ == 0) pairsCap_ = cand + cand / 4 + 1024;
  if (int rc = out.pairs.reserve(size_t(pairsCap_) * 8)) return rc;
  if (int rc = directTasks_.reserve(size_t(std::max<int64_t>(lqCap_, 1)) * sizeof(gat_range_task))) return rc;
  if (int rc = timer_.start(s)) return rc;
  GATH_TRY(gat_leaf_tasks(lq_.as<uint32_t>(), reinterpret_cast<const uint64_t*>(ctr), lqCap_,
                          tree.nodes.as<gat_tree_node>(), N, directTasks_.as<gat_range_task>(), s),
           "leaf tasks");
  GATH_TRY(gat_range_refine_pairs(&store_, directTasks_.as<gat_range_task>(), ntasks, devBoxes_.as<float>(), nq,
                                  out.qbase.as<uint64_t>(), out.qcount.as<uint32_t>(), out.pairs.as<int32_t>(),
                                  pairsCap_, s),
           "range refine");
  if (int rc = timer_.stop(s)) return rc;
  out.count.resize(size_t(nq));
  out.base.resize(size_t(nq));
  GATH_TRY(gat_memcpy_d2h(out.count.data(), out.qcount.get(), size_t(nq) * 4, s), "count readback");
  GATH_TRY(gat_memcpy_d2h(out.base.data(), out.qbase.get(), size_t(nq) * 8, s), "region readback");
  if (int rc = stream_.sync()) return rc;
  if (int rc = timer_.elapsed(&out.kernelMs)) return rc;
  out.selectMs = selMs;
  out.refineMs = out.kernelMs;
  out.hostStale = false;
  out.runIndexed = false;
  out.runTotal = 0;
  out.leafQueries = ntasks;
  out.taskNum = ntasks;
  out.candidates = cand;
  out.workItems = 0;
  out.pairTotal = 0;
  for (int q = 0; q < nq; ++q) out.pairTotal += out.count[size_t(q)];
  return GAT_OK;
}

RangeEngine::GraphKey RangeEngine::layoutKey(const DeviceTree& tree, int nq, const RangeBatch& out) const {
  GraphKey key;
  const void* p[GraphKey::kPtrs] = {tree.nodes.get(), tree.cellWords.get(), devBoxes_.get(), out.pairs.get(),
                                    out.qbase.get(),  out.qcount.get(),     out.runs.get(),  lq_.get(),
                                    nodeCount_.get(), lqRank_.get(),        nodeOff_.get(),  tiles_.get(),
                                    cand_.get(),      qlist_.get(),         items_.get(),    itemBase_.get(),
                                    itemTiles_.get(), counters_.get(),      store_.xy_dev,   store_.id_dev};
  std::copy(p, p + GraphKey::kPtrs, key.ptr);
  key.runs = runIndexed;
  key.treeGeneration = tree.generation;
  key.nq = nq;
  key.lqCap = lqCap_;
  key.qlistCap = qlistCap_;
  key.itemsCap = itemsCap_;
  key.pairsCap = pairsCap_;
  key.runsCap = int64_t(out.runs.size() / sizeof(gat_range_run));
  key.qb = queriesPerItem;
  return key;
}

void RangeEngine::dropGraph() {
  if (graph_) gat_graph_destroy(graph_);
  graph_ = nullptr;
  graphKey_ = GraphKey{};
}

// Layout of the batch's rectangle buffer: the boxes, the order the selection
// takes them in, and the walk tasks of the largest rectangles of a tree with a
// start table (their count and the order entries they cover, then the tasks).
RangeEngine::BoxLayout RangeEngine::boxLayout(int nq) {
  BoxLayout l;
  l.tasksCap = 1024 + size_t(nq) / 2;
  l.order = align_up(size_t(nq) * 16, 256);
  l.ntasks = align_up(l.order + size_t(nq) * 4, 256);
  l.tasks = l.ntasks + 8;
  l.bytes = l.tasks + l.tasksCap * 8;
  return l;
}

// Block of a coordinate offset in blocks, moved by d and clamped: the float
// steps the device's tree walk takes.
static uint32_t start_block(float v, int d, uint32_t nb) {
  const float t = std::floor(v) + float(d);
  return uint32_t(std::min(std::max(t, 0.0f), float(nb - 1)));
}

static float clamped_area(const MBB& b) {
  return std::max(0.0f, b.xmax - b.xmin) * std::max(0.0f, b.ymax - b.ymin);
}

// Uploads the rectangles and the order the selection takes them in: largest
// area first, so the longest traversals start first and the tail is short.
// With a start table, each of the largest rectangles whose candidate blocks
// span several 64-block chunks gets one walk task per chunk (while the task
// list has room), so its walk runs on several waves; its last task also takes
// any blocks past its chunks.  Rectangles under four blocks of area are not
// looked at (a thin one among them keeps one task: slower, never wrong).
int RangeEngine::uploadBoxes(const DeviceTree* tree, const MBB* bounds, int nq) {
  static_assert(sizeof(MBB) == 16, "MBB must be four packed floats");
  const BoxLayout lay = boxLayout(nq);
  const size_t bytes = lay.bytes;
  if (int rc = hostStage_.reserve(std::max<size_t>(bytes, 256))) return rc;
  if (devBoxes_.size() < bytes) dropGraph();  // the captured pipeline reads the old buffer
  if (int rc = devBoxes_.reserve(std::max<size_t>(bytes, 256))) return rc;
  char* h = hostStage_.as<char>();
  std::memcpy(h, bounds, size_t(nq) * 16);
  // Counting sort on the binary exponent of the area, largest first: an order
  // within a factor of two is all the scheduling needs, in O(n).
  auto* order = reinterpret_cast<uint32_t*>(h + lay.order);
  std::vector<uint8_t> key(static_cast<size_t>(nq));
  uint32_t start[257] = {};
  for (int q = 0; q < nq; ++q) {
    const float area = clamped_area(bounds[q]);
    uint32_t bits;
    std::memcpy(&bits, &area, 4);
    key[q] = uint8_t(255 - std::min<uint32_t>(bits >> 23, 255));  // NaN areas are 0: last
    ++start[key[q] + 1];
  }
  for (int b = 0; b < 256; ++b) start[b + 1] += start[b];
  for (int q = 0; q < nq; ++q) order[start[key[q]]++] = uint32_t(q);
  size_t upload = lay.order + size_t(nq) * 4;
  if (tree && tree->startLevel > 0) {
    auto* counts = reinterpret_cast<uint32_t*>(h + lay.ntasks);
    auto* tasks = reinterpret_cast<uint32_t*>(h + lay.tasks);
    const uint32_t nb = 1u << tree->startLevel;
    const float inv = 1.0f / tree->startBlock;
    const float small = 4.0f * tree->startBlock * tree->startBlock;
    size_t n = 0;
    int i = 0;
    for (; i < nq; ++i) {
      const uint32_t q = order[i];
      const MBB& b = bounds[q];
      if (clamped_area(b) < small) break;  // order is by area: every later one is under 8 blocks
      size_t nch = 1;
      if (b.xmin <= b.xmax && b.ymin <= b.ymax) {
        const uint32_t c0 = start_block((b.xmin - tree->xmin) * inv, -1, nb);
        const uint32_t c1 = start_block((b.xmax - tree->xmin) * inv, 1, nb);
        const uint32_t r0 = start_block((tree->ymax - b.ymax) * inv, -1, nb);
        const uint32_t r1 = start_block((tree->ymax - b.ymin) * inv, 1, nb);
        const size_t ncand = size_t(c1 - c0 + 1) * size_t(r1 - r0 + 1);
        nch = (ncand + 63) / 64;
      }
      if (n + nch > lay.tasksCap) break;
      for (size_t c = 0; c < nch; ++c, ++n) {
        tasks[2 * n] = q;
        tasks[2 * n + 1] = uint32_t(c) | (c + 1 == nch ? 0x80000000u : 0u);
      }
    }
    counts[0] = uint32_t(n);  // tasks
    counts[1] = uint32_t(i);  // the order entries they cover
    upload = lay.tasks + n * 8;
  }
  GATH_TRY(gat_memcpy_h2d(devBoxes_.get(), h, upload, stream_.get()), "box upload");
  return GAT_OK;
}

int RangeEngine::enqueueTreeSelect(const DeviceTree& tree, int nq, uint32_t* nodeCount, bool candCleared) {
  void* s = stream_.get();
  uint32_t* ctr = counters_.as<uint32_t>();
  auto* total = reinterpret_cast<uint64_t*>(ctr);
  const auto* order =
      reinterpret_cast<const uint32_t*>(devBoxes_.as<char>() + align_up(size_t(nq) * 16, 256));
  if (tree.flat()) {
    GATH_TRY(gat_grid_select(tree.nodes.as<gat_tree_node>(), tree.nx, tree.ny, tree.xmin, tree.ymax, tree.cellSize,
                             devBoxes_.as<float>(), order, 0, nq, nodeCount, cand_.as<uint64_t>(),
                             lq_.as<uint32_t>(), nodeCount ? lqRank_.as<uint32_t>() : nullptr, lqCap_, total, s),
             "grid select");
  } else if (tree.implicit()) {
    GATH_TRY(gat_morton_select(tree.cellWords.as<uint32_t>(), tree.implicitLevel, tree.nnodes, tree.xmin, tree.ymax,
                               tree.cellSize, devBoxes_.as<float>(), order, 0, nq, nodeCount, cand_.as<uint64_t>(),
                               lq_.as<uint32_t>(), nodeCount ? lqRank_.as<uint32_t>() : nullptr, lqCap_, total,
                               ctr + 4, s),
             "implicit tree select");
  } else {
    const BoxLayout lay = boxLayout(nq);
    char* d = devBoxes_.as<char>();
    const gat_tree_start start{tree.startOwner.as<uint32_t>(),
                               tree.startLevel,
                               tree.xmin,
                               tree.ymax,
                               tree.startBlock,
                               reinterpret_cast<const uint32_t*>(d + lay.tasks),
                               reinterpret_cast<const uint32_t*>(d + lay.ntasks),
                               int64_t(lay.tasksCap)};
    // Several tasks add to one rectangle's candidate count.
    if (tree.startLevel > 0 && !candCleared)
      GATH_TRY(gat_fill_u32(cand_.as<uint32_t>(), 0, int64_t(nq) * 2, s), "batch reset");
    GATH_TRY(gat_tree_select(tree.nodes.as<gat_tree_node>(), tree.nnodes, tree.nchild,
                             tree.startLevel > 0 ? &start : nullptr, devBoxes_.as<float>(), order,
                             0, nq, nodeCount, cand_.as<uint64_t>(), lq_.as<uint32_t>(),
                             nodeCount ? lqRank_.as<uint32_t>() : nullptr, lqCap_, total, ctr + 4, s),
             "tree select");
  }
  return GAT_OK;
}

int RangeEngine::enqueueSelect(const DeviceTree& tree, int nq, RangeBatch& out) {
  void* s = stream_.get();
  const int64_t N = tree.nnodes;
  uint32_t* ctr = counters_.as<uint32_t>();
  // The batch's resets in one launch: per-node counts, per-query counts, the
  // counters and, for the start-table walk, the candidate counts.
  const bool clearCand = selectAddsCandidates(tree);
  uint32_t* const spans[4] = {nodeCount_.as<uint32_t>(), out.qcount.as<uint32_t>(), ctr, cand_.as<uint32_t>()};
  const int64_t lens[4] = {N, nq, 16, clearCand ? int64_t(nq) * 2 : 0};
  GATH_TRY(gat_fill_u32_spans(4, spans, lens, 0, s), "batch reset");
  if (int rc = enqueueTreeSelect(tree, nq, nodeCount_.as<uint32_t>(), clearCand)) return rc;
  // The run-indexed join places its output by work item: it needs the
  // candidate total but no per-query regions.  (The walk adding the total
  // with one atomic per wave cost it 20 us: every wave on one address.)
  GATH_TRY(gat_scan_u64(cand_.as<uint64_t>(), nq, runIndexed ? nullptr : out.qbase.as<uint64_t>(),
                        reinterpret_cast<uint64_t*>(ctr + 2), tiles_.as<uint64_t>(), s),
           "region scan");
  GATH_TRY(gat_scan_leaf_lists(nodeCount_.as<uint32_t>(), N, queriesPerItem, nodeOff_.as<uint32_t>(), ctr + 5,
                               tiles_.as<uint32_t>(), s),
           "leaf list scan");
  return GAT_OK;
}

int RangeEngine::enqueueJoin(const DeviceTree& tree, int nq, RangeBatch& out) {
  void* s = stream_.get();
  const int64_t N = tree.nnodes;
  uint32_t* ctr = counters_.as<uint32_t>();
  GATH_TRY(gat_leaf_scatter(lq_.as<uint32_t>(), reinterpret_cast<const uint64_t*>(ctr), lqCap_,
                            tree.nodes.as<gat_tree_node>(), N, nodeOff_.as<uint32_t>(), nodeCount_.as<uint32_t>(),
                            lqRank_.as<uint32_t>(), queriesPerItem, nq, qlist_.as<uint32_t>(), qlistCap_,
                            items_.as<uint32_t>(), itemsCap_, s),
           "leaf scatter");
  if (runIndexed) {
    GATH_TRY(gat_leaf_item_regions(items_.as<uint32_t>(), ctr + 6, itemsCap_, itemBase_.as<uint64_t>(),
                                   itemTiles_.as<uint64_t>(), reinterpret_cast<uint64_t*>(ctr + 10), s),
             "item regions");
    if (timeRefine_)
      if (int rc = refineTimer_.start(s)) return rc;
    GATH_TRY(gat_leaf_refine_runs(&store_, items_.as<uint32_t>(), ctr + 6, itemsCap_, qlist_.as<uint32_t>(),
                                  qlistCap_, devBoxes_.as<float>(), nq, out.pairs.as<int32_t>(), pairsCap_,
                                  out.runs.as<gat_range_run>(), int64_t(out.runs.size() / sizeof(gat_range_run)),
                                  itemBase_.as<uint64_t>(), s),
             "leaf refine");
    if (timeRefine_)
      if (int rc = refineTimer_.stop(s)) return rc;
    GATH_TRY(gat_run_counts(&store_, items_.as<uint32_t>(), ctr + 6, itemsCap_, qlistCap_,
                            out.runs.as<gat_range_run>(), int64_t(out.runs.size() / sizeof(gat_range_run)), nq,
                            out.qcount.as<uint32_t>(), reinterpret_cast<uint64_t*>(ctr + 8), s),
             "run counts");
  } else {
    if (timeRefine_)
      if (int rc = refineTimer_.start(s)) return rc;
    GATH_TRY(gat_leaf_refine(&store_, items_.as<uint32_t>(), ctr + 6, itemsCap_, qlist_.as<uint32_t>(), qlistCap_,
                             devBoxes_.as<float>(), out.qbase.as<uint64_t>(), nq, out.qcount.as<uint32_t>(),
                             out.pairs.as<int32_t>(), pairsCap_, reinterpret_cast<uint64_t*>(ctr + 8), s),
             "leaf refine");
    if (timeRefine_)
      if (int rc = refineTimer_.stop(s)) return rc;
  }
  out.runIndexed = runIndexed;
  return GAT_OK;
}

int RangeEngine::uniqueCandidates(const DeviceTree& tree, int64_t* points) {
  const int64_t N = tree.nnodes;
  if (N <= 0 || nodeCount_.size() < size_t(N) * 4) {
    gath_set_error("uniqueCandidates: no batched run on this tree");
    return GAT_ERR_STATE;
  }
  std::vector<uint32_t> cnt(static_cast<size_t>(N));
  std::vector<gat_tree_node> nodes(static_cast<size_t>(N));
  GATH_TRY(gat_memcpy_d2h(cnt.data(), nodeCount_.get(), size_t(N) * 4, stream_.get()), "leaf count readback");
  GATH_TRY(gat_memcpy_d2h(nodes.data(), tree.nodes.get(), size_t(N) * sizeof(gat_tree_node), stream_.get()),
           "tree readback");
  if (int rc = stream_.sync()) return rc;
  int64_t u = 0;
  for (int64_t i = 0; i < N; ++i)
    if (cnt[size_t(i)] > 0 && nodes[size_t(i)].first_child < 0) u += nodes[size_t(i)].count;
  *points = u;
  return GAT_OK;
}

// Copies the batch counters back; per-query counts and regions stay on the
// device until RangeBatch::syncHost.
int RangeEngine::readCounters(int nq, RangeBatch& out) {
  GATH_TRY(gat_memcpy_d2h(countersHost_.get(), counters_.get(), 64, stream_.get()), "counter readback");
  out.hostStale = nq > 0;
  return stream_.sync();
}

bool RangeEngine::countersFit() const {
  const uint32_t* ch = countersHost_.as<uint32_t>();
  auto u64 = [&](int w) { return uint64_t(ch[w]) | (uint64_t(ch[w + 1]) << 32); };
  return u64(0) <= uint64_t(lqCap_) && int64_t(ch[5]) <= qlistCap_ && int64_t(ch[6]) <= itemsCap_ &&
         u64(2) <= uint64_t(pairsCap_) && u64(10) <= uint64_t(pairsCap_);
}

int RangeEngine::finishCounters(RangeBatch& out, bool replay) {
  const uint32_t* ch = countersHost_.as<uint32_t>();
  auto u64 = [&](int w) { return uint64_t(ch[w]) | (uint64_t(ch[w + 1]) << 32); };
  if (ch[4] & GAT_STATUS_STACK) {
    gath_set_error("batched range: traversal stack exceeded");
    return GAT_ERR_STATE;
  }
  out.replayed = replay;
  out.leafQueries = int64_t(u64(0));
  out.candidates = int64_t(u64(2));
  out.workItems = int64_t(ch[6]);
  out.pairTotal = int64_t(u64(8));
  out.runTotal = out.runIndexed ? int64_t(ch[5]) : 0;
  out.pairSpan = out.runIndexed ? int64_t(u64(10)) : 0;
  return GAT_OK;
}

int RangeEngine::stageBatch(const DeviceTree& tree, const MBB* bounds, int nq) {
  if (nq < 1 || !bounds || devBoxes_.size() < boxLayout(nq).bytes) {
    gath_set_error("stageBatch: the workspace was not sized for %d rectangles", nq);
    return GAT_ERR_STATE;
  }
  return uploadBoxes(&tree, bounds, nq);
}

int RangeEngine::enqueueBatch(const DeviceTree& tree, int nq, RangeBatch& out) {
  if (int rc = enqueueSelect(tree, nq, out)) return rc;
  return enqueueJoin(tree, nq, out);
}

int RangeEngine::finishBatch(int nq, RangeBatch& out, bool* fits) {
  out.queryNum = nq;
  out.count.assign(size_t(nq), 0);
  out.base.assign(size_t(nq), 0);
  out.stream = stream_.get();
  out.taskNum = 0;
  out.selectMs = out.uploadMs = 0;
  out.kernelMs = out.refineMs = 0;
  if (int rc = readCounters(nq, out)) return rc;
  *fits = countersFit();
  return *fits ? finishCounters(out, true) : GAT_OK;
}

int RangeEngine::runTree(const DeviceTree& tree, const MBB* bounds, int nq, RangeBatch& out, bool graph) {
  if (int rc = stream_.create()) return rc;
  if (tree.nnodes <= 0 || nq < 0 || (nq > 0 && !bounds) || queriesPerItem < 1 || queriesPerItem > 64) {
    gath_set_error("batched range: bad arguments");
    return GAT_ERR_ARG;
  }
  out.queryNum = nq;
  out.count.assign(size_t(nq), 0);
  out.base.assign(size_t(nq), 0);
  out.hostStale = false;
  out.stream = stream_.get();
  out.candidates = out.taskNum = out.pairTotal = out.leafQueries = out.workItems = 0;
  out.runTotal = out.pairSpan = 0;
  out.runIndexed = false;
  out.replayed = false;
  out.selectMs = out.uploadMs = 0;
  out.kernelMs = out.refineMs = 0;
  if (nq == 0) return GAT_OK;
  const double t0 = now_ms();
  if (int rc = uploadBoxes(&tree, bounds, nq)) return rc;
  const int64_t N = tree.nnodes;
  if (nq <= directMaxQueries) {
    // Direct form: workspace sized by the batch only.
    if (int rc = tiles_.reserve(size_t(gat_scan_tiles(nq)) * 8)) return rc;
    if (int rc = cand_.reserve(size_t(nq) * 8)) return rc;
    if (int rc = out.qbase.reserve(size_t(nq) * 8)) return rc;
    if (int rc = out.qcount.reserve(size_t(nq) * 4)) return rc;
    if (int rc = counters_.reserve(256)) return rc;
    if (int rc = countersHost_.reserve(256)) return rc;
    if (lqCap_ < int64_t(nq) * 16) {
      lqCap_ = std::max<int64_t>(int64_t(nq) * 16, 1 << 16);
      if (int rc = lq_.reserve(size_t(lqCap_) * 8)) return rc;
    }
    const int rc = runDirect(tree, nq, out);
    out.uploadMs = now_ms() - t0;
    return rc;
  }
  out.runIndexed = runIndexed;
  const size_t tileBytes = size_t(std::max(2 * gat_scan_tiles(N) * 4, gat_scan_tiles(nq) * 8));
  if (nodeCount_.size() < size_t(N) * 4 || tiles_.size() < tileBytes || cand_.size() < size_t(nq) * 8 ||
      out.qbase.size() < size_t(nq) * 8 || out.qcount.size() < size_t(nq) * 4)
    dropGraph();
  if (int rc = nodeCount_.reserve(size_t(N) * 4)) return rc;
  if (int rc = nodeOff_.reserve(size_t(N) * 8)) return rc;
  if (int rc = tiles_.reserve(tileBytes)) return rc;
  if (int rc = cand_.reserve(size_t(nq) * 8)) return rc;
  if (int rc = out.qbase.reserve(size_t(nq) * 8)) return rc;
  if (int rc = out.qcount.reserve(size_t(nq) * 4)) return rc;
  if (int rc = counters_.reserve(256)) return rc;
  if (int rc = countersHost_.reserve(256)) return rc;
  const uint32_t* ch = countersHost_.as<uint32_t>();
  auto u64 = [&](int w) { return uint64_t(ch[w]) | (uint64_t(ch[w + 1]) << 32); };
  auto fits = [&] { return countersFit(); };
  auto finish = [&](bool replay) -> int { return finishCounters(out, replay); };
  GraphKey key;
  auto make_key = [&] { key = layoutKey(tree, nq, out); };
  if (graph && graph_) {
    make_key();
    if (key == graphKey_) {
      if (int rc = timer_.start(stream_.get())) return rc;
      GATH_TRY(gat_graph_launch(graph_, stream_.get()), "graph launch");
      if (int rc = timer_.stop(stream_.get())) return rc;
      if (int rc = readCounters(nq, out)) return rc;
      if (int rc = timer_.elapsed(&out.kernelMs)) return rc;
      if (fits()) {
        out.uploadMs = now_ms() - t0;
        return finish(true);
      }
      // The batch outgrew the captured capacities: run it step by step below.
    }
  }
  // Selection, then capacities from its totals, then the join.
  if (lqCap_ < int64_t(nq) * 16) lqCap_ = std::max<int64_t>(int64_t(nq) * 16, 1 << 16);
  if (int rc = lq_.reserve(size_t(lqCap_) * 8)) return rc;
  if (int rc = lqRank_.reserve(size_t(lqCap_) * 4)) return rc;
  for (int attempt = 0;; ++attempt) {
    if (int rc = selTimer_.start(stream_.get())) return rc;
    if (int rc = enqueueSelect(tree, nq, out)) return rc;
    if (int rc = selTimer_.stop(stream_.get())) return rc;
    GATH_TRY(gat_memcpy_d2h(countersHost_.get(), counters_.get(), 64, stream_.get()), "counter readback");
    if (int rc = stream_.sync()) return rc;
    if (u64(0) <= uint64_t(lqCap_)) break;
    if (attempt > 0 || u64(0) >= uint64_t(UINT32_MAX)) {
      gath_set_error("batched range: %llu (leaf, query) entries do not fit", (unsigned long long)u64(0));
      return GAT_ERR_CAPACITY;
    }
    lqCap_ = int64_t(u64(0) + u64(0) / 4 + 1024);
    if (int rc = lq_.reserve(size_t(lqCap_) * 8)) return rc;
    if (int rc = lqRank_.reserve(size_t(lqCap_) * 4)) return rc;
  }
  float selMs = 0;
  if (int rc = selTimer_.elapsed(&selMs)) return rc;
  out.selectMs = selMs;
  if (ch[4] & GAT_STATUS_STACK) return finish(false);
  // Room with a quarter to spare, so a captured pipeline serves similar batches.
  auto grow = [](int64_t need, int64_t& cap) {
    if (need > cap || cap == 0) cap = need + need / 4 + 1024;
  };
  grow(int64_t(ch[5]), qlistCap_);
  grow(int64_t(ch[6]), itemsCap_);
  // Run-indexed regions round each work item up to a line: 15 pairs each at most.
  grow(int64_t(u64(2)) + (runIndexed ? 16 * int64_t(ch[6]) : 0), pairsCap_);
  if (int rc = qlist_.reserve(size_t(qlistCap_) * 4)) return rc;
  if (int rc = items_.reserve(size_t(itemsCap_) * 16)) return rc;
  if (int rc = out.pairs.reserve(size_t(pairsCap_) * 8)) return rc;
  if (int rc = out.runs.reserve(size_t(qlistCap_) * sizeof(gat_range_run))) return rc;
  if (int rc = itemBase_.reserve(size_t(itemsCap_) * 8)) return rc;
  if (int rc = itemTiles_.reserve(size_t(gat_scan_tiles(itemsCap_)) * 8)) return rc;
  if (int rc = timer_.start(stream_.get())) return rc;
  timeRefine_ = true;
  const int rc_join = enqueueJoin(tree, nq, out);
  timeRefine_ = false;
  if (rc_join) return rc_join;
  if (int rc = timer_.stop(stream_.get())) return rc;
  if (int rc = readCounters(nq, out)) return rc;
  if (int rc = timer_.elapsed(&out.kernelMs)) return rc;
  if (int rc = refineTimer_.elapsed(&out.refineMs)) return rc;
  out.uploadMs = now_ms() - t0;
  if (int rc = finish(false)) return rc;
  if (graph) {
    make_key();
    if (!(key == graphKey_) || !graph_) {
      dropGraph();
      void* g = nullptr;
      GATH_TRY(gat_graph_begin(stream_.get()), "graph capture");
      int rc = enqueueSelect(tree, nq, out);
      if (rc == GAT_OK) rc = enqueueJoin(tree, nq, out);
      const int rc_end = gat_graph_end(stream_.get(), &g);
      if (rc != GAT_OK) return rc;
      GATH_TRY(rc_end, "graph capture");
      graph_ = g;
      graphKey_ = key;
    }
  }
  return GAT_OK;
}
