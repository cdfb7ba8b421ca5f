#include "RangeEngine.h"

#include <algorithm>
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

int RangeBatch::fetch(int32_t* out, int64_t capacityPairs) const {
  // Query regions are laid out back to back by candidate count; copy each
  // query's filled prefix to its slot in `out`.
  int64_t need = queryNum ? int64_t(base.back()) + int64_t(count.back()) : 0;
  if (need > capacityPairs) {
    gath_set_error("fetch: %lld pair slots needed, %lld available", (long long)need, (long long)capacityPairs);
    return GAT_ERR_CAPACITY;
  }
  for (int q = 0; q < queryNum; ++q) {
    if (count[q] == 0) continue;
    GATH_TRY(gat_memcpy_d2h(out + 2 * base[q], pairs.as<int32_t>() + 2 * base[q], size_t(count[q]) * 8, nullptr),
             "pair readback");
  }
  GATH_TRY(gat_device_sync(), "pair readback");
  return GAT_OK;
}

int RangeBatch::digests(uint64_t* out) const {
  if (queryNum == 0) return GAT_OK;
  DeviceBuffer d;
  if (int rc = d.reserve(size_t(queryNum) * 8)) return rc;
  GATH_TRY(gat_pairs_digest(pairs.as<int32_t>(), qbase.as<uint64_t>(), qcount.as<uint32_t>(), queryNum,
                            int64_t(pairs.size() / 8), d.as<uint64_t>(), nullptr),
           "pair digest");
  GATH_TRY(gat_memcpy_d2h(out, d.get(), size_t(queryNum) * 8, nullptr), "digest readback");
  GATH_TRY(gat_device_sync(), "digest readback");
  return GAT_OK;
}
