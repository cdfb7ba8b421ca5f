#include "Server.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

namespace {
using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

// Everything the generator and the slots share.
struct Shared {
  std::mutex m;
  std::condition_variable cv;
  std::deque<int> range, topk;  // waiting request indices, in arrival order
  bool arrivalsDone = false;
  bool failed = false;
  int rc = GAT_OK;
  std::string error;

  void fail(int code) {
    std::lock_guard<std::mutex> lk(m);
    if (!failed) {
      failed = true;
      rc = code;
      error = gath_last_error();
    }
    cv.notify_all();
  }
};
}  // namespace

double Server::rangeBytes(const float* b) const {
  int64_t points = 0, leaves = 0;
  grid_.selectionCost(MBB(b[0], b[1], b[2], b[3]), &points, &leaves);
  return kRangeBytesPerPair * double(rangePairBound(points, leaves));
}

int64_t Server::rangePairBound(int64_t points, int64_t leaves) {
  // Result pairs are at most the candidate points; each selected leaf's work
  // item rounds its pair region up to whole lines of 16 pairs.
  return points + 16 * leaves;
}

double Server::topkBytes(int qlen, double survivors) const {
  // Query points (8 B each) and box, k result ids and distances, and the
  // filter survivors: a (query, trajectory) pair and its distance (12 B each).
  return 8.0 * qlen + 16.0 + 8.0 * cfg_.k + 12.0 * survivors;
}

int Server::run(const gath_request* reqs, int n, const float* qlon, const float* qlat, int64_t nqpoints,
                gath_outcome* out, int32_t* topkIds, int32_t* topkDists, ServeStats* stats) {
  if (n < 0 || (n > 0 && (!reqs || !out)) || cfg_.k < 1 || cfg_.rangeSlots < 1 || cfg_.maxRangeBatch < 1 ||
      cfg_.maxTopkBatch < 1 || cfg_.maxTopkBatch > 65535) {
    gath_set_error("serve: bad arguments");
    return GAT_ERR_ARG;
  }
  bool anyTopk = false;
  for (int i = 0; i < n; ++i) {
    const gath_request& r = reqs[i];
    if (r.kind != GATH_REQ_RANGE && r.kind != GATH_REQ_TOPK) {
      gath_set_error("serve: request %d has kind %d", i, r.kind);
      return GAT_ERR_ARG;
    }
    if (i > 0 && r.arrival_ms < reqs[i - 1].arrival_ms) {
      gath_set_error("serve: arrivals must be in time order (request %d)", i);
      return GAT_ERR_ARG;
    }
    if (r.kind == GATH_REQ_TOPK) {
      anyTopk = true;
      if (r.qlen < 1 || r.qlen > GAT_MAX_TRAJ_LEN || r.qoff < 0 || !qlon || !qlat || !topkIds || !topkDists) {
        gath_set_error("serve: bad top-k request %d", i);
        return GAT_ERR_ARG;
      }
      if (r.qoff > nqpoints - r.qlen) {
        gath_set_error("serve: top-k request %d reads query points [%lld, %lld) of %lld", i, (long long)r.qoff,
                       (long long)r.qoff + r.qlen, (long long)nqpoints);
        return GAT_ERR_ARG;
      }
    }
  }
  ServeStats st;
  // Each range request's buffer need, from the leaves it selects (the device
  // selection applies the same rule): computed as it arrives.
  std::vector<int64_t> rangeNeed(size_t(n), 0);
  // Admission is bounded by the device memory that is actually free: the
  // budgets of every slot together stay within 80% of it.
  size_t freeB = 0, totalB = 0;
  GATH_TRY(gat_mem_info(&freeB, &totalB), "serve: memory query");
  const double need = cfg_.batched ? cfg_.rangeSlots * cfg_.rangeBudgetBytes + cfg_.topkBudgetBytes : 0.0;
  const double scale = need > 0.8 * double(freeB) ? 0.8 * double(freeB) / need : 1.0;
  st.rangeBudgetBytes = cfg_.rangeBudgetBytes * scale;
  st.topkBudgetBytes = cfg_.topkBudgetBytes * scale;
  grid_.epsilon = cfg_.eps;
  if (anyTopk && grid_.mismatchTable().empty()) return GAT_ERR_HIP;

  Shared sh;
  Clock::time_point t0 = Clock::now();
  auto finish_topk = [&](const std::vector<int>& idx, const std::vector<TopK>& res) {
    for (size_t j = 0; j < idx.size(); ++j) {
      const size_t at = size_t(idx[j]) * size_t(cfg_.k);
      for (int c = 0; c < cfg_.k; ++c) {
        const bool has = size_t(c) < res[j].ids.size();
        topkIds[at + c] = has ? res[j].ids[size_t(c)] : 0;
        topkDists[at + c] = has ? res[j].dist[size_t(c)] : -1;
      }
    }
  };
  // The query points of a set of top-k requests as a trajectory view.
  auto gather_queries = [&](const std::vector<int>& idx, TrajectoryDB& db) {
    db.offsets.assign(1, 0);
    db.lon.clear();
    db.lat.clear();
    for (int i : idx) {
      const gath_request& r = reqs[i];
      db.lon.insert(db.lon.end(), qlon + r.qoff, qlon + r.qoff + r.qlen);
      db.lat.insert(db.lat.end(), qlat + r.qoff, qlat + r.qoff + r.qlen);
      db.offsets.push_back(int64_t(db.lon.size()));
    }
  };

  if (!cfg_.batched) {
    // Plain per-request path: one request at a time in arrival order, range
    // through the plain drop-in (host selection, task refine), top-k through
    // one EDR launch per query over every trajectory.
    RangeBatch rb;
    TrajectoryDB qdb;
    DeviceBuffer dig;
    if (int rc = dig.reserve(8)) return rc;
    for (int i = 0; i < n; ++i) {
      const gath_request& r = reqs[i];
      const double wait = r.arrival_ms - ms_since(t0);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
      gath_outcome& o = out[i];
      o.admit_ms = ms_since(t0);
      if (r.kind == GATH_REQ_RANGE) {
        const MBB box(r.box[0], r.box[1], r.box[2], r.box[3]);
        if (int rc = grid_.rangeQueryBatchGPUPairs(&box, 1, rb)) return rc;
        if (int rc = rb.syncHost()) return rc;
        uint64_t d = 0;
        if (int rc = rb.digestsInto(dig.as<uint64_t>(), &d)) return rc;
        o.count = rb.count[0];
        o.digest = d;
        o.batch = int32_t(st.rangeBatches++);
        ++st.rangeDone;
      } else {
        gather_queries({i}, qdb);
        std::vector<TopK> res;
        if (int rc = grid_.similarityPlain(TrajectoryView::of(qdb), cfg_.k, res)) return rc;
        finish_topk({i}, res);
        o.batch = int32_t(st.topkBatches++);
        ++st.topkDone;
      }
      o.done_ms = ms_since(t0);
    }
    st.wallMs = ms_since(t0);
    if (stats) *stats = st;
    return GAT_OK;
  }

  // Range slots: engines over the grid's resident store, each with its own
  // stream and a workspace sized up front for its budget, so serving does no
  // device allocation (a free would synchronise the whole device).
  struct RangeSlot {
    RangeEngine engine;
    RangeBatch batch;
    DeviceBuffer digests;
    std::vector<uint64_t> digestsHost;
    std::vector<MBB> boxes;
  };
  std::vector<std::unique_ptr<RangeSlot>> slots;
  // A slot's workspace holds maxCand result pairs and maxCand / 16 list
  // entries; a batch is admitted while its requests' pair bounds fit.
  const int64_t maxCand = std::max<int64_t>(int64_t(st.rangeBudgetBytes / kRangeBytesPerPair), 1);
  for (int s = 0; s < cfg_.rangeSlots; ++s) {
    auto slot = std::make_unique<RangeSlot>();
    if (int rc = slot->engine.attachStore(grid_.engine())) return rc;
    if (int rc = slot->engine.reserveBatched(grid_.deviceTree(), cfg_.maxRangeBatch, maxCand, slot->batch))
      return rc;
    if (int rc = slot->digests.reserve(size_t(cfg_.maxRangeBatch) * 8)) return rc;
    slots.push_back(std::move(slot));
  }

  auto range_worker = [&](RangeSlot* slot) {
    for (;;) {
      std::vector<int> take;
      double est = 0;
      int32_t batchNo = 0;
      {
        std::unique_lock<std::mutex> lk(sh.m);
        sh.cv.wait(lk, [&] { return sh.failed || !sh.range.empty() || sh.arrivalsDone; });
        if (sh.failed || (sh.range.empty() && sh.arrivalsDone)) return;
        // Every waiting rectangle that fits the budget (at least one).
        while (!sh.range.empty() && int(take.size()) < cfg_.maxRangeBatch) {
          const double e = kRangeBytesPerPair * double(rangeNeed[size_t(sh.range.front())]);
          if (!take.empty() && est + e > st.rangeBudgetBytes) break;
          est += e;
          take.push_back(sh.range.front());
          sh.range.pop_front();
        }
        batchNo = int32_t(st.rangeBatches++);
        st.peakRangeBytes = std::max(st.peakRangeBytes, est);
      }
      const double admit = ms_since(t0);
      slot->boxes.clear();
      for (int i : take) slot->boxes.emplace_back(reqs[i].box[0], reqs[i].box[1], reqs[i].box[2], reqs[i].box[3]);
      const int nq = int(take.size());
      int rc = slot->engine.runTree(grid_.deviceTree(), slot->boxes.data(), nq, slot->batch, false);
      if (rc == GAT_OK) rc = slot->batch.syncHost();
      slot->digestsHost.resize(size_t(nq));
      if (rc == GAT_OK) rc = slot->batch.digestsInto(slot->digests.as<uint64_t>(), slot->digestsHost.data());
      if (rc != GAT_OK) {
        sh.fail(rc);
        return;
      }
      const double done = ms_since(t0);
      for (int j = 0; j < nq; ++j) {
        gath_outcome& o = out[take[size_t(j)]];
        o.admit_ms = admit;
        o.done_ms = done;
        o.batch = batchNo;
        o.count = slot->batch.count[size_t(j)];
        o.digest = slot->digestsHost[size_t(j)];
      }
      std::lock_guard<std::mutex> lk(sh.m);
      st.rangeDone += nq;
    }
  };

  // Filter survivors per top-k query, as observed: seeded by the warm-up
  // batch below, then updated after every batch.
  double survivorsPerQuery = 0;
  auto topk_worker = [&]() {
    TrajectoryDB qdb;
    std::vector<TopK> res;
    for (;;) {
      std::vector<int> take;
      double est = 0;
      int32_t batchNo = 0;
      {
        std::unique_lock<std::mutex> lk(sh.m);
        sh.cv.wait(lk, [&] { return sh.failed || !sh.topk.empty() || sh.arrivalsDone; });
        if (sh.failed || (sh.topk.empty() && sh.arrivalsDone)) return;
        while (!sh.topk.empty() && int(take.size()) < cfg_.maxTopkBatch) {
          const double e = topkBytes(reqs[sh.topk.front()].qlen, survivorsPerQuery);
          if (!take.empty() && est + e > st.topkBudgetBytes) break;
          est += e;
          take.push_back(sh.topk.front());
          sh.topk.pop_front();
        }
        batchNo = int32_t(st.topkBatches++);
        st.peakTopkBytes = std::max(st.peakTopkBytes, est);
      }
      const double admit = ms_since(t0);
      gather_queries(take, qdb);
      EdrBatchStats es;
      const int rc = grid_.similarityBatched(TrajectoryView::of(qdb), cfg_.k, res, &es);
      if (rc != GAT_OK) {
        sh.fail(rc);
        return;
      }
      finish_topk(take, res);
      {
        // The survivors this batch really had, folded into the estimate the
        // next admissions charge.
        std::lock_guard<std::mutex> lk(sh.m);
        survivorsPerQuery = 0.75 * survivorsPerQuery + 0.25 * double(es.pairs) / double(take.size());
        st.peakTopkSurvivors = std::max<int64_t>(st.peakTopkSurvivors, es.pairs);
      }
      const double done = ms_since(t0);
      for (int i : take) {
        out[i].admit_ms = admit;
        out[i].done_ms = done;
        out[i].batch = batchNo;
        out[i].count = 0;
        out[i].digest = 0;
      }
      std::lock_guard<std::mutex> lk(sh.m);
      st.topkDone += int64_t(take.size());
    }
  };

  if (anyTopk) {
    // Size the top-k buffers before the clock starts with one batch of the
    // run's own queries (its answers are discarded).
    std::vector<int> first;
    for (int i = 0; i < n && int(first.size()) < cfg_.maxTopkBatch; ++i)
      if (reqs[i].kind == GATH_REQ_TOPK) first.push_back(i);
    TrajectoryDB qdb;
    std::vector<TopK> res;
    gather_queries(first, qdb);
    EdrBatchStats es;
    if (int rc = grid_.similarityBatched(TrajectoryView::of(qdb), cfg_.k, res, &es)) return rc;
    survivorsPerQuery = double(es.pairs) / double(std::max<size_t>(first.size(), 1));
  }
  std::vector<std::thread> workers;
  const int64_t allocs0 = gath_device_alloc_count();
  t0 = Clock::now();
  for (auto& s : slots) workers.emplace_back(range_worker, s.get());
  if (anyTopk) workers.emplace_back(topk_worker);
  // Arrivals: each request joins its queue at its arrival time.
  for (int i = 0; i < n; ++i) {
    const double wait = reqs[i].arrival_ms - ms_since(t0);
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
    if (reqs[i].kind == GATH_REQ_RANGE) {
      int64_t points = 0, leaves = 0;
      const float* b = reqs[i].box;
      grid_.selectionCost(MBB(b[0], b[1], b[2], b[3]), &points, &leaves);
      rangeNeed[size_t(i)] = rangePairBound(points, leaves);
    }
    std::lock_guard<std::mutex> lk(sh.m);
    if (sh.failed) break;
    (reqs[i].kind == GATH_REQ_RANGE ? sh.range : sh.topk).push_back(i);
    sh.cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(sh.m);
    sh.arrivalsDone = true;
    sh.cv.notify_all();
  }
  for (auto& w : workers) w.join();
  st.wallMs = ms_since(t0);
  st.deviceAllocs = gath_device_alloc_count() - allocs0;
  if (sh.failed) {
    gath_set_error("%s", sh.error.c_str());
    return sh.rc;
  }
  if (stats) *stats = st;
  return GAT_OK;
}
