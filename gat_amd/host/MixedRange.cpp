#include "MixedRange.h"

#include <algorithm>
#include <chrono>

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

MixedRange::~MixedRange() {
  if (graph_) gat_graph_destroy(graph_);
  if (fork_) gat_event_destroy(fork_);
  for (void* e : join_)
    if (e) gat_event_destroy(e);
}

RangeEngine& MixedRange::engine(int i) { return i == 0 ? fsg_.engine() : i == 1 ? grid_.engine() : stig_.engine(); }

const DeviceTree& MixedRange::tree(int i) {
  return i == 0 ? fsg_.deviceTree() : i == 1 ? grid_.deviceTree() : stig_.deviceTree();
}

// One graph, three branches: the coordinating stream records `fork_`, each
// index's stream waits for it, queues its pipeline and records its join
// event, and the coordinating stream waits for all three.
int MixedRange::capture(const int nq[3]) {
  if (graph_) gat_graph_destroy(graph_);
  graph_ = nullptr;
  void* c = coord_.get();
  GATH_TRY(gat_graph_begin(c), "mixed capture");
  int rc = gat_event_record(fork_, c);
  for (int i = 0; i < 3 && rc == GAT_OK; ++i) {
    if (nq[i] == 0) continue;
    void* s = engine(i).stream();
    rc = gat_stream_wait_event(s, fork_);
    if (rc == GAT_OK) rc = engine(i).enqueueBatch(tree(i), nq[i], batch_[i]);
    if (rc == GAT_OK) rc = gat_event_record(join_[i], s);
    if (rc == GAT_OK) rc = gat_stream_wait_event(c, join_[i]);
  }
  void* g = nullptr;
  const int rc_end = gat_graph_end(c, &g);
  if (rc != GAT_OK) return gath_fail_hip(rc, "mixed capture");
  GATH_TRY(rc_end, "mixed capture");
  graph_ = g;
  std::copy(nq, nq + 3, shape_);
  for (int i = 0; i < 3; ++i) keys_[i] = engine(i).layoutKey(tree(i), nq[i], batch_[i]);
  ++captures;
  return GAT_OK;
}

bool MixedRange::keysMatch(const int nq[3]) {
  for (int i = 0; i < 3; ++i)
    if (nq[i] > 0 && engine(i).layoutKey(tree(i), nq[i], batch_[i]) != keys_[i]) return false;
  return true;
}

int MixedRange::run(const MBB* bounds, int n, const MixedConfig& cfg, MixedResult* out) {
  if (!out || n < 0 || (n > 0 && !bounds) || cfg.groupSize < 1) {
    gath_set_error("MixedRange::run: bad arguments");
    return GAT_ERR_ARG;
  }
  for (int i = 0; i < 3; ++i)
    if (tree(i).nnodes == 0) {
      gath_set_error("MixedRange::run: index %d has no device tree (build it with upload)", i);
      return GAT_ERR_STATE;
    }
  const double t0 = now_ms();
  if (int rc = coord_.create()) return rc;
  if (!fork_) GATH_TRY(gat_event_create(&fork_), "mixed events");
  for (void*& e : join_)
    if (!e) GATH_TRY(gat_event_create(&e), "mixed events");
  *out = MixedResult{};
  out->count.assign(size_t(n), 0);
  out->digest.assign(size_t(n), 0);
  out->route.assign(size_t(n), 0);
  std::vector<int> list[3];
  for (int q = 0; q < n; ++q) {
    const double area = std::max(0.0, double(bounds[q].xmax) - bounds[q].xmin) *
                        std::max(0.0, double(bounds[q].ymax) - bounds[q].ymin);
    const int r = area < cfg.fsgMaxArea ? 0 : area < cfg.gridMaxArea ? 1 : 2;
    out->route[q] = int8_t(r);
    list[r].push_back(q);
  }
  // The captured pipelines are the leaf-major form: keep small groups off the
  // direct form while this runs.
  int saved[3];
  for (int i = 0; i < 3; ++i) {
    saved[i] = engine(i).directMaxQueries;
    engine(i).directMaxQueries = 0;
  }
  auto restore = [&] {
    for (int i = 0; i < 3; ++i) engine(i).directMaxQueries = saved[i];
  };
  size_t groups = 0;
  for (auto& l : list) groups = std::max(groups, (l.size() + size_t(cfg.groupSize) - 1) / size_t(cfg.groupSize));
  std::vector<MBB> boxes[3];
  std::vector<uint64_t> dig;
  for (size_t gi = 0; gi < groups; ++gi) {
    int nq[3];
    for (int i = 0; i < 3; ++i) {
      const size_t b = gi * size_t(cfg.groupSize);
      nq[i] = b >= list[i].size() ? 0 : int(std::min(list[i].size() - b, size_t(cfg.groupSize)));
      boxes[i].clear();
      for (int k = 0; k < nq[i]; ++k) boxes[i].push_back(bounds[list[i][b + size_t(k)]]);
    }
    bool done = false;
    if (cfg.graph && graph_ && std::equal(nq, nq + 3, shape_)) {
      // Replay: stage every index's rectangles, launch the one graph, check capacities.
      int rc = GAT_OK;
      for (int i = 0; i < 3 && rc == GAT_OK; ++i)
        if (nq[i] > 0) rc = engine(i).stageBatch(tree(i), boxes[i].data(), nq[i]);
      for (int i = 0; i < 3 && rc == GAT_OK; ++i)
        if (nq[i] > 0) rc = gat_stream_sync(engine(i).stream());
      // A workspace that moved since the capture: the graph's kernels would
      // read freed memory, so this group runs step by step and is recaptured.
      const bool current = rc == GAT_OK && keysMatch(nq);
      if (rc == GAT_OK && current) rc = gat_graph_launch(graph_, coord_.get());
      if (rc == GAT_OK) rc = coord_.sync();
      bool fits = current;
      for (int i = 0; i < 3 && rc == GAT_OK && current; ++i) {
        bool f = true;
        if (nq[i] > 0) rc = engine(i).finishBatch(nq[i], batch_[i], &f);
        fits = fits && f;
      }
      if (rc != GAT_OK) {
        restore();
        return gath_fail_hip(rc, "mixed replay");
      }
      done = fits;
      if (done) ++out->replays;
    }
    if (!done) {
      // Step by step (sizes every workspace, which may move its buffers, so
      // the captured graph is dropped first), then capture this shape.
      if (graph_) gat_graph_destroy(graph_);
      graph_ = nullptr;
      std::fill(shape_, shape_ + 3, -1);
      int rc = GAT_OK;
      if (rc == GAT_OK && nq[0] > 0) rc = fsg_.rangeQueryBatchedGPU(boxes[0].data(), nq[0], batch_[0], false);
      if (rc == GAT_OK && nq[1] > 0) rc = grid_.rangeQueryBatchedGPU(boxes[1].data(), nq[1], batch_[1], false);
      if (rc == GAT_OK && nq[2] > 0) rc = stig_.rangeQueryBatchedGPU(boxes[2].data(), nq[2], batch_[2], false);
      // A shape met again (every group but the last) is worth a graph.
      if (rc == GAT_OK && cfg.graph && gi + 2 < groups) rc = capture(nq);
      if (rc != GAT_OK) {
        restore();
        return rc;
      }
    }
    const double d0 = now_ms();
    for (int i = 0; i < 3; ++i) {
      if (nq[i] == 0) continue;
      RangeBatch& b = batch_[i];
      int rc = b.syncHost();
      if (rc == GAT_OK) rc = digests_[i].reserve(size_t(nq[i]) * 8);
      dig.resize(size_t(nq[i]));
      if (rc == GAT_OK) rc = b.digestsInto(digests_[i].as<uint64_t>(), dig.data());
      if (rc != GAT_OK) {
        restore();
        return rc;
      }
      const size_t base = gi * size_t(cfg.groupSize);
      for (int k = 0; k < nq[i]; ++k) {
        const int q = list[i][base + size_t(k)];
        out->count[size_t(q)] = b.count[size_t(k)];
        out->digest[size_t(q)] = dig[size_t(k)];
      }
      out->pairs += b.pairTotal;
      out->candidates += b.candidates;
    }
    out->digestMs += now_ms() - d0;
    ++out->groups;
  }
  restore();
  out->ms = now_ms() - t0;
  return GAT_OK;
}
