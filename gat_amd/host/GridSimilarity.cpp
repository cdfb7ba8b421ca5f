// Similarity queries of the Grid (Grid.cpp:852-1755): EDR between query
// trajectories and every indexed trajectory, then top-k.
#include <algorithm>
#include <numeric>

#include "Grid.h"

int Grid::uploadTrajectories(const TrajectoryView& db) {
  const int64_t N = db.pointNum();
  const int T = db.trajNum;
  trajLenHost_.resize(size_t(T));
  std::vector<float> xy(size_t(N) * 2);
  for (int64_t i = 0; i < N; ++i) {
    xy[2 * i] = db.lon[i];
    xy[2 * i + 1] = db.lat[i];
  }
  std::vector<gat_edr_task> tasks(static_cast<size_t>(T));
  for (int t = 0; t < T; ++t) {
    trajLenHost_[t] = int32_t(db.offsets[t + 1] - db.offsets[t]);
    tasks[t] = gat_edr_task{0, t + 1};
  }
  void* s = engine_.stream();
  if (int rc = trajXY_.reserve(std::max<size_t>(xy.size() * 4, 8))) return rc;
  if (int rc = trajOff_.reserve(std::max<size_t>(size_t(T) * 8, 8))) return rc;
  if (int rc = trajLen_.reserve(std::max<size_t>(size_t(T) * 4, 4))) return rc;
  if (int rc = allTasks_.reserve(std::max<size_t>(size_t(T) * 8, 8))) return rc;
  if (int rc = dist_.reserve(std::max<size_t>(size_t(T) * 4, 4))) return rc;
  GATH_TRY(gat_memcpy_h2d(trajXY_.get(), xy.data(), xy.size() * 4, s), "trajectory upload");
  GATH_TRY(gat_memcpy_h2d(trajOff_.get(), db.offsets, size_t(T) * 8, s), "trajectory upload");
  GATH_TRY(gat_memcpy_h2d(trajLen_.get(), trajLenHost_.data(), size_t(T) * 4, s), "trajectory upload");
  GATH_TRY(gat_memcpy_h2d(allTasks_.get(), tasks.data(), size_t(T) * 8, s), "trajectory upload");
  GATH_TRY(gat_stream_sync(s), "trajectory upload");
  return GAT_OK;
}

int Grid::ensureQueryStore(const TrajectoryView& q) {
  const int64_t N = q.pointNum();
  std::vector<float> xy(size_t(N) * 2);
  for (int64_t i = 0; i < N; ++i) {
    xy[2 * i] = q.lon[i];
    xy[2 * i + 1] = q.lat[i];
  }
  std::vector<int32_t> off(static_cast<size_t>(q.trajNum)), len(size_t(q.trajNum));
  for (int i = 0; i < q.trajNum; ++i) {
    off[i] = int32_t(q.offsets[i]);
    len[i] = int32_t(q.offsets[i + 1] - q.offsets[i]);
    if (len[i] < 1 || len[i] > GAT_MAX_TRAJ_LEN) {
      gath_set_error("similarity: query %d has %d points (supported: 1..%d)", i, len[i], GAT_MAX_TRAJ_LEN);
      return GAT_ERR_ARG;
    }
  }
  void* s = engine_.stream();
  if (int rc = queryXY_.reserve(std::max<size_t>(xy.size() * 4, 8))) return rc;
  if (int rc = queryOff_.reserve(std::max<size_t>(off.size() * 4, 4))) return rc;
  if (int rc = queryLen_.reserve(std::max<size_t>(len.size() * 4, 4))) return rc;
  GATH_TRY(gat_memcpy_h2d(queryXY_.get(), xy.data(), xy.size() * 4, s), "query upload");
  GATH_TRY(gat_memcpy_h2d(queryOff_.get(), off.data(), off.size() * 4, s), "query upload");
  GATH_TRY(gat_memcpy_h2d(queryLen_.get(), len.data(), len.size() * 4, s), "query upload");
  GATH_TRY(gat_stream_sync(s), "query upload");
  return GAT_OK;
}

namespace {
void topk_of(const std::vector<int32_t>& dist, int k, TopK& out) {
  const int n = int(dist.size());
  std::vector<int32_t> idx(static_cast<size_t>(n));
  std::iota(idx.begin(), idx.end(), 0);
  const int kk = std::min(k, n);
  // Ascending (distance, trajectory id); index i is trajectory i + 1.
  std::partial_sort(idx.begin(), idx.begin() + kk, idx.end(),
                    [&](int32_t a, int32_t b) { return dist[a] != dist[b] ? dist[a] < dist[b] : a < b; });
  out.ids.resize(size_t(kk));
  out.dist.resize(size_t(kk));
  for (int i = 0; i < kk; ++i) {
    out.ids[i] = idx[i] + 1;
    out.dist[i] = dist[idx[i]];
  }
}
}  // namespace

int Grid::similarityPlain(const TrajectoryView& queries, int kValue, std::vector<TopK>& out) {
  for (int t = 0; t < trajNum; ++t)
    if (trajLenHost_[t] < 1 || trajLenHost_[t] > GAT_MAX_TRAJ_LEN) {
      gath_set_error("similarity: trajectory %d has %d points (supported: 1..%d)", t + 1, trajLenHost_[t],
                     GAT_MAX_TRAJ_LEN);
      return GAT_ERR_ARG;
    }
  if (int rc = ensureQueryStore(queries)) return rc;
  out.assign(size_t(queries.trajNum), TopK{});
  std::vector<int32_t> dist(static_cast<size_t>(trajNum));
  void* s = engine_.stream();
  for (int q = 0; q < queries.trajNum; ++q) {
    // The candidate list is every trajectory; task i pairs "query 0" of the
    // shifted query arrays (i.e. query q) with trajectory i + 1.
    GATH_TRY(gatEDRDistanceBatchHandler(trajNum, allTasks_.as<gat_edr_task>(), queryXY_.as<float>(),
                                        queryOff_.as<int32_t>() + q, queryLen_.as<int32_t>() + q, trajXY_.as<float>(),
                                        trajOff_.as<int64_t>(), trajLen_.as<int32_t>(), epsilon, dist_.as<int32_t>(),
                                        s),
             "EDR batch");
    GATH_TRY(gat_memcpy_d2h(dist.data(), dist_.get(), size_t(trajNum) * 4, s), "EDR readback");
    GATH_TRY(gat_stream_sync(s), "EDR readback");
    topk_of(dist, kValue, out[q]);
  }
  return GAT_OK;
}

int Grid::SimilarityQueryBatchOnGPU(Trajectory* qTra, int queryTrajNum, int* topKSimilarityTraj, int kValue) {
  TrajectoryDB q;
  for (int i = 0; i < queryTrajNum; ++i) {
    for (int k = 0; k < qTra[i].length; ++k) {
      q.lon.push_back(qTra[i].points[k].lon);
      q.lat.push_back(qTra[i].points[k].lat);
    }
    q.offsets.push_back(int64_t(q.lon.size()));
  }
  std::vector<TopK> res;
  if (int rc = similarityPlain(TrajectoryView::of(q), kValue, res)) return rc;
  for (int i = 0; i < queryTrajNum; ++i)
    for (int k = 0; k < kValue; ++k)
      topKSimilarityTraj[i * kValue + k] = k < int(res[i].ids.size()) ? res[i].ids[k] : 0;
  return GAT_OK;
}

int Grid::edrAll(const float* qlon, const float* qlat, int qlen, int32_t* dist) {
  int64_t off[2] = {0, qlen};
  TrajectoryView q{off, 1, qlon, qlat};
  if (int rc = ensureQueryStore(q)) return rc;
  void* s = engine_.stream();
  GATH_TRY(gatEDRDistanceBatchHandler(trajNum, allTasks_.as<gat_edr_task>(), queryXY_.as<float>(),
                                      queryOff_.as<int32_t>(), queryLen_.as<int32_t>(), trajXY_.as<float>(),
                                      trajOff_.as<int64_t>(), trajLen_.as<int32_t>(), epsilon, dist_.as<int32_t>(), s),
           "EDR batch");
  GATH_TRY(gat_memcpy_d2h(dist, dist_.get(), size_t(trajNum) * 4, s), "EDR readback");
  GATH_TRY(gat_stream_sync(s), "EDR readback");
  return GAT_OK;
}
