// Similarity queries of the Grid (Grid.cpp:852-1755): EDR between query
// trajectories and every indexed trajectory, then top-k.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "Grid.h"
#include "Parallel.h"

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// A NaN-padded, interleaved copy of a trajectory set: GAT_EDR_PAD NaN points
// before every trajectory and GAT_EDR_TAIL_PAD after the last.  offsets[i] is
// the first point of trajectory i + 1; boxes[4i..4i+3] its bounding box.
struct Padded {
  std::vector<float> xy;
  std::vector<int64_t> offsets;
  std::vector<int32_t> lengths;
  std::vector<float> boxes;
};

Padded pad_store(const TrajectoryView& v) {
  Padded p;
  const float nan = std::numeric_limits<float>::quiet_NaN();
  const int64_t total = v.pointNum() + int64_t(GAT_EDR_PAD) * v.trajNum + GAT_EDR_TAIL_PAD + GAT_EDR_PAD;
  p.xy.assign(size_t(total) * 2, nan);
  p.offsets.resize(size_t(v.trajNum));
  p.lengths.resize(size_t(v.trajNum));
  p.boxes.resize(size_t(v.trajNum) * 4);
  int64_t at = GAT_EDR_PAD;
  for (int t = 0; t < v.trajNum; ++t) {
    const int64_t b = v.offsets[t], e = v.offsets[t + 1];
    p.offsets[t] = at;
    p.lengths[t] = int32_t(e - b);
    float xl = INFINITY, yl = INFINITY, xh = -INFINITY, yh = -INFINITY;
    for (int64_t i = b; i < e; ++i, ++at) {
      p.xy[2 * at] = v.lon[i];
      p.xy[2 * at + 1] = v.lat[i];
      xl = std::min(xl, v.lon[i]);
      xh = std::max(xh, v.lon[i]);
      yl = std::min(yl, v.lat[i]);
      yh = std::max(yh, v.lat[i]);
    }
    p.boxes[4 * t] = xl;
    p.boxes[4 * t + 1] = yl;
    p.boxes[4 * t + 2] = xh;
    p.boxes[4 * t + 3] = yh;
    at += GAT_EDR_PAD;
  }
  return p;
}

template <class T>
int upload_vec(DeviceBuffer& d, const std::vector<T>& v, void* s, const char* what) {
  if (int rc = d.reserve(std::max<size_t>(v.size() * sizeof(T), 16))) return rc;
  GATH_TRY(gat_memcpy_h2d(d.get(), v.data(), v.size() * sizeof(T), s), what);
  return GAT_OK;
}

int strips_of(int n) { return (n + 15) / 16; }

}  // namespace

int Grid::uploadTrajectories(const TrajectoryView& db) {
  const int T = db.trajNum;
  Padded p = pad_store(db);
  trajLenHost_ = p.lengths;
  badTrajLen_ = 0;
  for (int t = 0; t < T && badTrajLen_ == 0; ++t)
    if (trajLenHost_[t] < 1 || trajLenHost_[t] > GAT_MAX_TRAJ_LEN) badTrajLen_ = t + 1;
  std::vector<gat_edr_task> tasks(static_cast<size_t>(T));
  for (int t = 0; t < T; ++t) tasks[t] = gat_edr_task{0, t + 1};
  void* s = engine_.stream();
  if (int rc = upload_vec(trajXY_, p.xy, s, "trajectory upload")) return rc;
  if (int rc = upload_vec(trajOff_, p.offsets, s, "trajectory upload")) return rc;
  if (int rc = upload_vec(trajLen_, p.lengths, s, "trajectory upload")) return rc;
  if (int rc = upload_vec(trajBox_, p.boxes, s, "trajectory upload")) return rc;
  if (int rc = upload_vec(allTasks_, tasks, s, "trajectory upload")) return rc;
  if (int rc = dist_.reserve(std::max<size_t>(size_t(T) * 4, 4))) return rc;
  GATH_TRY(gat_stream_sync(s), "trajectory upload");
  lengthIds_.assign(GAT_MAX_TRAJ_LEN + 1, {});
  lengthOrder_.assign(GAT_MAX_TRAJ_LEN + 1, {});
  for (int t = 0; t < T; ++t)
    if (trajLenHost_[t] >= 1 && trajLenHost_[t] <= GAT_MAX_TRAJ_LEN) lengthIds_[trajLenHost_[t]].push_back(t + 1);
  // The same groups for the device merge: 0-based ids, offsets per length.
  std::vector<int32_t> lenOff(GAT_MAX_TRAJ_LEN + 2, 0), lenIds;
  for (int L = 0; L <= GAT_MAX_TRAJ_LEN; ++L) lenOff[L + 1] = lenOff[L] + int32_t(lengthIds_[L].size());
  lenIds.reserve(size_t(lenOff.back()));
  for (int L = 0; L <= GAT_MAX_TRAJ_LEN; ++L)
    for (int32_t id : lengthIds_[L]) lenIds.push_back(id - 1);
  if (int rc = upload_vec(lenIdsDev_, lenIds, s, "length groups upload")) return rc;
  if (int rc = upload_vec(lenOffDev_, lenOff, s, "length groups upload")) return rc;
  GATH_TRY(gat_stream_sync(s), "length groups upload");
  return buildCandidateGrid(p.boxes, T);
}

namespace {
// Cell column (or row) of v in the candidate grid, with the float operations
// gat_edr_filter_grid applies to the same boxes.
int cand_cell(float v, float o, float inv, int g) {
  float t = std::floor((v - o) * inv);
  t = std::fmin(std::fmax(t, 0.0f), float(g - 1));
  return int(t);
}
}  // namespace

// Cells about twice the median box side (at most 1024 per side) over the
// extent of the finite boxes; a trajectory whose box touches more than 64
// cells, or is not finite or inverted (an empty or all-NaN trajectory), is
// tested by every query instead.  On the 200k-trajectory top-k workload:
// 313 cells per side, 2.2 entries per trajectory.
int Grid::buildCandidateGrid(const std::vector<float>& boxes, int T) {
  constexpr int kMaxCellsPerBox = 64;
  auto finite_box = [&](int t) {
    const float* b = &boxes[size_t(t) * 4];
    return std::isfinite(b[0]) && std::isfinite(b[1]) && std::isfinite(b[2]) && std::isfinite(b[3]) &&
           b[0] <= b[2] && b[1] <= b[3];
  };
  float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  std::vector<float> side;
  side.reserve(size_t(T));
  for (int t = 0; t < T; ++t) {
    if (!finite_box(t)) continue;
    const float* b = &boxes[size_t(t) * 4];
    x0 = std::min(x0, b[0]);
    y0 = std::min(y0, b[1]);
    x1 = std::max(x1, b[2]);
    y1 = std::max(y1, b[3]);
    side.push_back(std::max(b[2] - b[0], b[3] - b[1]));
  }
  int g = 1;
  float invx = 0.0f, invy = 0.0f;
  if (!side.empty()) {
    std::nth_element(side.begin(), side.begin() + side.size() / 2, side.end());
    const float ex = x1 - x0, ey = y1 - y0, e = std::max(ex, ey);
    const float cell = std::max(2.0f * side[side.size() / 2], e / 1024.0f);
    if (cell > 0.0f) g = std::max(1, int(std::min(1024.0f, std::ceil(e / cell))));
    invx = ex > 0.0f ? float(g) / ex : 0.0f;
    invy = ey > 0.0f ? float(g) / ey : 0.0f;
  } else {
    x0 = y0 = 0.0f;
  }
  const size_t cells = size_t(g) * size_t(g);
  std::vector<uint32_t> start(cells + 1, 0);
  std::vector<int32_t> always;
  std::vector<int> c0(static_cast<size_t>(T), -1), c1(static_cast<size_t>(T), 0), r0(static_cast<size_t>(T), 0),
      r1(static_cast<size_t>(T), 0);
  for (int t = 0; t < T; ++t) {
    c0[t] = -1;
    if (!finite_box(t)) {
      always.push_back(t);
      continue;
    }
    const float* b = &boxes[size_t(t) * 4];
    c0[t] = cand_cell(b[0], x0, invx, g);
    c1[t] = cand_cell(b[2], x0, invx, g);
    r0[t] = cand_cell(b[1], y0, invy, g);
    r1[t] = cand_cell(b[3], y0, invy, g);
    if (int64_t(c1[t] - c0[t] + 1) * int64_t(r1[t] - r0[t] + 1) > kMaxCellsPerBox) {
      c0[t] = -1;
      always.push_back(t);
      continue;
    }
    for (int r = r0[t]; r <= r1[t]; ++r)
      for (int c = c0[t]; c <= c1[t]; ++c) ++start[size_t(r) * g + c + 1];
  }
  for (size_t i = 0; i < cells; ++i) start[i + 1] += start[i];
  const size_t n = start[cells];
  std::vector<uint32_t> at(start.begin(), start.end() - 1);
  std::vector<int32_t> id(std::max<size_t>(n, 1), 0);
  std::vector<float> box(std::max<size_t>(n, 1) * 4, 0.0f);
  for (int t = 0; t < T; ++t) {
    if (c0[t] < 0) continue;
    for (int r = r0[t]; r <= r1[t]; ++r)
      for (int c = c0[t]; c <= c1[t]; ++c) {
        const uint32_t e = at[size_t(r) * g + c]++;
        id[e] = t;
        std::memcpy(&box[size_t(e) * 4], &boxes[size_t(t) * 4], 16);
      }
  }
  if (always.empty()) always.push_back(-1);  // a valid pointer; nalways stays 0 below
  void* s = engine_.stream();
  if (int rc = upload_vec(candStart_, start, s, "candidate grid upload")) return rc;
  if (int rc = upload_vec(candId_, id, s, "candidate grid upload")) return rc;
  if (int rc = upload_vec(candBox_, box, s, "candidate grid upload")) return rc;
  if (int rc = upload_vec(candAlways_, always, s, "candidate grid upload")) return rc;
  GATH_TRY(gat_stream_sync(s), "candidate grid upload");
  candGrid_ = gat_edr_cand_grid{g,
                                x0,
                                y0,
                                invx,
                                invy,
                                candStart_.as<uint32_t>(),
                                candId_.as<int32_t>(),
                                candBox_.as<float>(),
                                candAlways_.as<int32_t>(),
                                always.size() == 1 && always[0] < 0 ? 0 : int32_t(always.size())};
  return GAT_OK;
}

int Grid::filterPairs(int nq, int64_t cap, void* stream) {
  GATH_TRY(gat_edr_filter_grid(queryBox_.as<float>(), nq, trajBox_.as<float>(), trajNum, &candGrid_, epsilon,
                               pairsDev_.as<int32_t>(), cap, counterDev_.as<uint64_t>(), stream),
           "candidate filter");
  return GAT_OK;
}

int Grid::ensureQueryStore(const TrajectoryView& q) {
  for (int i = 0; i < q.trajNum; ++i) {
    const int64_t len = q.offsets[i + 1] - q.offsets[i];
    if (len < 1 || len > GAT_MAX_TRAJ_LEN) {
      gath_set_error("similarity: query %d has %lld points (supported: 1..%d)", i, (long long)len, GAT_MAX_TRAJ_LEN);
      return GAT_ERR_ARG;
    }
  }
  // The offsets, then the raw coordinates, go up from pinned memory in two
  // DMAs (the second, of the latitudes, copied on the host while the first
  // runs), and the device packs the padded store and the boxes
  // (gat_edr_pack_trajectories).  The offsets and lengths the kernels read
  // are views into the same device buffer.
  const int T = q.trajNum;
  const int64_t o0 = T > 0 ? q.offsets[0] : 0;
  const int64_t n = T > 0 ? q.offsets[T] - o0 : 0;
  const int64_t total = n + int64_t(GAT_EDR_PAD) * T + GAT_EDR_TAIL_PAD + GAT_EDR_PAD;
  auto up16 = [](size_t b) { return (b + 15) & ~size_t(15); };
  const size_t braw = up16(size_t(n) * 4), bsrc = up16((size_t(T) + 1) * 8), b64 = up16(size_t(T) * 8);
  const size_t b32 = up16(size_t(T) * 4);
  const size_t meta = bsrc + b64 + 2 * b32;
  if (int rc = queryStage_.reserve(meta + 2 * braw)) return rc;
  char* base = queryStage_.as<char>();
  int64_t* src = reinterpret_cast<int64_t*>(base);
  int64_t* off64 = reinterpret_cast<int64_t*>(base + bsrc);
  int32_t* off32 = reinterpret_cast<int32_t*>(base + bsrc + b64);
  int32_t* lens = reinterpret_cast<int32_t*>(base + bsrc + b64 + b32);
  float* lon = reinterpret_cast<float*>(base + meta);
  float* lat = reinterpret_cast<float*>(base + meta + braw);
  queryLenHost_.resize(size_t(T));
  int64_t at = GAT_EDR_PAD;
  for (int t = 0; t < T; ++t) {
    const int64_t len = q.offsets[t + 1] - q.offsets[t];
    src[t] = q.offsets[t] - o0;
    off64[t] = at;
    off32[t] = int32_t(at);
    lens[t] = int32_t(len);
    queryLenHost_[size_t(t)] = int32_t(len);
    at += len + GAT_EDR_PAD;
  }
  src[T] = n;
  void* s = engine_.stream();
  if (int rc = queryUp_.reserve(std::max<size_t>(meta + 2 * braw, 16))) return rc;
  char* dev = queryUp_.as<char>();
  if (n > 0) std::memcpy(lon, q.lon + o0, size_t(n) * 4);
  GATH_TRY(gat_memcpy_h2d(dev, base, meta + braw, s), "query upload");
  if (n > 0) {
    std::memcpy(lat, q.lat + o0, size_t(n) * 4);
    GATH_TRY(gat_memcpy_h2d(dev + meta + braw, lat, braw, s), "query upload");
  }
  querySrcOffDev_ = reinterpret_cast<const int64_t*>(dev);
  queryOff64Dev_ = reinterpret_cast<const int64_t*>(dev + bsrc);
  queryOffDev_ = reinterpret_cast<const int32_t*>(dev + bsrc + b64);
  queryLenDev_ = reinterpret_cast<const int32_t*>(dev + bsrc + b64 + b32);
  if (int rc = queryXY_.reserve(size_t(total) * 8)) return rc;
  if (int rc = queryBox_.reserve(std::max<size_t>(size_t(T) * 16, 16))) return rc;
  const float* rlon = reinterpret_cast<const float*>(dev + meta);
  const float* rlat = reinterpret_cast<const float*>(dev + meta + braw);
  GATH_TRY(gat_edr_pack_trajectories(rlon, rlat, querySrcOffDev_, queryOff64Dev_, T, total, queryXY_.as<float>(),
                                     queryBox_.as<float>(), s),
           "query packing");
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

// Every indexed trajectory must have a length the EDR kernels take: one
// outside 1..GAT_MAX_TRAJ_LEN would come back as distance -1 (the kernels'
// marker for a malformed pair), which a top-k would rank first.
int Grid::checkTrajLengths(const char* what) const {
  if (badTrajLen_ == 0) return GAT_OK;
  gath_set_error("%s: trajectory %d has %d points (supported: 1..%d)", what, badTrajLen_,
                 trajLenHost_[size_t(badTrajLen_) - 1], GAT_MAX_TRAJ_LEN);
  return GAT_ERR_ARG;
}

int Grid::similarityPlain(const TrajectoryView& queries, int kValue, std::vector<TopK>& out) {
  if (int rc = checkTrajLengths("similarity")) return rc;
  if (int rc = ensureQueryStore(queries)) return rc;
  out.assign(size_t(queries.trajNum), TopK{});
  std::vector<int32_t> dist(static_cast<size_t>(trajNum));
  void* s = engine_.stream();
  for (int q = 0; q < queries.trajNum; ++q) {
    // The candidate list is every trajectory; task i pairs "query 0" of the
    // shifted query arrays (i.e. query q) with trajectory i + 1.
    GATH_TRY(gatEDRDistanceBatchHandler(trajNum, allTasks_.as<gat_edr_task>(), queryXY_.as<float>(),
                                        queryOffDev_ + q, queryLenDev_ + q, trajXY_.as<float>(),
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
                                      queryOffDev_, queryLenDev_, trajXY_.as<float>(),
                                      trajOff_.as<int64_t>(), trajLen_.as<int32_t>(), epsilon, dist_.as<int32_t>(), s),
           "EDR batch");
  GATH_TRY(gat_memcpy_d2h(dist, dist_.get(), size_t(trajNum) * 4, s), "EDR readback");
  GATH_TRY(gat_stream_sync(s), "EDR readback");
  return GAT_OK;
}

// ------------------------------------------------------------------ batched

const std::vector<int32_t>& Grid::mismatchTable() {
  if (!ftable_.empty()) return ftable_;
  const size_t w = GAT_MAX_TRAJ_LEN + 1;
  void* s = engine_.stream();
  std::vector<int32_t> t(w * w);
  if (ftableDev_.reserve(w * w * 4) == GAT_OK && gat_edr_mismatch_table(GAT_MAX_TRAJ_LEN, ftableDev_.as<int32_t>(), s) == GAT_OK &&
      gat_memcpy_d2h(t.data(), ftableDev_.get(), w * w * 4, s) == GAT_OK && gat_stream_sync(s) == GAT_OK)
    ftable_.swap(t);
  else
    gath_fail_hip(GAT_ERR_HIP, "mismatch table");
  return ftable_;
}

// Orders the pairs so that consecutive groups of four share a strip count
// (ceil(min length / 16)) and similar column counts, longest first, and runs
// the batched kernel; dist receives each pair's distance in the original order.
size_t Grid::runBatchedStaging(int64_t np) {
  // Sorted pairs, jobs (at most one per pair), distances, and the regions'
  // alignment.
  return size_t(np) * 8 + (size_t(np) + 2) * sizeof(gat_edr_job) + size_t(np) * 4 + 3 * 256;
}

int Grid::runBatched(std::vector<int32_t>& pairs, int queryNum, std::vector<int32_t>& dist, EdrBatchStats* st) {
  const double h0 = now_ms();
  const int64_t np = int64_t(pairs.size() / 2);
  dist.assign(size_t(np), -1);
  if (np == 0) return GAT_OK;
  // Length buckets: pairs whose longer trajectory is at most edrSplit points
  // run in a launch with a small LDS row buffer (more waves per CU), those up
  // to kMid points in one sized for kMid, four pairs per wave, and the rest one
  // pair per wave (64-lane strips, a quarter as many one after another).
  // Inside a bucket, a counting sort by (strips, longer length), descending,
  // and jobs of up to four pairs with the same strip count.
  constexpr int kM = GAT_MAX_TRAJ_LEN + 1;
  constexpr int kS = GAT_MAX_TRAJ_LEN / 16 + 1;
  constexpr int kBuckets = 4;  // 3: long pairs of a one-round batch over several waves
  constexpr int kMid = 1024;
  constexpr int64_t kLatencyJobsPerSimd = 4;  // the batched kernel holds 4 waves per SIMD: one round
  const int split = std::max(1, std::min(edrSplit, GAT_MAX_TRAJ_LEN));
  // A batch that fits one round of the device's waves (a top-k batch's
  // survivors, a serving batch) ends when its longest pair does, and a pair
  // runs its strips one after another: four pairs per wave on 16-lane strips
  // take a 1000 x 1000 pair through 63 strips, one pair per wave on 64-lane
  // strips through 16.  Such batches send every pair longer than the split to
  // the one-pair-per-wave bucket.
  int64_t props[5] = {0, 0, 0, 0, 0};
  const int64_t cus = gat_device_props(props) == GAT_OK && props[0] > 0 ? props[0] : 256;
  const bool latencyBound = (np + 3) / 4 < cus * 4 * kLatencyJobsPerSimd;
  const int mid = latencyBound ? split : std::max(split, kMid);
  const int waves = (edrWavesPerPair == 2 || edrWavesPerPair == 4 || edrWavesPerPair == 8) ? edrWavesPerPair : 1;
  for (int64_t i = 0; i < np; ++i) {
    const int q = pairs[2 * i], c = pairs[2 * i + 1];
    if (q < 0 || q >= queryNum || c < 0 || c >= trajNum) {
      gath_set_error("batched EDR: pair %lld out of range", (long long)i);
      return GAT_ERR_ARG;
    }
  }
  // In a one-round batch with several waves per long pair, only the pairs
  // that would outlast the batch's longest pair on one wave take several:
  // a pair of s 64-row strips over a b-block row takes s * b blocks on one
  // wave and pipe_blocks(s, b) over `waves` waves (the workgroup's schedule,
  // with the lag of 6 blocks between strips), and the batch cannot end before
  // its longest pair's pipelined schedule does.  The rest run one pair per
  // wave, beside it; several waves on a pair of fewer strips than waves leave
  // the surplus waves idle at every barrier.
  auto pipe_blocks = [&](int64_t s, int64_t b) {
    const int64_t P = std::max<int64_t>(b, int64_t(waves) * 6);
    return ((s - 1) / waves) * P + ((s - 1) % waves) * 6 + b;
  };
  const bool multiWave = latencyBound && waves > 1;
  int64_t critical = 0;
  if (multiWave)
    for (int64_t i = 0; i < np; ++i) {
      const int a = queryLenHost_[pairs[2 * i]], b = trajLenHost_[pairs[2 * i + 1]];
      const int n = std::min(a, b), m = std::max(a, b);
      if (m > mid) critical = std::max(critical, pipe_blocks((n + 63) / 64, (m + 63 + 15) / 16));
    }
  std::vector<int32_t> key(static_cast<size_t>(np));
  int64_t cells = 0;
  int maxlen[kBuckets] = {1, 1, 1, 1};
  for (int64_t i = 0; i < np; ++i) {
    const int a = queryLenHost_[pairs[2 * i]], b = trajLenHost_[pairs[2 * i + 1]];
    const int n = std::min(a, b), m = std::max(a, b);
    cells += int64_t(n) * m;
    int bucket = m <= split ? 0 : m <= mid ? 1 : 2;
    if (bucket == 2 && multiWave &&
        (edrWavesEveryLongPair || int64_t((n + 63) / 64) * ((m + 63 + 15) / 16) > critical))
      bucket = 3;
    maxlen[bucket] = std::max(maxlen[bucket], std::min(m, GAT_MAX_TRAJ_LEN));
    key[i] = (bucket * kS + (kS - 1 - strips_of(n))) * kM + (kM - 1 - m);
  }
  // Stable order by key: two counting passes over 11-bit digits (keys are
  // below kBuckets * kS * kM < 2^22).
  static_assert(int64_t(kBuckets) * kS * kM <= (int64_t(1) << 22), "pair keys must fit two 11-bit digits");
  std::vector<int64_t> perm(static_cast<size_t>(np)), tmp(static_cast<size_t>(np));
  {
    constexpr int kDigit = 11, kRadix = 1 << kDigit;
    std::vector<int64_t> cnt(kRadix + 1);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < np; ++i) ++cnt[size_t(key[i] & (kRadix - 1)) + 1];
    for (int d = 1; d <= kRadix; ++d) cnt[size_t(d)] += cnt[size_t(d) - 1];
    for (int64_t i = 0; i < np; ++i) tmp[size_t(cnt[size_t(key[i] & (kRadix - 1))]++)] = i;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < np; ++i) ++cnt[size_t(key[i] >> kDigit) + 1];
    for (int d = 1; d <= kRadix; ++d) cnt[size_t(d)] += cnt[size_t(d) - 1];
    for (int64_t j = 0; j < np; ++j) {
      const int64_t i = tmp[size_t(j)];
      perm[size_t(cnt[size_t(key[i] >> kDigit)]++)] = i;
    }
  }
  // Sorted pairs, jobs and distances pass through the pinned arena (its
  // callers size it for runBatchedStaging(np)).
  int32_t* sorted = stage_.take_as<int32_t>(size_t(np) * 2);
  gat_edr_job* jobs = stage_.take_as<gat_edr_job>(size_t(np) + 2);
  int32_t* outH = stage_.take_as<int32_t>(size_t(np));
  if (!sorted || !jobs || !outH) {
    gath_set_error("batched EDR: transfer arena too small for %lld pairs", (long long)np);
    return GAT_ERR_STATE;
  }
  int64_t njobs = 0;
  int64_t bucketJobs[kBuckets] = {0, 0, 0, 0};
  int64_t i = 0;
  while (i < np) {
    const int32_t k0 = key[perm[i]] / kM;  // bucket and strip count
    // Jobs of the long buckets run one pair over a whole wave (or workgroup).
    const int32_t per_job = k0 / kS >= 2 ? 1 : 4;
    int32_t c = 0;
    while (c < per_job && i + c < np && key[perm[i + c]] / kM == k0) ++c;
    jobs[njobs++] = gat_edr_job{int32_t(i), c};
    ++bucketJobs[k0 / kS];
    i += c;
  }
  for (int64_t j = 0; j < np; ++j) {
    sorted[2 * j] = pairs[2 * perm[j]];
    sorted[2 * j + 1] = pairs[2 * perm[j] + 1];
  }
  void* s = engine_.stream();
  if (int rc = pairsDev_.reserve(size_t(np) * 8)) return rc;
  if (int rc = jobsDev_.reserve(std::max<size_t>(size_t(njobs) * sizeof(gat_edr_job), 16))) return rc;
  GATH_TRY(gat_memcpy_h2d(pairsDev_.get(), sorted, size_t(np) * 8, s), "pair upload");
  GATH_TRY(gat_memcpy_h2d(jobsDev_.get(), jobs, size_t(njobs) * sizeof(gat_edr_job), s), "job upload");
  if (int rc = outDev_.reserve(size_t(np) * 4)) return rc;
  if (int rc = counterDev_.reserve(16)) return rc;
  const double h1 = now_ms();
  if (int rc = edrTimer_.start(s)) return rc;
  // Each bucket has its own job counter.  The longest bucket holds few jobs
  // whose strips run for milliseconds each: launched first on a side stream, it
  // overlaps the other buckets instead of trailing them on an idle device.
  int64_t j0[kBuckets] = {0, bucketJobs[0], bucketJobs[0] + bucketJobs[1],
                          bucketJobs[0] + bucketJobs[1] + bucketJobs[2]};
  auto launch = [&](int bucket, void* stream) -> int {
    GATH_TRY(gat_edr_batched(pairsDev_.as<int32_t>(), np, jobsDev_.as<gat_edr_job>() + j0[bucket],
                             bucketJobs[bucket], nullptr, counterDev_.as<int32_t>() + bucket, queryXY_.as<float>(),
                             queryOff64Dev_, queryLenDev_, queryNum, trajXY_.as<float>(),
                             trajOff_.as<int64_t>(), trajLen_.as<int32_t>(), trajNum, maxlen[bucket], epsilon,
                             bucket == 0 ? edrPerCuShort : bucket == 1 ? edrPerCuLong : 0,
                             bucket == 3 ? 64 * waves : bucket == 2 ? 64 : 16, outDev_.as<int32_t>(), stream),
             "batched EDR");
    return GAT_OK;
  };
  // Several-wave pairs (bucket 3) and one-wave long pairs (bucket 2) each get
  // a side stream when other buckets have jobs, so the three run together
  // instead of the long ones queueing behind the short ones.
  const int longest = bucketJobs[3] > 0 ? 3 : 2;
  int64_t others = 0;
  for (int bucket = 0; bucket < kBuckets; ++bucket) others += bucket == longest ? 0 : bucketJobs[bucket];
  const bool side = bucketJobs[longest] > 0 && others > 0;
  const bool side2 = side && longest == 3 && bucketJobs[2] > 0 && others > bucketJobs[2];
  if (side) {
    if (int rc = edrSide_.fork(s)) return rc;
    if (int rc = launch(longest, edrSide_.stream())) return rc;
  }
  if (side2) {
    if (int rc = edrSide2_.fork(s)) return rc;
    if (int rc = launch(2, edrSide2_.stream())) return rc;
  }
  for (int bucket = 0; bucket < kBuckets; ++bucket) {
    if (bucketJobs[bucket] == 0 || (side && bucket == longest) || (side2 && bucket == 2)) continue;
    if (int rc = launch(bucket, s)) return rc;
  }
  if (side)
    if (int rc = edrSide_.join(s)) return rc;
  if (side2)
    if (int rc = edrSide2_.join(s)) return rc;
  if (int rc = edrTimer_.stop(s)) return rc;
  GATH_TRY(gat_memcpy_d2h(outH, outDev_.get(), size_t(np) * 4, s), "batched EDR readback");
  GATH_TRY(gat_stream_sync(s), "batched EDR readback");
  float ms = 0;
  if (int rc = edrTimer_.elapsed(&ms)) return rc;
  for (int64_t j = 0; j < np; ++j) dist[perm[j]] = outH[j];
  if (st) {
    st->pairs += np;
    st->cells += cells;
    st->jobs += njobs;
    st->kernelMs += ms;
    st->hostMs += (h1 - h0);
  }
  return GAT_OK;
}

int Grid::edrBatchedPairs(const TrajectoryView& queries, std::vector<int32_t>& pairs, std::vector<int32_t>& dist,
                          EdrBatchStats* st) {
  if (int rc = ensureQueryStore(queries)) return rc;
  if (deviceChain) return runBatchedOnDevice(pairs.data(), int64_t(pairs.size() / 2), queries.trajNum, dist, st);
  if (int rc = stage_.begin(runBatchedStaging(int64_t(pairs.size() / 2)), engine_.stream())) return rc;
  return runBatched(pairs, queries.trajNum, dist, st);
}

int Grid::edrBatchedAll(const TrajectoryView& queries, std::vector<int32_t>& dist, EdrBatchStats* st) {
  if (int rc = ensureQueryStore(queries)) return rc;
  const int nq = queries.trajNum;
  if (deviceChain) return runBatchedOnDevice(nullptr, int64_t(nq) * trajNum, nq, dist, st);
  std::vector<int32_t> pairs(size_t(nq) * size_t(trajNum) * 2);
  for (int q = 0; q < nq; ++q)
    for (int t = 0; t < trajNum; ++t) {
      pairs[2 * (size_t(q) * trajNum + t)] = q;
      pairs[2 * (size_t(q) * trajNum + t) + 1] = t;
    }
  if (int rc = stage_.begin(runBatchedStaging(int64_t(pairs.size() / 2)), engine_.stream())) return rc;
  return runBatched(pairs, nq, dist, st);
}

// The top-k merge on the device (gat_edr_topk): each query's survivors,
// ascending by id with their DP distances, go up as one CSR; every other
// candidate takes the all-mismatch value of its length there.
int Grid::deviceTopk(const std::vector<int32_t>& pairs, const std::vector<int32_t>& dist, int nq, int kValue,
                     std::vector<TopK>& out) {
  const size_t count = pairs.size() / 2;
  std::vector<int32_t> sstart(size_t(nq) + 1, 0);
  for (size_t i = 0; i < count; ++i) ++sstart[size_t(pairs[2 * i]) + 1];
  for (int q = 0; q < nq; ++q) sstart[size_t(q) + 1] += sstart[size_t(q)];
  std::vector<std::pair<int32_t, int32_t>> surv(count);  // (id, distance), grouped by query
  {
    std::vector<int32_t> fill(sstart.begin(), sstart.end() - 1);
    for (size_t i = 0; i < count; ++i) surv[size_t(fill[size_t(pairs[2 * i])]++)] = {pairs[2 * i + 1], dist[i]};
  }
  // One upload: [start (nq + 1) | ids (count) | distances (count)]; one
  // readback: [ids (nq k) | distances (nq k) | status].
  const size_t ns = size_t(nq) + 1;
  const size_t cells = size_t(nq) * size_t(kValue);
  int32_t* up = stage_.take_as<int32_t>(ns + 2 * count);
  int32_t* down = stage_.take_as<int32_t>(2 * cells + 1);
  if (!up || !down) {
    gath_set_error("top-k merge: transfer arena too small");
    return GAT_ERR_STATE;
  }
  std::copy(sstart.begin(), sstart.end(), up);
  for (int q = 0; q < nq; ++q) {
    std::sort(surv.begin() + sstart[size_t(q)], surv.begin() + sstart[size_t(q) + 1]);
    for (int32_t i = sstart[size_t(q)]; i < sstart[size_t(q) + 1]; ++i) {
      up[ns + size_t(i)] = surv[size_t(i)].first;
      up[ns + count + size_t(i)] = surv[size_t(i)].second;
    }
  }
  int64_t eligible = 0;
  for (int L = 1; L <= GAT_MAX_TRAJ_LEN; ++L) eligible += int64_t(lengthIds_[L].size());
  const int kk = int(std::min<int64_t>(kValue, eligible));
  void* s = engine_.stream();
  if (int rc = survDev_.reserve((ns + 2 * count) * 4)) return rc;
  GATH_TRY(gat_memcpy_h2d(survDev_.get(), up, (ns + 2 * count) * 4, s), "survivor upload");
  if (int rc = topDev_.reserve((2 * cells + 1) * 4)) return rc;
  const int32_t* sv = survDev_.as<int32_t>();
  int32_t* top = topDev_.as<int32_t>();
  GATH_TRY(gat_edr_topk(ftableDev_.as<int32_t>(), GAT_MAX_TRAJ_LEN + 1, queryLenDev_, nq,
                        trajLen_.as<int32_t>(), trajNum, lenIdsDev_.as<int32_t>(), lenOffDev_.as<int32_t>(), sv,
                        sv + ns, sv + ns + count, kValue, kk, top, top + cells, top + 2 * cells, s),
           "top-k merge");
  GATH_TRY(gat_memcpy_d2h(down, topDev_.get(), (2 * cells + 1) * 4, s), "top-k readback");
  GATH_TRY(gat_stream_sync(s), "top-k readback");
  const int32_t status = down[2 * cells];
  if (status != 0) {
    gath_set_error("top-k merge: inconsistent inputs (status %d)", status);
    return GAT_ERR_HIP;
  }
  const int32_t* ids = down;
  const int32_t* dists = down + cells;
  out.assign(size_t(nq), TopK{});
  for (int q = 0; q < nq; ++q) {
    out[size_t(q)].ids.assign(ids + int64_t(q) * kValue, ids + int64_t(q) * kValue + kk);
    out[size_t(q)].dist.assign(dists + int64_t(q) * kValue, dists + int64_t(q) * kValue + kk);
  }
  return GAT_OK;
}

// The DP over a device-built job list (gat_edr_jobs, sortedDev_ / jobsDev_ /
// the windows in win): one launch per bucket reading its window on the
// device, buckets 3 (several waves per pair) and 2 (one pair per wave) on side
// streams beside the short buckets, as runBatched launches them.
int Grid::launchDeviceBuckets(int64_t cap, int nq, const int32_t* win) {
  void* s = engine_.stream();
  const int split = std::max(1, std::min(edrSplit, GAT_MAX_TRAJ_LEN));
  const int waves = (edrWavesPerPair == 2 || edrWavesPerPair == 4 || edrWavesPerPair == 8) ? edrWavesPerPair : 1;
  auto launch = [&](int bucket, void* stream) -> int {
    const int maxlen = bucket == 0 ? split : bucket == 1 ? std::max(split, 1024) : GAT_MAX_TRAJ_LEN;
    GATH_TRY(gat_edr_batched(sortedDev_.as<int32_t>(), cap, jobsDev_.as<gat_edr_job>(), cap + 4, win + 2 * bucket,
                             dpCounterDev_.as<int32_t>() + bucket, queryXY_.as<float>(), queryOff64Dev_,
                             queryLenDev_, nq, trajXY_.as<float>(), trajOff_.as<int64_t>(),
                             trajLen_.as<int32_t>(), trajNum, maxlen, epsilon,
                             bucket == 0 ? edrPerCuShort : bucket == 1 ? edrPerCuLong : 0,
                             bucket == 3 ? 64 * waves : bucket == 2 ? 64 : 16, outDev_.as<int32_t>(), stream),
             "batched EDR");
    return GAT_OK;
  };
  if (int rc = edrSide_.fork(s)) return rc;
  if (waves > 1)
    if (int rc = launch(3, edrSide_.stream())) return rc;
  if (int rc = edrSide2_.fork(s)) return rc;
  if (int rc = launch(2, edrSide2_.stream())) return rc;
  if (int rc = launch(0, s)) return rc;
  if (int rc = launch(1, s)) return rc;
  if (int rc = edrSide_.join(s)) return rc;
  return edrSide2_.join(s);
}

// Batched EDR over np pairs with the job list built on the device: the pairs
// (pairsHost, or every query x trajectory pair when it is null, generated on
// the device) go through gat_edr_jobs, the bucket launches and gat_edr_unsort,
// and dist comes back in the pairs' order.
int Grid::runBatchedOnDevice(const int32_t* pairsHost, int64_t np, int queryNum, std::vector<int32_t>& dist,
                             EdrBatchStats* st) {
  dist.assign(size_t(np), -1);
  if (np == 0) return GAT_OK;
  if (np > int64_t(INT32_MAX) - 8) {
    gath_set_error("batched EDR: %lld pairs (at most %d per call)", (long long)np, INT32_MAX - 8);
    return GAT_ERR_CAPACITY;
  }
  const double h0 = now_ms();
  if (pairsHost)
    for (int64_t i = 0; i < np; ++i) {
      const int q = pairsHost[2 * i], c = pairsHost[2 * i + 1];
      if (q < 0 || q >= queryNum || c < 0 || c >= trajNum) {
        gath_set_error("batched EDR: pair %lld out of range", (long long)i);
        return GAT_ERR_ARG;
      }
    }
  void* s = engine_.stream();
  const int split = std::max(1, std::min(edrSplit, GAT_MAX_TRAJ_LEN));
  const int waves = (edrWavesPerPair == 2 || edrWavesPerPair == 4 || edrWavesPerPair == 8) ? edrWavesPerPair : 1;
  if (int rc = pairsDev_.reserve(size_t(np) * 8)) return rc;
  if (int rc = sortedDev_.reserve(size_t(np) * 8)) return rc;
  if (int rc = nmDev_.reserve((size_t(np) * 2 + GAT_EDR_JOBS_SCRATCH_EXTRA) * 4)) return rc;
  if (int rc = originDev_.reserve(size_t(np) * 4)) return rc;
  if (int rc = jobsDev_.reserve((size_t(np) + 4) * sizeof(gat_edr_job))) return rc;
  if (int rc = outDev_.reserve(size_t(np) * 4)) return rc;
  if (int rc = unsortDev_.reserve(size_t(np) * 4)) return rc;
  if (int rc = counterDev_.reserve(16)) return rc;
  if (int rc = windowsDev_.reserve(64)) return rc;
  if (int rc = dpCounterDev_.reserve(16)) return rc;
  if (int rc = stage_.begin(size_t(np) * 4 + (pairsHost ? size_t(np) * 8 + 8 : 0) + 3 * 256 + 64, s)) return rc;
  int32_t* win = windowsDev_.as<int32_t>();
  auto* count = counterDev_.as<uint64_t>();
  if (pairsHost) {
    int32_t* up = stage_.take_as<int32_t>(size_t(np) * 2);
    auto* cnt = stage_.take_as<uint64_t>(1);
    if (!up || !cnt) {
      gath_set_error("batched EDR: transfer arena too small");
      return GAT_ERR_STATE;
    }
    std::memcpy(up, pairsHost, size_t(np) * 8);
    *cnt = uint64_t(np);
    GATH_TRY(gat_memcpy_h2d(pairsDev_.get(), up, size_t(np) * 8, s), "pair upload");
    GATH_TRY(gat_memcpy_h2d(count, cnt, 8, s), "pair upload");
  } else {
    GATH_TRY(gat_edr_all_pairs(queryNum, trajNum, pairsDev_.as<int32_t>(), count, s), "all pairs");
  }
  GATH_TRY(gat_edr_jobs(pairsDev_.as<int32_t>(), count, np, queryLenDev_, queryNum, trajLen_.as<int32_t>(),
                        trajNum, split, waves, edrWavesEveryLongPair ? 1 : 0, sortedDev_.as<int32_t>(),
                        originDev_.as<int32_t>(), nmDev_.as<int32_t>(), jobsDev_.as<gat_edr_job>(), win, win + 8,
                        reinterpret_cast<uint64_t*>(win + 10), s),
           "device jobs");
  const double h1 = now_ms();
  if (int rc = edrTimer_.start(s)) return rc;
  if (int rc = launchDeviceBuckets(np, queryNum, win)) return rc;
  if (int rc = edrTimer_.stop(s)) return rc;
  GATH_TRY(gat_edr_unsort(originDev_.as<int32_t>(), outDev_.as<int32_t>(), count, np, unsortDev_.as<int32_t>(), s),
           "distance order");
  int32_t* down = stage_.take_as<int32_t>(size_t(np));
  int32_t* ctl = stage_.take_as<int32_t>(16);
  if (!down || !ctl) {
    gath_set_error("batched EDR: transfer arena too small");
    return GAT_ERR_STATE;
  }
  GATH_TRY(gat_memcpy_d2h(down, unsortDev_.get(), size_t(np) * 4, s), "batched EDR readback");
  GATH_TRY(gat_memcpy_d2h(ctl, windowsDev_.get(), 48, s), "batched EDR readback");
  GATH_TRY(gat_stream_sync(s), "batched EDR readback");
  if (ctl[8] != 0) {
    gath_set_error("batched EDR: device job list status %d", ctl[8]);
    return GAT_ERR_HIP;
  }
  std::memcpy(dist.data(), down, size_t(np) * 4);
  if (st) {
    float ms = 0;
    if (int rc = edrTimer_.elapsed(&ms)) return rc;
    uint64_t cells = 0;
    std::memcpy(&cells, ctl + 10, 8);
    st->pairs += np;
    st->cells += int64_t(cells);
    st->jobs += int64_t(ctl[1]) + ctl[3] + ctl[5] + ctl[7];
    st->kernelMs += ms;
    st->hostMs += (h1 - h0);
  }
  return GAT_OK;
}

// The batch chained on the device: the filter writes the survivors,
// gat_edr_jobs groups them into the DP's length buckets and jobs, the DP
// launches read their job windows on the device, gat_edr_survivor_lists builds
// each query's list and gat_edr_topk merges; one readback at the end carries
// the results and the statuses.  A survivor count above the capacity empties
// every window: the capacity grows and the batch runs again.
int Grid::similarityOnDevice(int nq, int kValue, std::vector<TopK>& out, EdrBatchStats* st, bool* done) {
  *done = false;
  void* s = engine_.stream();
  int64_t eligible = 0;
  for (int L = 1; L <= GAT_MAX_TRAJ_LEN; ++L) eligible += int64_t(lengthIds_[L].size());
  const int kk = int(std::min<int64_t>(kValue, eligible));
  const int split = std::max(1, std::min(edrSplit, GAT_MAX_TRAJ_LEN));
  const int waves = (edrWavesPerPair == 2 || edrWavesPerPair == 4 || edrWavesPerPair == 8) ? edrWavesPerPair : 1;
  const size_t cells = size_t(nq) * size_t(kValue);
  int64_t cap = std::max<int64_t>({deviceTopkCap_, int64_t(nq) * 64, 1024});
  for (int attempt = 0; attempt < 2; ++attempt) {
    const size_t ns = size_t(nq) + 1;
    if (int rc = pairsDev_.reserve(size_t(cap) * 8)) return rc;
    if (int rc = sortedDev_.reserve(size_t(cap) * 8)) return rc;
    if (int rc = nmDev_.reserve((size_t(cap) * 2 + GAT_EDR_JOBS_SCRATCH_EXTRA) * 4)) return rc;
    if (int rc = jobsDev_.reserve((size_t(cap) + 4) * sizeof(gat_edr_job))) return rc;
    if (int rc = outDev_.reserve(size_t(cap) * 4)) return rc;
    if (int rc = survDev_.reserve((ns + 2 * size_t(cap)) * 4)) return rc;
    if (int rc = topDev_.reserve((2 * cells + 1) * 4)) return rc;
    if (int rc = counterDev_.reserve(16)) return rc;
    if (int rc = windowsDev_.reserve(64)) return rc;
    if (int rc = dpCounterDev_.reserve(16)) return rc;
    // windows: [0..7] bucket windows, [8] status, [10..11] cell count.
    int32_t* win = windowsDev_.as<int32_t>();
    auto* cellCount = reinterpret_cast<uint64_t*>(win + 10);
    if (int rc = filterTimer_.start(s)) return rc;
    if (int rc = filterPairs(nq, cap, s)) return rc;
    if (int rc = filterTimer_.stop(s)) return rc;
    GATH_TRY(gat_edr_jobs(pairsDev_.as<int32_t>(), counterDev_.as<uint64_t>(), cap, queryLenDev_, nq,
                          trajLen_.as<int32_t>(), trajNum, split, waves, edrWavesEveryLongPair ? 1 : 0,
                          sortedDev_.as<int32_t>(), nullptr, nmDev_.as<int32_t>(), jobsDev_.as<gat_edr_job>(), win,
                          win + 8, cellCount, s),
             "device jobs");
    if (int rc = edrTimer_.start(s)) return rc;
    if (int rc = launchDeviceBuckets(cap, nq, win)) return rc;
    if (int rc = edrTimer_.stop(s)) return rc;
    if (int rc = mergeTimer_.start(s)) return rc;
    int32_t* sv = survDev_.as<int32_t>();
    GATH_TRY(gat_edr_survivor_lists(sortedDev_.as<int32_t>(), outDev_.as<int32_t>(), counterDev_.as<uint64_t>(), cap,
                                    nq, sv, sv + ns, sv + ns + cap, win + 8, s),
             "survivor lists");
    int32_t* top = topDev_.as<int32_t>();
    GATH_TRY(gat_edr_topk(ftableDev_.as<int32_t>(), GAT_MAX_TRAJ_LEN + 1, queryLenDev_, nq,
                          trajLen_.as<int32_t>(), trajNum, lenIdsDev_.as<int32_t>(), lenOffDev_.as<int32_t>(), sv,
                          sv + ns, sv + ns + cap, kValue, kk, top, top + cells, top + 2 * cells, s),
             "top-k merge");
    if (int rc = mergeTimer_.stop(s)) return rc;
    if (int rc = stage_.begin((2 * cells + 1) * 4 + 2 * 256, s)) return rc;
    int32_t* down = stage_.take_as<int32_t>(2 * cells + 1);
    int32_t* ctl = stage_.take_as<int32_t>(16);
    if (!down || !ctl) {
      gath_set_error("device top-k: transfer arena too small");
      return GAT_ERR_STATE;
    }
    GATH_TRY(gat_memcpy_d2h(down, topDev_.get(), (2 * cells + 1) * 4, s), "top-k readback");
    GATH_TRY(gat_memcpy_d2h(ctl, windowsDev_.get(), 48, s), "top-k readback");
    GATH_TRY(gat_memcpy_d2h(ctl + 12, counterDev_.get(), 8, s), "top-k readback");
    GATH_TRY(gat_stream_sync(s), "top-k readback");
    uint64_t count = 0;
    std::memcpy(&count, ctl + 12, 8);
    const int32_t status = ctl[8];
    if (status & 1) {  // more survivors than the capacity: grow and run again
      if (count > uint64_t(INT32_MAX) - 8) {
        gath_set_error("similarityBatched: %llu candidate pairs pass the filter (at most %d per call; split the "
                       "queries into smaller batches)",
                       (unsigned long long)count, INT32_MAX - 8);
        return GAT_ERR_CAPACITY;
      }
      cap = std::min<int64_t>(int64_t(count) + int64_t(count) / 4 + 1024, INT32_MAX - 8);
      deviceTopkCap_ = cap;
      continue;
    }
    if (status & 2) return GAT_OK;  // a query's list exceeds the device sort: the host-built path
    if (down[2 * cells] != 0) {
      gath_set_error("top-k merge: inconsistent inputs (status %d)", down[2 * cells]);
      return GAT_ERR_HIP;
    }
    out.assign(size_t(nq), TopK{});
    for (int q = 0; q < nq; ++q) {
      out[size_t(q)].ids.assign(down + int64_t(q) * kValue, down + int64_t(q) * kValue + kk);
      out[size_t(q)].dist.assign(down + cells + int64_t(q) * kValue, down + cells + int64_t(q) * kValue + kk);
    }
    if (st) {
      float fms = 0, dms = 0, mms = 0;
      if (int rc = filterTimer_.elapsed(&fms)) return rc;
      if (int rc = edrTimer_.elapsed(&dms)) return rc;
      if (int rc = mergeTimer_.elapsed(&mms)) return rc;
      uint64_t c = 0;
      std::memcpy(&c, ctl + 10, 8);
      st->pairs += int64_t(count);
      st->cells += int64_t(c);
      st->jobs += int64_t(ctl[1]) + ctl[3] + ctl[5] + ctl[7];
      st->pruned += int64_t(nq) * trajNum - int64_t(count);
      st->filterMs += fms;
      st->kernelMs += dms;
      st->mergeMs += mms;
    }
    *done = true;
    return GAT_OK;
  }
  return GAT_OK;  // still over capacity after growing once: the host-built path
}

int Grid::similarityBatched(const TrajectoryView& queries, int kValue, std::vector<TopK>& out, EdrBatchStats* st) {
  if (int rc = checkTrajLengths("similarityBatched")) return rc;
  const double u0 = now_ms();
  if (int rc = ensureQueryStore(queries)) return rc;
  if (st) st->uploadMs += now_ms() - u0;
  const int nq = queries.trajNum;
  if (nq > 65535) {
    gath_set_error("similarityBatched: at most 65535 queries per call");
    return GAT_ERR_ARG;
  }
  const std::vector<int32_t>& F = mismatchTable();
  if (F.empty()) return GAT_ERR_HIP;
  if (deviceChain && kValue >= 1 && kValue <= kDeviceTopkMax && nq <= kDeviceListQueries) {
    bool done = false;
    if (int rc = similarityOnDevice(nq, kValue, out, st, &done)) return rc;
    if (done) return GAT_OK;
  }
  void* s = engine_.stream();
  // Candidate filter on the device; grow the survivor buffer until it fits.
  int64_t cap = std::max<int64_t>(int64_t(nq) * 64, 1024);
  uint64_t count = 0;
  EventTimer ft;
  for (;;) {
    if (int rc = pairsDev_.reserve(size_t(cap) * 8)) return rc;
    if (int rc = counterDev_.reserve(16)) return rc;
    if (int rc = ft.start(s)) return rc;
    if (int rc = filterPairs(nq, cap, s)) return rc;
    if (int rc = ft.stop(s)) return rc;
    GATH_TRY(gat_memcpy_d2h(&count, counterDev_.get(), 8, s), "filter readback");
    GATH_TRY(gat_stream_sync(s), "filter readback");
    if (int64_t(count) <= cap) break;
    // Pair indexes, job offsets and the survivor lists are 32-bit.
    if (count > uint64_t(INT32_MAX)) {
      gath_set_error("similarityBatched: %llu candidate pairs pass the filter (at most %d per call; split the "
                     "queries into smaller batches)",
                     (unsigned long long)count, INT32_MAX);
      return GAT_ERR_CAPACITY;
    }
    cap = std::min<int64_t>(int64_t(count) + int64_t(count) / 4 + 1024, INT32_MAX);
  }
  float fms = 0;
  if (int rc = ft.elapsed(&fms)) return rc;
  // Every transfer of the batch from here passes through the pinned arena.
  const size_t need = size_t(count) * 8 + runBatchedStaging(int64_t(count)) +
                      (size_t(nq) + 1 + 2 * size_t(count)) * 4 + (2 * size_t(nq) * size_t(std::max(kValue, 0)) + 1) * 4 +
                      4 * 256;
  if (int rc = stage_.begin(need, s)) return rc;
  int32_t* pairsH = stage_.take_as<int32_t>(size_t(count) * 2);
  if (!pairsH) {
    gath_set_error("similarityBatched: transfer arena too small");
    return GAT_ERR_STATE;
  }
  GATH_TRY(gat_memcpy_d2h(pairsH, pairsDev_.get(), size_t(count) * 8, s), "survivor readback");
  GATH_TRY(gat_stream_sync(s), "survivor readback");
  std::vector<int32_t> pairs(pairsH, pairsH + size_t(count) * 2);
  std::vector<int32_t> dist;
  if (int rc = runBatched(pairs, nq, dist, st)) return rc;
  const double h0 = now_ms();
  if (kValue >= 1 && kValue <= kDeviceTopkMax) {
    if (int rc = deviceTopk(pairs, dist, nq, kValue, out)) return rc;
    if (st) {
      st->filterMs += fms;
      st->pruned += int64_t(nq) * trajNum - int64_t(count);
      st->mergeMs += now_ms() - h0;
    }
    return GAT_OK;
  }
  // Per query: the survivors with their exact distances, and every other
  // trajectory at the all-mismatch value of its length, merged by (distance, id).
  std::vector<int64_t> start(size_t(nq) + 1, 0);
  for (size_t i = 0; i < count; ++i) ++start[size_t(pairs[2 * i]) + 1];
  for (int q = 0; q < nq; ++q) start[q + 1] += start[q];
  std::vector<std::pair<int32_t, int32_t>> surv(count);  // (dist, id), grouped by query
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (size_t i = 0; i < count; ++i) surv[size_t(fill[pairs[2 * i]]++)] = {dist[i], pairs[2 * i + 1] + 1};
  }
  std::vector<int32_t> lens;  // lengths present in the index
  for (int L = 1; L <= GAT_MAX_TRAJ_LEN; ++L)
    if (!lengthIds_[L].empty()) lens.push_back(L);
  const int w = GAT_MAX_TRAJ_LEN + 1;
  // The length order of every query length of the batch not seen before (the
  // all-mismatch values do not follow the length difference monotonically, so
  // each query length needs its own sort; it is kept for later batches).
  {
    std::vector<int> fresh;
    std::vector<char> seen(size_t(w), 0);
    for (int q = 0; q < nq; ++q) {
      const int a = queryLenHost_[q];
      if (a >= 0 && a < w && !seen[size_t(a)] && lengthOrder_[size_t(a)].empty()) {
        seen[size_t(a)] = 1;
        fresh.push_back(a);
      }
    }
    parallel_for(int(fresh.size()), threads(), [&](int i) {
      const int a = fresh[size_t(i)];
      std::vector<std::pair<int32_t, int32_t>> byF;
      byF.reserve(lens.size());
      for (int L : lens) byF.push_back({F[size_t(std::min(a, L)) * w + std::max(a, L)], L});
      std::sort(byF.begin(), byF.end());
      lengthOrder_[size_t(a)] = std::move(byF);
    });
  }
  out.assign(size_t(nq), TopK{});
  parallel_for(nq, threads(), [&](int q) {
    auto sb = surv.begin() + start[q], se = surv.begin() + start[q + 1];
    std::sort(sb, se);
    std::vector<int32_t> excluded;
    excluded.reserve(size_t(se - sb));
    for (auto it = sb; it != se; ++it) excluded.push_back(it->second);
    std::sort(excluded.begin(), excluded.end());
    const std::vector<std::pair<int32_t, int32_t>>& byF = lengthOrder_[size_t(queryLenHost_[q])];
    const size_t need = size_t(kValue);  // the k best non-survivors suffice
    std::vector<std::pair<int32_t, int32_t>> rest;  // (dist, id), ascending
    rest.reserve(need);
    struct Cursor {
      const int32_t* p;
      const int32_t* e;
    };
    auto later = [](const Cursor& x, const Cursor& y) { return *x.p > *y.p; };
    std::vector<Cursor> heap;
    for (size_t g = 0; g < byF.size() && rest.size() < need;) {
      // One value of F: k-way merge of its lengths' ascending id lists.
      heap.clear();
      size_t h = g;
      for (; h < byF.size() && byF[h].first == byF[g].first; ++h) {
        const auto& ids = lengthIds_[byF[h].second];
        heap.push_back({ids.data(), ids.data() + ids.size()});
      }
      std::make_heap(heap.begin(), heap.end(), later);
      while (!heap.empty() && rest.size() < need) {
        std::pop_heap(heap.begin(), heap.end(), later);
        Cursor& c = heap.back();
        const int32_t id = *c.p++;
        if (!std::binary_search(excluded.begin(), excluded.end(), id)) rest.push_back({byF[g].first, id});
        if (c.p == c.e)
          heap.pop_back();
        else
          std::push_heap(heap.begin(), heap.end(), later);
      }
      g = h;
    }
    std::vector<std::pair<int32_t, int32_t>> merged;
    merged.reserve(size_t(se - sb) + rest.size());
    std::merge(sb, se, rest.begin(), rest.end(), std::back_inserter(merged));
    const size_t kk = std::min(size_t(kValue), merged.size());
    out[q].dist.resize(kk);
    out[q].ids.resize(kk);
    for (size_t i = 0; i < kk; ++i) {
      out[q].dist[i] = merged[i].first;
      out[q].ids[i] = merged[i].second;
    }
  });
  if (st) {
    st->filterMs += fms;
    st->pruned += int64_t(nq) * trajNum - int64_t(count);
    st->hostMs += now_ms() - h0;
  }
  return GAT_OK;
}
