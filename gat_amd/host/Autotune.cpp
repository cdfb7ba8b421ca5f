#include "Autotune.h"

#include <algorithm>

namespace {
double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.0 : v[v.size() / 2];
}
}  // namespace

template <class Index>
int tuneRangeBatched(Index& index, const MBB* boxes, int nq, int reps, TuneRangeResult* out) {
  if (!out || nq < 1 || !boxes || reps < 1) {
    gath_set_error("tuneRangeBatched: bad arguments");
    return GAT_ERR_ARG;
  }
  out->candidates = {4, 8, 16, 32, 64};
  out->ms.clear();
  RangeBatch batch;
  RangeEngine& e = index.engine();
  const int threshold = e.directMaxQueries;
  auto time_it = [&](double* ms) -> int {
    std::vector<double> t;
    for (int r = 0; r < reps + 1; ++r) {
      if (int rc = index.rangeQueryBatchedGPU(boxes, nq, batch, false)) return rc;
      if (r > 0) t.push_back(batch.selectMs + double(batch.kernelMs));  // the first run warms up
    }
    *ms = median(t);
    return GAT_OK;
  };
  double best = -1;
  e.directMaxQueries = 0;  // leaf-major form
  for (int qb : out->candidates) {
    e.queriesPerItem = qb;
    double m = 0;
    if (int rc = time_it(&m)) return rc;
    out->ms.push_back(m);
    if (best < 0 || m < best) {
      best = m;
      out->best = qb;
    }
  }
  e.queriesPerItem = out->best;
  e.directMaxQueries = nq;  // direct form
  if (int rc = time_it(&out->directMs)) return rc;
  e.directMaxQueries = out->directMs < best ? std::max(threshold, nq) : std::min(threshold, nq - 1);
  out->directMaxQueries = e.directMaxQueries;
  return GAT_OK;
}

template int tuneRangeBatched<Grid>(Grid&, const MBB*, int, int, TuneRangeResult*);
template int tuneRangeBatched<STIG>(STIG&, const MBB*, int, int, TuneRangeResult*);
template int tuneRangeBatched<FSG>(FSG&, const MBB*, int, int, TuneRangeResult*);

int tuneEdrBatched(Grid& grid, const TrajectoryView& queries, int reps, TuneEdrResult* out) {
  if (!out || queries.trajNum < 1 || reps < 1) {
    gath_set_error("tuneEdrBatched: bad arguments");
    return GAT_ERR_ARG;
  }
  std::vector<int32_t> dist;
  auto time_it = [&](double* ms) -> int {
    std::vector<double> t;
    for (int r = 0; r < reps + 1; ++r) {
      EdrBatchStats st;
      if (int rc = grid.edrBatchedAll(queries, dist, &st)) return rc;
      if (r > 0) t.push_back(double(st.kernelMs));
    }
    *ms = median(t);
    ++out->evaluations;
    return GAT_OK;
  };
  out->evaluations = 0;
  int split = grid.edrSplit, pcs = grid.edrPerCuShort, pcl = grid.edrPerCuLong;
  double best = 0;
  grid.edrSplit = split;
  if (int rc = time_it(&best)) return rc;
  // One coordinate at a time; each keeps its best value for the next.
  const int splits[] = {256, 512, 1024, GAT_MAX_TRAJ_LEN};
  for (int v : splits) {
    grid.edrSplit = v;
    double ms = 0;
    if (int rc = time_it(&ms)) return rc;
    if (ms < best) {
      best = ms;
      split = v;
    }
  }
  grid.edrSplit = split;
  const int occupancy[] = {0, 6, 9, 12, 16, 24};
  for (int v : occupancy) {
    grid.edrPerCuShort = v;
    double ms = 0;
    if (int rc = time_it(&ms)) return rc;
    if (ms < best) {
      best = ms;
      pcs = v;
    }
  }
  grid.edrPerCuShort = pcs;
  for (int v : occupancy) {
    grid.edrPerCuLong = v;
    double ms = 0;
    if (int rc = time_it(&ms)) return rc;
    if (ms < best) {
      best = ms;
      pcl = v;
    }
  }
  grid.edrPerCuLong = pcl;
  // Waves per long pair, on a batch that fits one round of the device's
  // waves and has a top-k batch's mix of pair shapes: the longest tuning
  // query against 8192 trajectories spread evenly over the database (most
  // pairs of a few strips, one wave each; the longest over several).
  {
    int qbest = 0;
    for (int q = 1; q < queries.trajNum; ++q)
      if (queries.offsets[q + 1] - queries.offsets[q] > queries.offsets[qbest + 1] - queries.offsets[qbest]) qbest = q;
    std::vector<int32_t> pairs;
    const int64_t spread = std::min<int64_t>(8192, grid.trajNum);
    for (int64_t i = 0; i < spread; ++i) {
      pairs.push_back(qbest);
      pairs.push_back(int32_t(spread > 1 ? i * (grid.trajNum - 1) / (spread - 1) : 0));
    }
    int wbest = grid.edrWavesPerPair;
    double sbest = -1;
    for (int w : {1, 2, 4, 8}) {
      grid.edrWavesPerPair = w;
      std::vector<double> t;
      for (int r = 0; r < reps + 1; ++r) {
        EdrBatchStats st;
        if (int rc = grid.edrBatchedPairs(queries, pairs, dist, &st)) return rc;
        if (r > 0) t.push_back(double(st.kernelMs));
      }
      const double ms = median(t);
      ++out->evaluations;
      if (sbest < 0 || ms < sbest) {
        sbest = ms;
        wbest = w;
      }
    }
    grid.edrWavesPerPair = wbest;
    out->wavesPerPair = wbest;
    out->smallMs = sbest;
  }
  out->split = split;
  out->perCuShort = pcs;
  out->perCuLong = pcl;
  out->ms = best;
  return GAT_OK;
}
