#include "FSG.h"

#include <algorithm>
#include <thread>

FSG::FSG(const MBB& mbb, float val_cell_size) {
  range = mbb;
  cell_size = val_cell_size;
  cell_num_x = int((mbb.xmax - mbb.xmin) / val_cell_size) + 1;
  cell_num_y = int((mbb.ymax - mbb.ymin) / val_cell_size) + 1;
  cellnum = int64_t(cell_num_x) * int64_t(cell_num_y);
}

int FSG::threads() const {
  if (hostThreads > 0) return hostThreads;
  unsigned h = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(h ? h : 1u, 16u)));
}

// FSG.cpp:88-93; a point whose row or column is off the grid has no cell.
int64_t FSG::cellOf(float lon, float lat) const {
  const float fr = (range.ymax - lat) / cell_size;
  const float fc = (lon - range.xmin) / cell_size;
  if (!(fr > -1.0f && fr < float(cell_num_y)) || !(fc > -1.0f && fc < float(cell_num_x))) return -1;
  return int64_t(int(fr)) * cell_num_x + int(fc);
}

int FSG::WhichCellPointIn(SamplePoint p) const { return int(cellOf(p.lon, p.lat)); }

int FSG::addDatasetToGrid(Trajectory* db, int traNum) {
  TrajectoryDB flat = TrajectoryDB::fromArray(db, traNum);
  return addDatasetToGrid(TrajectoryView::of(flat), true);
}

int FSG::addDatasetToGrid(const TrajectoryView& db, bool upload) {
  if (cell_num_x <= 0 || cell_num_y <= 0 || cellnum > (int64_t(1) << 30)) {
    gath_set_error("FSG: %d x %d cells is outside the supported range", cell_num_x, cell_num_y);
    return GAT_ERR_ARG;
  }
  trajNum = db.trajNum;
  const int64_t N = db.pointNum();
  totalPointNum = N;
  if (N >= int64_t(UINT32_MAX)) {
    gath_set_error("FSG: %lld points exceed the 32-bit store index", (long long)N);
    return GAT_ERR_ARG;
  }
  if (upload && deviceBuild) {
    const gat_cell_grid g{0, cell_num_x, cell_num_y, range.xmin, range.ymax, cell_size};
    int64_t bad = -1;
    if (int rc = engine_.buildStore(g, db, &cellStart, &bad)) return rc;
    if (bad >= 0) {
      gath_set_error("FSG: point %lld (%g, %g) lies outside the grid", (long long)bad, double(db.lon[bad]),
                     double(db.lat[bad]));
      return GAT_ERR_ARG;
    }
    return uploadCells();
  }
  // Stable bucketing by cell: cells in id order, points of a cell in
  // trajectory order (FSG.cpp:122-141).
  std::vector<uint32_t> cell(static_cast<size_t>(N));
  cellStart.assign(size_t(cellnum) + 1, 0);
  for (int t = 1; t <= db.trajNum; ++t)
    for (int64_t i = db.offsets[t - 1]; i < db.offsets[t]; ++i) {
      const int64_t c = cellOf(db.lon[i], db.lat[i]);
      if (c < 0) {
        gath_set_error("FSG: point %lld (%g, %g) lies outside the grid", (long long)i, double(db.lon[i]),
                       double(db.lat[i]));
        return GAT_ERR_ARG;
      }
      cell[i] = uint32_t(c);
      ++cellStart[c + 1];
    }
  for (int64_t c = 0; c < cellnum; ++c) cellStart[c + 1] += cellStart[c];
  if (!upload) return GAT_OK;
  std::vector<int64_t> cur(cellStart.begin(), cellStart.end() - 1);
  std::vector<float> xy(size_t(N) * 2);
  std::vector<int32_t> tp(size_t(N) * 2);
  for (int t = 1; t <= db.trajNum; ++t) {
    const int64_t b = db.offsets[t - 1];
    for (int64_t i = b; i < db.offsets[t]; ++i) {
      const int64_t d = cur[cell[i]]++;
      xy[2 * d] = db.lon[i];
      xy[2 * d + 1] = db.lat[i];
      tp[2 * d] = t;
      tp[2 * d + 1] = int32_t(i - b);
    }
  }
  if (int rc = engine_.upload(xy.data(), tp.data(), N)) return rc;
  return uploadCells();
}

// The device cell table over the cell offsets, once the point store is resident.
int FSG::uploadCells() {
  std::vector<gat_tree_node> cells(static_cast<size_t>(cellnum));
  for (int64_t c = 0; c < cellnum; ++c) {
    gat_tree_node& d = cells[size_t(c)];
    d = gat_tree_node{{0.0f, 0.0f, 0.0f, 0.0f}, -1, uint32_t(cellStart[c]), uint32_t(cellStart[c + 1] - cellStart[c]), 0u};
  }
  return tree_.uploadGrid(cells, cell_num_x, cell_num_y, range.xmin, range.ymax, cell_size, engine_.stream());
}

int FSG::rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("FSG::rangeQueryBatchedGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  if (tree_.nnodes == 0) {
    gath_set_error("FSG::rangeQueryBatchedGPU: no device grid (build the index with upload)");
    return GAT_ERR_STATE;
  }
  return engine_.runTree(tree_, bounds, rangeNum, out, graph);
}

bool FSG::cellRange(const MBB& q, int* g1, int* g2, int* g3, int* g4) const {
  // FSG.cpp:183-186 (monotone float formulas, so no point inside the
  // rectangle can fall in a cell outside the range), clamped to the grid.
  auto clampi = [](float v, int hi) {
    if (!(v > -1.0f)) return 0;        // also NaN
    if (!(v < float(hi))) return hi - 1;
    return std::min(std::max(int(v), 0), hi - 1);
  };
  *g1 = clampi((q.xmin - range.xmin) / cell_size, cell_num_x);
  *g2 = clampi((q.xmax - range.xmin) / cell_size, cell_num_x);
  *g3 = clampi((range.ymax - q.ymax) / cell_size, cell_num_y);
  *g4 = clampi((range.ymax - q.ymin) / cell_size, cell_num_y);
  return *g1 <= *g2 && *g3 <= *g4;
}

// The reference scans the rectangle's border cells and interior cells alike
// (FSG.cpp:204-312); together they are every cell of rows g3..g4, columns
// g1..g2, and each row's cells are one contiguous store run.
void FSG::addRows(const MBB& q, uint32_t qid, TaskList& list) const {
  int g1, g2, g3, g4;
  if (!cellRange(q, &g1, &g2, &g3, &g4)) return;
  for (int row = g3; row <= g4; ++row) {
    const int64_t a = int64_t(row) * cell_num_x + g1, b = int64_t(row) * cell_num_x + g2;
    list.addRun(uint64_t(cellStart[a]), uint64_t(cellStart[b + 1] - cellStart[a]), qid);
  }
}

void FSG::selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                           std::vector<uint64_t>& candidates) const {
  RangeEngine::select(
      rangeNum, [&](int q, TaskList& list) { addRows(bounds[q], uint32_t(q), list); }, threads(), tasks, candidates);
}

int FSG::rangeQueryBatchGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                            int device_idx) {
  if (device_idx != 0 || rangeNum < 0 || (rangeNum > 0 && (!bounds || !ResultTable))) {
    gath_set_error("FSG::rangeQueryBatchGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  const int stride = trajNum + 1;
  std::vector<uint8_t> bitmap(size_t(rangeNum) * size_t(stride));
  auto selector = [&](int q, TaskList& list) { addRows(bounds[q], uint32_t(q), list); };
  if (int rc = engine_.runBitmap(bounds, rangeNum, selector, threads(), stride, bitmap.data(), false)) return rc;
  for (int q = 0; q < rangeNum; ++q) {
    ResultTable[q].assign(size_t(stride), false);
    int n = 0;
    for (int t = 0; t < stride; ++t)
      if (bitmap[size_t(q) * stride + t]) {
        ResultTable[q][t] = true;
        ++n;
      }
    if (resultSetSize) resultSetSize[q] = n;
  }
  return GAT_OK;
}

int FSG::rangeQueryBatchGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("FSG::rangeQueryBatchGPUPairs: bad arguments");
    return GAT_ERR_ARG;
  }
  auto selector = [&](int q, TaskList& list) { addRows(bounds[q], uint32_t(q), list); };
  return engine_.runPairs(bounds, rangeNum, selector, threads(), out);
}
