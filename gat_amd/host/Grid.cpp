#include "Grid.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace {
int hw_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return int(std::min(h, 16u));
}
}  // namespace

Grid::Grid(const MBB& mbb, float val_cell_size, int /*VITURAL_CELL_PARAM*/, int maxPoints) {
  range = mbb;
  cell_size = val_cell_size;
  maxPointsInNode = maxPoints;
  // Grid.cpp:104-109: the cells needed per axis, rounded up to the power of two
  // strictly above the highest set bit of the larger count.
  const int divX = int((mbb.xmax - mbb.xmin) / val_cell_size) + 1;
  const int divY = int((mbb.ymax - mbb.ymin) / val_cell_size) + 1;
  const int maxValue = std::max(divX, divY);
  const int highBit = maxValue > 0 ? 31 - __builtin_clz(unsigned(maxValue)) : 0;
  cellNum_axis = 1 << (highBit + 1);
  cellnum = cellNum_axis * cellNum_axis;
  totalLevel = highBit + 1;  // log4(cellnum), Grid.cpp:60
}

int Grid::threads() const { return hostThreads > 0 ? hostThreads : hw_threads(); }

// Grid.cpp:202-208: row counted down from ymax, column right from xmin, both
// by float division and truncation.  A point whose row or column falls outside
// the grid has no cell (the reference would index past its cell array).
uint32_t Grid::cellOf(float lon, float lat) const {
  const float fr = (range.ymax - lat) / cell_size;
  const float fc = (lon - range.xmin) / cell_size;
  if (!(fr > -1.0f && fr < float(cellNum_axis)) || !(fc > -1.0f && fc < float(cellNum_axis))) return UINT32_MAX;
  return getIdxFromXY(uint32_t(int(fc)), uint32_t(int(fr)));
}

// Grid.cpp:116-122: cell (row i, column j) spans
// [xmin + cs*j, xmin + cs*(j+1)] x [ymax - cs*(i+1), ymax - cs*i] in float.
MBB Grid::cellMBB(uint32_t cellId) const {
  uint32_t j = 0, i = 0;
  xyFromIdx(cellId, &j, &i);
  const float xl = range.xmin + cell_size * float(int(j));
  const float yl = range.ymax - cell_size * float(int(i) + 1);
  const float xh = range.xmin + cell_size * float(int(j) + 1);
  const float yh = range.ymax - cell_size * float(int(i));
  return MBB(xl, yl, xh, yh);
}

int Grid::addDatasetToGrid(Trajectory* db, int traNum) {
  TrajectoryDB flat = TrajectoryDB::fromArray(db, traNum);
  return addDatasetToGrid(TrajectoryView::of(flat), true);
}

int Grid::addDatasetToGrid(const TrajectoryView& db, bool upload) {
  if (cellNum_axis <= 0 || cellNum_axis > 32768) {
    gath_set_error("Grid: %d cells per axis is outside the supported range", cellNum_axis);
    return GAT_ERR_ARG;
  }
  trajNum = db.trajNum;
  const int64_t N = db.pointNum();
  totalPointNum = N;
  if (N >= int64_t(UINT32_MAX)) {
    gath_set_error("Grid: %lld points exceed the 32-bit store index", (long long)N);
    return GAT_ERR_ARG;
  }
  if (upload && deviceBuild) {
    const gat_cell_grid g{1, cellNum_axis, cellNum_axis, range.xmin, range.ymax, cell_size};
    int64_t bad = -1;
    if (int rc = engine_.buildStore(g, db, &cellStart, &bad)) return rc;
    if (bad >= 0) {
      gath_set_error("Grid: point %lld (%g, %g) lies outside the grid", (long long)bad, double(db.lon[bad]),
                     double(db.lat[bad]));
      return GAT_ERR_ARG;
    }
    return finishBuild(db);
  }
  // Split the trajectories into contiguous groups of about equal point count;
  // each thread owns one group, so the order within a cell stays the
  // reference's (sub-trajectories in trajectory-id order, Grid.cpp:242-261).
  const int want = threads();
  const int T = std::max(1, std::min<int>(want, int(std::max<int64_t>(1, int64_t(1) << 26) / std::max(cellnum, 1))));
  std::vector<int> groupStart(T + 1, 1);
  for (int g = 1; g < T; ++g) {
    const int64_t target = N * g / T;
    groupStart[g] = int(std::upper_bound(db.offsets, db.offsets + db.trajNum + 1, target) - db.offsets);
    groupStart[g] = std::max(groupStart[g], groupStart[g - 1]);
  }
  groupStart[T] = db.trajNum + 1;

  std::vector<uint32_t> cell(static_cast<size_t>(N));
  std::vector<std::vector<uint32_t>> hist(static_cast<size_t>(T));
  std::vector<int64_t> bad(size_t(T), -1);
  auto countPart = [&](int g) {
    auto& h = hist[g];
    h.assign(size_t(cellnum), 0);
    for (int t = groupStart[g]; t < groupStart[g + 1]; ++t)
      for (int64_t i = db.offsets[t - 1]; i < db.offsets[t]; ++i) {
        const uint32_t c = cellOf(db.lon[i], db.lat[i]);
        if (c >= uint32_t(cellnum)) {
          if (bad[g] < 0) bad[g] = i;
          cell[i] = 0;
          continue;
        }
        cell[i] = c;
        ++h[c];
      }
  };
  {
    std::vector<std::thread> pool;
    for (int g = 0; g < T; ++g) pool.emplace_back(countPart, g);
    for (auto& th : pool) th.join();
  }
  for (int g = 0; g < T; ++g)
    if (bad[g] >= 0) {
      gath_set_error("Grid: point %lld (%g, %g) lies outside the grid", (long long)bad[g], double(db.lon[bad[g]]),
                     double(db.lat[bad[g]]));
      return GAT_ERR_ARG;
    }
  cellStart.assign(size_t(cellnum) + 1, 0);
  for (int c = 0; c < cellnum; ++c) {
    int64_t s = 0;
    for (int g = 0; g < T; ++g) s += hist[g][c];
    cellStart[c + 1] = cellStart[c] + s;
  }
  std::vector<float> xy;
  std::vector<int32_t> tp;
  if (upload) {
    xy.resize(size_t(N) * 2);
    tp.resize(size_t(N) * 2);
    auto scatter = [&](int g) {
      // Cursor of group g in each cell: after the cell's start and the points
      // of the groups before it.
      std::vector<int64_t> cur(static_cast<size_t>(cellnum));
      for (int c = 0; c < cellnum; ++c) {
        int64_t at = cellStart[c];
        for (int h = 0; h < g; ++h) at += hist[h][c];
        cur[c] = at;
      }
      for (int t = groupStart[g]; t < groupStart[g + 1]; ++t) {
        const int64_t b = db.offsets[t - 1];
        for (int64_t i = b; i < db.offsets[t]; ++i) {
          const int64_t d = cur[cell[i]]++;
          xy[2 * d] = db.lon[i];
          xy[2 * d + 1] = db.lat[i];
          tp[2 * d] = t;
          tp[2 * d + 1] = int32_t(i - b);
        }
      }
    };
    std::vector<std::thread> pool;
    for (int g = 0; g < T; ++g) pool.emplace_back(scatter, g);
    for (auto& th : pool) th.join();
  }
  hist.clear();
  cell.clear();
  cell.shrink_to_fit();

  if (!upload) return buildHostTree();
  if (int rc = engine_.upload(xy.data(), tp.data(), N)) return rc;
  xy = std::vector<float>();
  tp = std::vector<int32_t>();
  return finishBuild(db);
}

// The quadtree over the cell offsets, its device copy and the trajectory
// store, once the point store is resident.
int Grid::buildHostTree() {
  nodes.clear();
  nodeNum = 1;
  buildNode(0, 0);
  return GAT_OK;
}

int64_t Grid::leafNum() const {
  int64_t n = 0;
  for (const QuadtreeNode& q : nodes) n += q.isLeaf;
  return n;
}

int Grid::finishBuild(const TrajectoryView& db) {
  if (int rc = buildHostTree()) return rc;
  int depth = 0;
  const std::vector<gat_tree_node> tn = treeNodes(&depth);
  if (int rc = tree_.upload(tn, 4, depth, engine_.stream())) return rc;
  int k = 0;
  float block = 0;
  const std::vector<uint32_t> owner = startTable(depth, &k, &block);
  if (!owner.empty())
    if (int rc = tree_.uploadStart(owner, k, range.xmin, range.ymax, block, engine_.stream())) return rc;
  return uploadTrajectories(db);
}

// The device walk starts at the owners of a level-k table, k = min(7, depth -
// 2): on 1M trajectories / 10k rectangles (depth 10) it takes the walk from
// 125k to 67k steps, and the selection from 250 to 229 us with the largest
// rectangles' walks split over several waves (level 8: 257 us, its blocks
// too small for the largest rectangles' chunks).
std::vector<uint32_t> Grid::startTable(int depth, int* level, float* block) const {
  *level = 0;
  *block = 0;
  const int k = std::min(7, depth - 2);
  if (k < 2) return {};
  std::vector<uint32_t> owner = startOwners(k);
  if (owner.empty()) return owner;
  *level = k;
  *block = cell_size * float(1 << (totalLevel - k));
  return owner;
}

// Every block of level k, in Morton order, gets the node that owns it: its
// node at level k, or the leaf above level k that covers it (a node's blocks
// at level k are the Morton range [id * 4^(k-l), (id+1) * 4^(k-l))).  The walk
// may start at the owners only when the outcome cannot differ from the walk
// from the root: every box holds its children's boxes and is wider than zero
// in x; BoxIntersect (MBB.cpp:28-73) is then monotone along every root path,
// so a node is reached exactly when its own box intersects.  An empty table
// means the tree does not qualify.
std::vector<uint32_t> Grid::startOwners(int k) const {
  std::vector<uint32_t> owner;
  if (nodes.empty() || k < 1 || k > totalLevel || k > 15) return owner;
  std::vector<int> order{0};
  for (size_t h = 0; h < order.size(); ++h) {
    const QuadtreeNode& n = nodes[size_t(order[h])];
    if (!n.isLeaf)
      for (int c = 0; c < 4; ++c) order.push_back(n.child[c]);
  }
  if (order.size() >= (size_t(1) << 27)) return owner;
  for (const QuadtreeNode& n : nodes) {
    if (!(n.mbb.xmax > n.mbb.xmin)) return owner;
    if (n.isLeaf) continue;
    for (int c = 0; c < 4; ++c) {
      const MBB& b = nodes[size_t(n.child[c])].mbb;
      if (!(b.xmin >= n.mbb.xmin && b.xmax <= n.mbb.xmax && b.ymin >= n.mbb.ymin && b.ymax <= n.mbb.ymax))
        return owner;
    }
  }
  owner.assign(size_t(1) << (2 * k), UINT32_MAX);
  for (size_t i = 0; i < order.size(); ++i) {
    const QuadtreeNode& n = nodes[size_t(order[i])];
    if (n.level > k || (n.level < k && !n.isLeaf)) continue;
    const uint64_t span = uint64_t(1) << (2 * (k - n.level));
    const uint64_t first = uint64_t(n.NodeID) * span;
    if (first + span > owner.size()) return {};
    std::fill(owner.begin() + int64_t(first), owner.begin() + int64_t(first + span),
              (uint32_t(n.level) << 27) | uint32_t(i));
  }
  for (uint32_t o : owner)
    if (o == UINT32_MAX) return {};
  return owner;
}

// Grid.cpp:58-94.
int Grid::buildNode(int level, int id) {
  const int64_t span = int64_t(1) << (2 * (totalLevel - level));
  const int64_t first = int64_t(id) * span, last = first + span - 1;
  QuadtreeNode n;
  n.NodeID = id;
  n.level = level;
  n.pointStart = cellStart[first];
  n.numPoints = cellStart[last + 1] - cellStart[first];
  const MBB a = cellMBB(uint32_t(first)), b = cellMBB(uint32_t(last));
  n.mbb = MBB(a.xmin, b.ymin, b.xmax, a.ymax);
  n.isLeaf = n.numPoints < maxPointsInNode || level == totalLevel;
  const int me = int(nodes.size());
  nodes.push_back(n);
  if (!n.isLeaf) {
    nodeNum += 4;
    for (int k = 0; k < 4; ++k) {
      const int c = buildNode(level + 1, (id << 2) + k);
      nodes[me].child[k] = c;
    }
  }
  return me;
}

std::vector<gat_tree_node> Grid::treeNodes(int* depth) const {
  std::vector<gat_tree_node> out;
  *depth = 0;
  if (nodes.empty()) return out;
  // Breadth-first order: the four children of a node are appended together,
  // so they are contiguous and keep the UL, UR, DL, DR order.
  std::vector<int> order{0};
  for (size_t h = 0; h < order.size(); ++h) {
    const QuadtreeNode& n = nodes[size_t(order[h])];
    if (!n.isLeaf)
      for (int k = 0; k < 4; ++k) order.push_back(n.child[k]);
  }
  std::vector<int32_t> pos(nodes.size(), -1);
  for (size_t i = 0; i < order.size(); ++i) pos[size_t(order[i])] = int32_t(i);
  out.resize(order.size());
  for (size_t i = 0; i < order.size(); ++i) {
    const QuadtreeNode& n = nodes[size_t(order[i])];
    gat_tree_node& d = out[i];
    d.box[0] = n.mbb.xmin;
    d.box[1] = n.mbb.ymin;
    d.box[2] = n.mbb.xmax;
    d.box[3] = n.mbb.ymax;
    d.first_child = n.isLeaf ? -1 : pos[size_t(n.child[0])];
    d.start = n.isLeaf ? uint32_t(n.pointStart) : 0u;
    d.count = n.isLeaf ? uint32_t(n.numPoints) : 0u;
    d.flags = i == 0 ? GAT_NODE_UNTESTED : 0u;  // the root is taken as it is
    *depth = std::max(*depth, n.level);
  }
  return out;
}

int Grid::rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("Grid::rangeQueryBatchedGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  if (tree_.nnodes == 0) {
    gath_set_error("Grid::rangeQueryBatchedGPU: no device tree (build the index with upload)");
    return GAT_ERR_STATE;
  }
  const int rc = engine_.runTree(tree_, bounds, rangeNum, out, graph);
  lastSelectMs = out.selectMs;
  lastKernelMs = out.kernelMs;
  return rc;
}

// Grid.cpp:428-446, iteratively: the root is taken as it is; every other node
// is visited only when its box intersects the rectangle (BoxIntersect != 0),
// children in UL, UR, DL, DR order.
void Grid::findMatchNodeInQuadTree(const MBB& bound, std::vector<int>* leaves) const {
  if (nodes.empty()) return;
  int stack[4 * 24];
  int top = 0;
  stack[top++] = 0;
  while (top > 0) {
    const QuadtreeNode& n = nodes[stack[--top]];
    if (n.isLeaf) {
      leaves->push_back(int(&n - nodes.data()));
      continue;
    }
    for (int k = 3; k >= 0; --k)
      if (BoxIntersect(bound, nodes[n.child[k]].mbb)) stack[top++] = n.child[k];
  }
}

void Grid::addLeafRuns(const MBB& bound, uint32_t q, TaskList& list) const {
  std::vector<int> leaves;
  findMatchNodeInQuadTree(bound, &leaves);
  for (int li : leaves) list.addRun(uint64_t(nodes[li].pointStart), uint64_t(nodes[li].numPoints), q);
}

void Grid::selectionCost(const MBB& bound, int64_t* points, int64_t* leaves) const {
  TaskList list;
  list.candidates.assign(1, 0);
  addLeafRuns(bound, 0, list);
  *points = int64_t(list.pointsAdded);
  *leaves = int64_t(list.runsAdded);
}

void Grid::selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                            std::vector<uint64_t>& candidates) const {
  RangeEngine::select(
      rangeNum, [&](int q, TaskList& list) { addLeafRuns(bounds[q], uint32_t(q), list); }, threads(), tasks,
      candidates);
}

int Grid::rangeQueryBatchGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                             int device_idx) {
  if (device_idx != 0) {
    gath_set_error("Grid::rangeQueryBatchGPU: only device 0 is supported");
    return GAT_ERR_ARG;
  }
  if (rangeNum < 0 || (rangeNum > 0 && (!bounds || !ResultTable))) {
    gath_set_error("Grid::rangeQueryBatchGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  const int stride = trajNum + 1;
  std::vector<uint8_t> bitmap(size_t(rangeNum) * size_t(stride));
  auto selector = [&](int q, TaskList& list) { addLeafRuns(bounds[q], uint32_t(q), list); };
  if (int rc = engine_.runBitmap(bounds, rangeNum, selector, threads(), stride, bitmap.data(), false)) return rc;
  for (int q = 0; q < rangeNum; ++q) {
    ResultTable[q].assign(size_t(stride), false);
    int n = 0;
    const uint8_t* row = bitmap.data() + size_t(q) * size_t(stride);
    for (int t = 0; t < stride; ++t)
      if (row[t]) {
        ResultTable[q][t] = true;
        ++n;
      }
    if (resultSetSize) resultSetSize[q] = n;
  }
  return GAT_OK;
}

int Grid::rangeQueryBatchGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("Grid::rangeQueryBatchGPUPairs: bad arguments");
    return GAT_ERR_ARG;
  }
  auto selector = [&](int q, TaskList& list) { addLeafRuns(bounds[q], uint32_t(q), list); };
  const int rc = engine_.runPairs(bounds, rangeNum, selector, threads(), out);
  lastSelectMs = out.selectMs;
  lastKernelMs = out.kernelMs;
  return rc;
}
