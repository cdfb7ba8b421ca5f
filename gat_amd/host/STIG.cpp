#include "STIG.h"

#include <algorithm>
#include <limits>
#include <thread>

int STIG::threads() const {
  if (hostThreads > 0) return hostThreads;
  unsigned h = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(h ? h : 1u, 16u)));
}

int STIG::initial(int blk, int totalDim, Trajectory* db, int trajNum) {
  if (totalDim != 2) {
    gath_set_error("STIG: only two dimensions are supported");
    return GAT_ERR_ARG;
  }
  TrajectoryDB flat = TrajectoryDB::fromArray(db, trajNum);
  return initial(blk, TrajectoryView::of(flat), true);
}

// The tree's shape depends only on the node sizes, so it is laid out first and
// the partitioning of disjoint subtrees then runs in parallel.
int STIG::skeleton(int depth, int64_t lo, int64_t hi, bool leaf) {
  Node n;
  n.leaf = leaf;
  n.dim = depth % 2;
  n.lo = lo;
  n.hi = hi;
  const int me = int(nodes.size());
  nodes.push_back(n);
  if (!leaf) {
    const int64_t mid = (lo + hi) / 2;
    const int l = skeleton(depth + 1, lo, mid - 1, mid - lo <= blockSize);
    const int r = skeleton(depth + 1, mid, hi, hi - mid + 1 <= blockSize);
    nodes[me].left = l;
    nodes[me].right = r;
  }
  return me;
}

void STIG::partition(int ni, std::vector<P>& pts, int spawnDepth) {
  Node& n = nodes[ni];
  if (n.leaf) {
    float xl = 999, xh = -999, yl = 999, yh = -999;  // STIG.cpp:77-90
    for (int64_t i = n.lo; i <= n.hi; ++i) {
      xl = std::min(xl, pts[i].x);
      xh = std::max(xh, pts[i].x);
      yl = std::min(yl, pts[i].y);
      yh = std::max(yh, pts[i].y);
    }
    n.box = MBB(xl, yl, xh, yh);
    return;
  }
  const int64_t mid = (n.lo + n.hi) / 2;
  const int dim = n.dim;
  std::nth_element(pts.begin() + n.lo, pts.begin() + mid, pts.begin() + n.hi + 1,
                   [dim](const P& a, const P& b) { return dim == 0 ? a.x < b.x : a.y < b.y; });
  n.median = dim == 0 ? pts[mid].x : pts[mid].y;
  const int l = n.left, r = n.right;
  if (spawnDepth > 0) {
    std::thread th([&, l] { partition(l, pts, spawnDepth - 1); });
    partition(r, pts, spawnDepth - 1);
    th.join();
  } else {
    partition(l, pts, 0);
    partition(r, pts, 0);
  }
}

int STIG::initial(int blk, const TrajectoryView& db, bool upload) {
  if (blk < 2) {
    gath_set_error("STIG: block size must be at least 2");
    return GAT_ERR_ARG;
  }
  blockSize = blk;
  maxTid = db.trajNum;
  totalPointsNum = db.pointNum();
  if (totalPointsNum >= int64_t(UINT32_MAX)) {
    gath_set_error("STIG: %lld points exceed the 32-bit store index", (long long)totalPointsNum);
    return GAT_ERR_ARG;
  }
  std::vector<P> pts(static_cast<size_t>(totalPointsNum));
  for (int t = 1; t <= db.trajNum; ++t)
    for (int64_t i = db.offsets[t - 1]; i < db.offsets[t]; ++i)
      pts[i] = P{db.lon[i], db.lat[i], t, int32_t(i - db.offsets[t - 1])};
  nodes.clear();
  if (totalPointsNum == 0) return GAT_OK;
  skeleton(0, 0, totalPointsNum - 1, totalPointsNum <= blockSize);
  if (upload && deviceBuild) return buildOnDevice(db);
  int spawn = 0;
  for (int t = threads(); t > 1; t >>= 1) ++spawn;
  partition(0, pts, spawn);
  if (!upload) return GAT_OK;
  std::vector<float> xy(size_t(totalPointsNum) * 2);
  std::vector<int32_t> tp(size_t(totalPointsNum) * 2);
  for (int64_t i = 0; i < totalPointsNum; ++i) {
    xy[2 * i] = pts[i].x;
    xy[2 * i + 1] = pts[i].y;
    tp[2 * i] = pts[i].tid;
    tp[2 * i + 1] = pts[i].pid;
  }
  if (int rc = engine_.upload(xy.data(), tp.data(), totalPointsNum)) return rc;
  int depth = 0;
  const std::vector<gat_tree_node> tn = treeNodes(&depth);
  return tree_.upload(tn, 2, depth, engine_.stream());
}

// The skeleton's split nodes depth by depth (each depth ascending in
// position) and its leaves go to the device, which orders the points; the
// medians and leaf boxes come back into the nodes.
int STIG::buildOnDevice(const TrajectoryView& db) {
  RangeEngine::KdSkeleton k;
  std::vector<int> splitNode, leafNode;
  std::vector<int> level{0};
  k.depthFirst.push_back(0);
  while (!level.empty()) {
    std::vector<int> next;
    std::sort(level.begin(), level.end(), [&](int a, int b) { return nodes[a].lo < nodes[b].lo; });
    for (int ni : level) {
      const Node& n = nodes[size_t(ni)];
      if (n.leaf) {
        leafNode.push_back(ni);
        k.leafLo.push_back(n.lo);
        k.leafHi.push_back(n.hi);
        continue;
      }
      splitNode.push_back(ni);
      k.splitLo.push_back(n.lo);
      k.splitHi.push_back(n.hi);
      next.push_back(n.left);
      next.push_back(n.right);
    }
    if (next.empty()) break;
    k.depthFirst.push_back(int64_t(k.splitLo.size()));
    level.swap(next);
  }
  if (k.depthFirst.back() != int64_t(k.splitLo.size())) k.depthFirst.push_back(int64_t(k.splitLo.size()));
  std::vector<float> med, boxes;
  if (int rc = engine_.buildKdStore(db, k, &med, &boxes)) return rc;
  for (size_t i = 0; i < splitNode.size(); ++i) nodes[size_t(splitNode[i])].median = med[i];
  for (size_t i = 0; i < leafNode.size(); ++i)
    nodes[size_t(leafNode[i])].box = MBB(boxes[4 * i], boxes[4 * i + 1], boxes[4 * i + 2], boxes[4 * i + 3]);
  int depth = 0;
  const std::vector<gat_tree_node> tn = treeNodes(&depth);
  return tree_.upload(tn, 2, depth, engine_.stream());
}

std::vector<gat_tree_node> STIG::treeNodes(int* depth) const {
  std::vector<gat_tree_node> out;
  *depth = 0;
  if (nodes.empty()) return out;
  const float inf = std::numeric_limits<float>::infinity();
  struct Entry {
    int node, level;
    gat_tree_node d;
  };
  auto leaf_entry = [&](const Node& n, gat_tree_node& d) {
    d.box[0] = n.box.xmin;
    d.box[1] = n.box.ymin;
    d.box[2] = n.box.xmax;
    d.box[3] = n.box.ymax;
    d.first_child = -1;
    d.start = uint32_t(n.lo);
    d.count = n.hi >= n.lo ? uint32_t(n.hi - n.lo + 1) : 0u;
    d.flags = GAT_NODE_INCLUSIVE;
  };
  // A leaf's box test implies every split test above it (the left part of a
  // split holds values <= the median, the right part values >= it), so a
  // leaf is tested against its box alone.
  std::vector<Entry> q;
  Entry root{0, 0, {}};
  if (nodes[0].leaf) {
    leaf_entry(nodes[0], root.d);
  } else {
    root.d = gat_tree_node{{-inf, -inf, inf, inf}, 0, 0, 0, GAT_NODE_UNTESTED};
  }
  q.push_back(root);
  for (size_t h = 0; h < q.size(); ++h) {
    const Node& n = nodes[size_t(q[h].node)];
    *depth = std::max(*depth, q[h].level);
    if (n.leaf) continue;
    q[h].d.first_child = int32_t(q.size());
    const int kids[2] = {n.left, n.right};
    for (int k = 0; k < 2; ++k) {
      const Node& c = nodes[size_t(kids[k])];
      Entry e{kids[k], q[h].level + 1, {}};
      if (c.leaf) {
        leaf_entry(c, e.d);
      } else {
        // searchTree (STIG.cpp:179-252): left when qmin <= median, right when qmax >= median.
        float b[4] = {-inf, -inf, inf, inf};
        b[n.dim + (k == 0 ? 2 : 0)] = n.median;
        e.d = gat_tree_node{{b[0], b[1], b[2], b[3]}, 0, 0, 0, GAT_NODE_INCLUSIVE};
      }
      q.push_back(e);
    }
  }
  out.reserve(q.size());
  for (const Entry& e : q) out.push_back(e.d);
  return out;
}

int STIG::rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("STIG::rangeQueryBatchedGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  if (tree_.nnodes == 0) {
    gath_set_error("STIG::rangeQueryBatchedGPU: no device tree (build the index with upload)");
    return GAT_ERR_STATE;
  }
  return engine_.runTree(tree_, bounds, rangeNum, out, graph);
}

int STIG::searchTree(const MBB& q, std::vector<STIGBlock>* out) const {
  if (nodes.empty()) return 0;
  int stack[128];
  int top = 0;
  stack[top++] = 0;
  while (top > 0) {
    const Node& n = nodes[stack[--top]];
    if (n.leaf) {
      if (n.lo <= n.hi && n.box.xmin <= q.xmax && n.box.xmax >= q.xmin && n.box.ymin <= q.ymax &&
          n.box.ymax >= q.ymin)
        out->push_back(STIGBlock{n.lo, n.hi});
      continue;
    }
    const float qmin = n.dim == 0 ? q.xmin : q.ymin;
    const float qmax = n.dim == 0 ? q.xmax : q.ymax;
    // Push right first so the left subtree is visited first (store order).
    if (qmax >= n.median) stack[top++] = n.right;
    if (qmin <= n.median) stack[top++] = n.left;
  }
  return 0;
}

void STIG::addLeaves(const MBB& q, uint32_t qid, TaskList& list) const {
  std::vector<STIGBlock> blocks;
  searchTree(q, &blocks);
  for (const STIGBlock& b : blocks) list.addRun(uint64_t(b.startIdx), uint64_t(b.endIdx - b.startIdx + 1), qid);
}

void STIG::selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                            std::vector<uint64_t>& candidates) const {
  RangeEngine::select(
      rangeNum, [&](int q, TaskList& list) { addLeaves(bounds[q], uint32_t(q), list); }, threads(), tasks,
      candidates);
}

int STIG::rangeQueryGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                        int device_idx) {
  if (device_idx != 0 || rangeNum < 0 || (rangeNum > 0 && (!bounds || !ResultTable))) {
    gath_set_error("STIG::rangeQueryGPU: bad arguments");
    return GAT_ERR_ARG;
  }
  const int stride = maxTid + 1;
  std::vector<uint8_t> bitmap(size_t(rangeNum) * size_t(stride));
  auto selector = [&](int q, TaskList& list) { addLeaves(bounds[q], uint32_t(q), list); };
  if (int rc = engine_.runBitmap(bounds, rangeNum, selector, threads(), stride, bitmap.data(), true)) return rc;
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

int STIG::rangeQueryGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out) {
  if (rangeNum < 0 || (rangeNum > 0 && !bounds)) {
    gath_set_error("STIG::rangeQueryGPUPairs: bad arguments");
    return GAT_ERR_ARG;
  }
  auto selector = [&](int q, TaskList& list) { addLeaves(bounds[q], uint32_t(q), list); };
  return engine_.runPairs(bounds, rangeNum, selector, threads(), out);
}
