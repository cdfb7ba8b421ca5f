// TEST INFRASTRUCTURE ONLY — the parity oracle.
//
// A CPU restatement of azh18/GAT's trajectory-query path (range scan over the
// GT-quadtree / FSG / STIG indexes, and the EDR recurrence behind top-k
// similarity).  Every function cites the reference file:line it follows.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker or as the timed CPU baseline; the
// product path (gat_amd) never links, loads or calls it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, as the reference's
// Makefile:17 compiles its host code with g++ -O2; -ffp-contract=off keeps every
// float operation individually rounded, as the reference's x86-64 SSE code is).
//
// Deliberate departures from the reference, each forced by the parity contract:
//  * STIG is a correct kd-split (the reference's sort range drops one element,
//    STIG.cpp:101/135, and duplicates the median, STIG.cpp:107-118), its search
//    uses unbounded half-planes instead of the hard-coded [0,180]
//    (STIG.cpp:206-230), and its leaf test is an inclusive box overlap instead
//    of the strict BoxIntersect (STIG.cpp:243), so its answer equals the
//    inclusive brute-force set.
//  * FSG clamps the rectangle's cell range to the grid (FSG.cpp:183-190 does not).
//  * eps is a run-time double instead of the EPSILON macro (ConstDefine.h:28).
//  * top-k is exact brute force ordered by (distance, trajectory id).
//  * The reference's own pruned CPU top-k (FVTable + SimilarityMultiThreadHandler
//    + SimilarityQueryBatchCPUParallel) is restated too, as the timed baseline
//    the reference runs (SystemTest.cpp:207); it is not exact and is never used
//    as the parity answer.  Two points of it are undefined behaviour in C++ and
//    are restated as the reference's x86-64 build evaluates them or as its
//    evident intent (see fv_neighbors and fd_value).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <thread>
#include <vector>

namespace {

struct Box {
  float xmin, ymin, xmax, ymax;
};

inline Box box_from(const float* b) { return Box{b[0], b[1], b[2], b[3]}; }

// MBB::pInBox — MBB.cpp:21-26 (inclusive on all four edges).
inline bool pin_box(const Box& b, float x, float y) {
  return x <= b.xmax && x >= b.xmin && y <= b.ymax && y >= b.ymin;
}

// BoxIntersect — MBB.cpp:28-73.  0 = disjoint or only touching, 1 = overlap,
// 2 = first contains second, 3 = second contains first.
int box_intersect(const Box& first, const Box& second) {
  // Order the pair so that `lo` starts further left (ties: the wider one).
  const Box* lo;
  const Box* hi;
  bool swapped;
  if (first.xmin < second.xmin) {
    lo = &first; hi = &second; swapped = false;
  } else if (first.xmin == second.xmin && first.xmax > second.xmax) {
    lo = &first; hi = &second; swapped = false;
  } else {
    lo = &second; hi = &first; swapped = true;
  }
  if (hi->xmin >= lo->xmax) return 0;
  if (hi->ymax <= lo->ymin) return 0;
  if (hi->ymin >= lo->ymax) return 0;
  bool contains = pin_box(*lo, hi->xmin, hi->ymin) && pin_box(*lo, hi->xmin, hi->ymax) &&
                  pin_box(*lo, hi->xmax, hi->ymin) && pin_box(*lo, hi->xmax, hi->ymax);
  if (!contains) return 1;
  return swapped ? 3 : 2;
}

// getIdxFromXY — Grid.cpp:25-54.  Bit i (1-based) of the result takes the next
// bit of x when i is odd and the next bit of y when i is even, for
// i = 1 .. 2*max(bitlen(x), bitlen(y)) (bitlen(0) = 1).
int morton(int x, int y) {
  int lenx = (x == 0) ? 1 : int(std::log2(x)) + 1;
  int leny = (y == 0) ? 1 : int(std::log2(y)) + 1;
  int total = 2 * std::max(lenx, leny);
  int result = 0, xb = 0, yb = 0;
  for (int i = 1; i <= total; ++i) {
    if (i & 1) {
      result += ((x >> xb) & 1) * (1 << (i - 1));
      ++xb;
    } else {
      result += ((y >> yb) & 1) * (1 << (i - 1));
      ++yb;
    }
  }
  return result;
}

uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Order-independent digest of a (tid, pid) set: the wrapping sum of mix64.
inline uint64_t pair_digest(int32_t tid, int32_t pid) {
  return mix64((uint64_t(uint32_t(tid)) << 32) | uint32_t(pid));
}

// A cell-ordered point store shared by the Grid and FSG restatements.
// Sub-trajectories (maximal runs of consecutive points in one cell) are appended
// to their cell's list in trajectory-id order (Grid.cpp:128-199,
// Cell.cpp:36-50); cells are then laid out in cell-id order, runs in list order,
// points in trajectory order (Grid.cpp:242-261, FSG.cpp:122-141).
struct Run {
  int32_t tid;
  int32_t first;  // index of the run's first point within its trajectory
  int32_t count;
};

struct CellStore {
  std::vector<std::vector<Run>> runs;  // per cell
  std::vector<int64_t> start;          // per cell: first index in the store
  std::vector<int64_t> count;          // per cell: number of points
  std::vector<float> x, y;
  std::vector<int32_t> tid, pid;

  template <class CellOf>
  void build(int ncell, const int64_t* offs, int ntraj, const float* lon, const float* lat,
             CellOf cell_of) {
    runs.assign(ncell, {});
    for (int t = 1; t <= ntraj; ++t) {
      int64_t b = offs[t - 1], e = offs[t];
      if (e <= b) continue;
      int last = cell_of(lon[b], lat[b]);
      int32_t run_first = 0;
      for (int64_t i = b + 1; i < e; ++i) {
        int now = cell_of(lon[i], lat[i]);
        if (now != last) {
          runs[last].push_back(Run{t, run_first, int32_t(i - b) - run_first});
          last = now;
          run_first = int32_t(i - b);
        }
      }
      runs[last].push_back(Run{t, run_first, int32_t(e - b) - run_first});
    }
    start.assign(ncell, 0);
    count.assign(ncell, 0);
    int64_t total = 0;
    for (int c = 0; c < ncell; ++c)
      for (const Run& r : runs[c]) total += r.count;
    x.resize(total); y.resize(total); tid.resize(total); pid.resize(total);
    int64_t at = 0;
    for (int c = 0; c < ncell; ++c) {
      start[c] = at;
      for (const Run& r : runs[c]) {
        int64_t base = offs[r.tid - 1] + r.first;
        for (int k = 0; k < r.count; ++k) {
          x[at] = lon[base + k];
          y[at] = lat[base + k];
          tid[at] = r.tid;
          pid[at] = r.first + k;
          ++at;
        }
      }
      count[c] = at - start[c];
    }
  }
};

// Per-query answer accumulator: either the trajectory bitmap of the reference's
// CPURangeQueryResult (QueryResult.h:4), or (tid, pid) pairs, or a count+digest.
struct Sink {
  uint8_t* bitmap = nullptr;   // nq * stride bytes
  int64_t stride = 0;
  int64_t* counts = nullptr;   // nq
  uint64_t* digests = nullptr; // nq
  std::vector<std::vector<int32_t>>* pairs = nullptr;
  void hit(int q, int32_t t, int32_t p) {
    if (bitmap) bitmap[int64_t(q) * stride + t] = 1;
    if (counts) counts[q] += 1;
    if (digests) digests[q] += pair_digest(t, p);
    if (pairs) {
      (*pairs)[q].push_back(t);
      (*pairs)[q].push_back(p);
    }
  }
};

struct GridIndex {
  Box range{};
  float cell_size = 0;
  int axis = 0, cellnum = 0, total_level = 0, max_pts = 0;
  std::vector<Box> cell_box;  // by Morton id
  CellStore store;
  struct Node {
    int id = 0, level = 0;
    int64_t npts = 0;
    bool leaf = true;
    int child[4] = {-1, -1, -1, -1};  // UL, UR, DL, DR
    Box mbb{};
  };
  std::vector<Node> nodes;

  // Grid::Grid — Grid.cpp:97-124.
  void init(const Box& r, float cs, int maxp) {
    range = r;
    cell_size = cs;
    max_pts = maxp;
    int divx = int((r.xmax - r.xmin) / cs) + 1;
    int divy = int((r.ymax - r.ymin) / cs) + 1;
    int mv = std::max(divx, divy);
    int lg = int(std::log2(mv));
    axis = (mv >> lg) << (lg + 1);
    cellnum = axis * axis;
    total_level = int(std::log2(double(cellnum)) / std::log2(4.0));
    cell_box.assign(cellnum, Box{});
    for (int i = 0; i < axis; ++i)
      for (int j = 0; j < axis; ++j) {
        float xl = range.xmin + cell_size * float(j);
        float yl = range.ymax - cell_size * float(i + 1);
        float xh = range.xmin + cell_size * float(j + 1);
        float yh = range.ymax - cell_size * float(i);
        cell_box[morton(j, i)] = Box{xl, yl, xh, yh};
      }
  }

  // Grid::WhichCellPointIn — Grid.cpp:202-208.
  int cell_of(float lon, float lat) const {
    int row = int((range.ymax - lat) / cell_size);
    int col = int((lon - range.xmin) / cell_size);
    return morton(col, row);
  }

  // Grid::buildQuadTree — Grid.cpp:58-94.
  int build_node(int level, int id) {
    int span = 1;
    for (int k = 0; k < total_level - level; ++k) span *= 4;
    int first = id * span, last = (id + 1) * span - 1;
    int64_t pts = 0;
    for (int c = first; c <= last; ++c) pts += store.count[c];
    Node n;
    n.id = id;
    n.level = level;
    n.npts = pts;
    n.mbb = Box{cell_box[first].xmin, cell_box[last].ymin, cell_box[last].xmax, cell_box[first].ymax};
    n.leaf = (pts < max_pts) || (level == total_level);
    int me = int(nodes.size());
    nodes.push_back(n);
    if (!n.leaf)
      for (int k = 0; k < 4; ++k) {
        int c = build_node(level + 1, (id << 2) + k);
        nodes[me].child[k] = c;
      }
    return me;
  }

  // Grid::addDatasetToGrid — Grid.cpp:210-316 (store and quadtree; the
  // frequency vectors and cellBasedTrajectory only serve the FD pruning and EDR
  // reassembly, which the oracle does not need).
  void add(const int64_t* offs, int ntraj, const float* lon, const float* lat) {
    store.build(cellnum, offs, ntraj, lon, lat, [this](float a, float b) { return cell_of(a, b); });
    nodes.clear();
    build_node(0, 0);
  }

  // Grid::findMatchNodeInQuadTree — Grid.cpp:428-446.
  void match(int ni, const Box& q, std::vector<int>& out) const {
    const Node& n = nodes[ni];
    if (n.leaf) {
      out.push_back(ni);
      return;
    }
    for (int k = 0; k < 4; ++k)
      if (box_intersect(q, nodes[n.child[k]].mbb)) match(n.child[k], q, out);
  }

  int64_t leaf_start(const Node& n) const {
    int span = 1;
    for (int k = 0; k < total_level - n.level; ++k) span *= 4;
    return store.start[n.id * span];
  }

  // Grid::rangeQueryBatch — Grid.cpp:341-405 (one query).
  void query(int qi, const Box& q, Sink& sink) const {
    std::vector<int> leaves;
    match(0, q, leaves);
    for (int li : leaves) {
      const Node& n = nodes[li];
      int span = 1;
      for (int k = 0; k < total_level - n.level; ++k) span *= 4;
      int first = n.id * span, last = (n.id + 1) * span - 1;
      for (int c = first; c <= last; ++c) {
        int64_t b = store.start[c], e = b + store.count[c];
        for (int64_t i = b; i < e; ++i)
          if (pin_box(q, store.x[i], store.y[i])) sink.hit(qi, store.tid[i], store.pid[i]);
      }
    }
  }
};

// MortonGrid — MortonGrid.cpp:8-13 (leaf flag index), 23-49 (geometry),
// 125-131 (cell of a point), 134-214 (store), 238-268 (leaf flags, node
// boxes), 270-335 (rangeQueryBatch), 355-393 (findMatchNodeInQuadTree).  The
// tree is only a flag per (level, id); the reference's flag set is a
// std::bitset of (4^10 - 1) / 3 entries (MortonGrid.h:17), which a grid of
// 2^10 cells per axis (leaves at level 10) overruns, so here it is sized by
// the grid.
struct MortonIndex {
  Box range{};
  float cell_size = 0;
  int axis = 0, cellnum = 0, total_level = 0, max_pts = 0;
  std::vector<Box> cell_box;  // cellPtr[c].mbb
  CellStore store;
  std::vector<bool> is_leaf;

  static int bitmap_index(int level, int id) {  // findIdxInBitmap
    if (level <= 0) return 0;
    return (int(std::pow(4, level) + 0.5) - 1) / 3 + id;
  }

  void init(const Box& r, float cs, int maxp) {
    range = r;
    cell_size = cs;
    max_pts = maxp;
    int divx = int((r.xmax - r.xmin) / cs) + 1;
    int divy = int((r.ymax - r.ymin) / cs) + 1;
    int mv = std::max(divx, divy);
    axis = mv >> int(std::log2(mv)) << (int(std::log2(mv)) + 1);
    cellnum = axis * axis;
    total_level = int(std::log2(cellnum) / std::log2(4));
    cell_box.assign(cellnum, Box{});
    for (int i = 0; i <= axis - 1; i++)
      for (int j = 0; j <= axis - 1; j++)
        cell_box[morton(j, i)] = Box{range.xmin + cell_size * j, range.ymax - cell_size * (i + 1),
                                     range.xmin + cell_size * (j + 1), range.ymax - cell_size * i};
  }

  int cell_of(float lon, float lat) const {
    int row = int((range.ymax - lat) / cell_size);
    int col = int((lon - range.xmin) / cell_size);
    return morton(col, row);
  }

  void build_tree(int level, int id) {
    int span = int(std::pow(4, total_level - level));
    int64_t pts = 0;
    for (int i = id * span; i <= (id + 1) * span - 1; i++) pts += store.count[i];
    if (pts < max_pts || level == total_level) {
      is_leaf[bitmap_index(level, id)] = true;
      return;
    }
    for (int k = 0; k < 4; ++k) build_tree(level + 1, (id << 2) + k);
  }

  Box node_box(int level, int id) const {  // generateMBBfromNode
    int span = int(std::pow(4, total_level - level));
    const Box& first = cell_box[id * span];
    const Box& last = cell_box[(id + 1) * span - 1];
    return Box{first.xmin, last.ymin, last.xmax, first.ymax};
  }

  void add(const int64_t* offs, int ntraj, const float* lon, const float* lat) {
    store.build(cellnum, offs, ntraj, lon, lat, [this](float a, float b) { return cell_of(a, b); });
    is_leaf.assign(size_t(bitmap_index(total_level + 1, 0)), false);
    build_tree(0, 0);
  }

  void match(int level, int id, const Box& q, std::vector<std::pair<int, int>>& out) const {
    if (is_leaf[bitmap_index(level, id)]) {
      out.push_back({level, id});
      return;
    }
    for (int k = 0; k < 4; ++k)
      if (box_intersect(q, node_box(level + 1, (id << 2) + k))) match(level + 1, (id << 2) + k, q, out);
  }

  void query(int qi, const Box& q, Sink& sink) const {
    std::vector<std::pair<int, int>> nodes;
    match(0, 0, q, nodes);
    for (const auto& n : nodes) {
      int span = int(std::pow(4, total_level - n.first) + 0.5);
      for (int c = n.second * span; c <= (n.second + 1) * span - 1; c++) {
        int64_t b = store.start[c], e = b + store.count[c];
        for (int64_t i = b; i < e; ++i)
          if (pin_box(q, store.x[i], store.y[i])) sink.hit(qi, store.tid[i], store.pid[i]);
      }
    }
  }
};

// FSG — FSG.cpp:7-28 (geometry), 88-93 (cell of a point), 95-143 (store),
// 183-261 (candidate cells of a rectangle).
struct FsgIndex {
  Box range{};
  float cell_size = 0;
  int nx = 0, ny = 0;
  CellStore store;

  void init(const Box& r, float cs) {
    range = r;
    cell_size = cs;
    nx = int((r.xmax - r.xmin) / cs) + 1;
    ny = int((r.ymax - r.ymin) / cs) + 1;
  }
  int cell_of(float lon, float lat) const {
    int row = int((range.ymax - lat) / cell_size);
    int col = int((lon - range.xmin) / cell_size);
    return row * nx + col;
  }
  void add(const int64_t* offs, int ntraj, const float* lon, const float* lat) {
    store.build(nx * ny, offs, ntraj, lon, lat, [this](float a, float b) { return cell_of(a, b); });
  }
  // The reference enumerates the rectangle's border cells, then its interior
  // cells, and scans both (the direct-result shortcut is commented out,
  // FSG.cpp:313-322).  Column and row ranges are clamped to the grid here.
  void candidate_cells(const Box& q, std::vector<int>& cells) const {
    int g1 = int((q.xmin - range.xmin) / cell_size);
    int g2 = int((q.xmax - range.xmin) / cell_size);
    int g3 = int((range.ymax - q.ymax) / cell_size);
    int g4 = int((range.ymax - q.ymin) / cell_size);
    g1 = std::min(std::max(g1, 0), nx - 1);
    g2 = std::min(std::max(g2, 0), nx - 1);
    g3 = std::min(std::max(g3, 0), ny - 1);
    g4 = std::min(std::max(g4, 0), ny - 1);
    if (g2 < g1 || g4 < g3) return;
    int m = nx;
    int a = g1 + g3 * m, b = g2 + g3 * m, c = g1 + g4 * m, d = g2 + g4 * m;
    for (int i = a; i <= b; ++i) cells.push_back(i);                        // top border row
    if (g4 - g3 >= 1)
      for (int i = c; i <= d; ++i) cells.push_back(i);                      // bottom border row
    if (g4 - g3 >= 2) {
      for (int i = a + m; i <= a + (g4 - g3 - 1) * m; i += m) cells.push_back(i);  // left column
      if (g2 > g1)
        for (int i = b + m; i <= b + (g4 - g3 - 1) * m; i += m) cells.push_back(i);  // right column
    }
    if (b >= a + 2 && c >= a + 2 * m)                                        // interior
      for (int i = a + 1; i <= b - 1; ++i)
        for (int j = 1; j <= g4 - g3 - 1; ++j) cells.push_back(i + j * m);
  }
  void query(int qi, const Box& q, Sink& sink) const {
    std::vector<int> cells;
    candidate_cells(q, cells);
    for (int c : cells) {
      int64_t b = store.start[c], e = b + store.count[c];
      for (int64_t i = b; i < e; ++i)
        if (pin_box(q, store.x[i], store.y[i])) sink.hit(qi, store.tid[i], store.pid[i]);
    }
  }
};

// STIG — STIG.cpp:35-66 (all points in trajectory order), 96-168 (alternating
// x/y median split down to leaves of at most blockSize points), 71-93 (leaf
// box from its points), 179-252 (search) — corrected as stated in the header.
struct StigIndex {
  struct P {
    float x, y;
    int32_t tid, pid;
  };
  int block = 0;
  std::vector<P> pts;
  struct Node {
    bool leaf = true;
    int dim = 0;
    float median = 0;
    int left = -1, right = -1;
    int64_t lo = 0, hi = 0;  // leaf: [lo, hi]
    Box box{};
  };
  std::vector<Node> nodes;

  int make_leaf(int64_t lo, int64_t hi) {
    Node n;
    n.leaf = true;
    n.lo = lo;
    n.hi = hi;
    float xl = 999, xh = -999, yl = 999, yh = -999;
    for (int64_t i = lo; i <= hi; ++i) {
      xl = std::min(xl, pts[i].x); xh = std::max(xh, pts[i].x);
      yl = std::min(yl, pts[i].y); yh = std::max(yh, pts[i].y);
    }
    n.box = Box{xl, yl, xh, yh};
    nodes.push_back(n);
    return int(nodes.size()) - 1;
  }

  // The reference sorts the node's points on the split axis and takes the
  // element at (start+end)/2 as the median (STIG.cpp:99-105); a selection of
  // that element gives the same partition property: every point left of it has
  // a key <= the median and every point from it on has a key >= the median.
  int make_inner(int depth, int64_t lo, int64_t hi) {
    int dim = depth % 2;
    int64_t mid = (lo + hi) / 2;
    auto by_key = [dim](const P& a, const P& b) { return dim == 0 ? a.x < b.x : a.y < b.y; };
    std::nth_element(pts.begin() + lo, pts.begin() + mid, pts.begin() + hi + 1, by_key);
    Node n;
    n.leaf = false;
    n.dim = dim;
    n.median = dim == 0 ? pts[mid].x : pts[mid].y;
    int me = int(nodes.size());
    nodes.push_back(n);
    int l = (mid - lo <= block) ? make_leaf(lo, mid - 1) : make_inner(depth + 1, lo, mid - 1);
    int r = (hi - mid + 1 <= block) ? make_leaf(mid, hi) : make_inner(depth + 1, mid, hi);
    nodes[me].left = l;
    nodes[me].right = r;
    return me;
  }

  void add(const int64_t* offs, int ntraj, const float* lon, const float* lat, int blk) {
    block = blk;
    int64_t total = offs[ntraj];
    pts.resize(total);
    for (int t = 1; t <= ntraj; ++t)
      for (int64_t i = offs[t - 1]; i < offs[t]; ++i) pts[i] = P{lon[i], lat[i], t, int32_t(i - offs[t - 1])};
    nodes.clear();
    if (total == 0) return;
    if (total <= block) make_leaf(0, total - 1);
    else make_inner(0, 0, total - 1);
  }

  void search(int ni, const Box& q, std::vector<int>& out) const {
    const Node& n = nodes[ni];
    if (n.leaf) {
      if (n.box.xmin <= q.xmax && n.box.xmax >= q.xmin && n.box.ymin <= q.ymax && n.box.ymax >= q.ymin)
        out.push_back(ni);
      return;
    }
    float qmin = n.dim == 0 ? q.xmin : q.ymin;
    float qmax = n.dim == 0 ? q.xmax : q.ymax;
    if (qmin <= n.median) search(n.left, q, out);
    if (qmax >= n.median) search(n.right, q, out);
  }

  void query(int qi, const Box& q, Sink& sink) const {
    if (nodes.empty()) return;
    std::vector<int> leaves;
    search(0, q, leaves);
    for (int li : leaves)
      for (int64_t i = nodes[li].lo; i <= nodes[li].hi; ++i)
        if (pin_box(q, pts[i].x, pts[i].y)) sink.hit(qi, pts[i].tid, pts[i].pid);
  }
};

// Run `body(q)` for queries [0, nq) on `nthreads` threads, slicing the query
// list as Grid::rangeQueryBatchMultiThread does (Grid.cpp:407-426).
template <class F>
void for_queries(int nq, int nthreads, F body) {
  if (nthreads <= 1 || nq <= 1) {
    for (int q = 0; q < nq; ++q) body(q);
    return;
  }
  std::vector<std::thread> th;
  int at = 0;
  for (int t = 0; t < nthreads && at < nq; ++t) {
    int n = nq / nthreads + 1;
    if (at + n >= nq) n = nq - at;
    th.emplace_back([=] {
      for (int q = at; q < at + n; ++q) body(q);
    });
    at += n;
  }
  for (auto& t : th) t.join();
}

// Shared driver for every range-query restatement.  mode 0: bitmap,
// 1: pairs (grouped by query), 2: counts + digests only.
template <class Index>
int64_t run_range(const Index& ix, const float* boxes, int nq, int nthreads, uint8_t* bitmap,
                  int64_t stride, int64_t* counts, uint64_t* digests, int32_t* pairs, int64_t cap) {
  std::vector<std::vector<int32_t>> buf;
  bool want_pairs = pairs != nullptr;
  if (want_pairs) buf.resize(nq);
  if (counts) std::memset(counts, 0, sizeof(int64_t) * nq);
  if (digests) std::memset(digests, 0, sizeof(uint64_t) * nq);
  if (bitmap) std::memset(bitmap, 0, size_t(nq) * size_t(stride));
  for_queries(nq, nthreads, [&](int q) {
    Sink s;
    s.bitmap = bitmap;
    s.stride = stride;
    s.counts = counts;
    s.digests = digests;
    s.pairs = want_pairs ? &buf : nullptr;
    ix.query(q, box_from(boxes + 4 * int64_t(q)), s);
  });
  int64_t total = 0;
  if (want_pairs) {
    for (int q = 0; q < nq; ++q) total += int64_t(buf[q].size() / 2);
    if (total > cap) return -total;
    int64_t at = 0;
    for (int q = 0; q < nq; ++q) {
      std::memcpy(pairs + 2 * at, buf[q].data(), buf[q].size() * sizeof(int32_t));
      at += int64_t(buf[q].size() / 2);
    }
  } else if (counts) {
    for (int q = 0; q < nq; ++q) total += counts[q];
  }
  return total;
}

struct BruteIndex {
  const int64_t* offs;
  int ntraj;
  const float* lon;
  const float* lat;
  void query(int qi, const Box& q, Sink& sink) const {
    for (int t = 1; t <= ntraj; ++t)
      for (int64_t i = offs[t - 1]; i < offs[t]; ++i)
        if (pin_box(q, lon[i], lat[i])) sink.hit(qi, t, int32_t(i - offs[t - 1]));
  }
};

// The EDR match test of Grid.cpp:1281: |p1.x - p2.x| < EPSILON && |p1.y - p2.y| <
// EPSILON, with the difference computed in float and compared as a double.
inline bool edr_match(float ax, float ay, float bx, float by, double eps) {
  float dx = ax - bx;
  float dy = ay - by;
  return double(std::fabs(dx)) < eps && double(std::fabs(dy)) < eps;
}

// Rows are the shorter trajectory; on equal lengths the query is the rows
// (Grid.cpp:1234-1247).
inline void orient(const float*& r, int& nr, const float*& c, int& nc, const float* q, int nq,
                   const float* cand, int ncand) {
  if (ncand >= nq) {
    r = q; nr = nq; c = cand; nc = ncand;
  } else {
    r = cand; nr = ncand; c = q; nc = nq;
  }
}

// The cell rule of Grid.cpp:1288-1300.
inline int edr_rule(int m, int u, int l) {
  bool c1 = (m < u) && (m < l);
  bool c2 = (l < u) && (l < m);
  return c1 ? m : (c2 ? l : u);
}

// Grid::SimilarityExecuter — Grid.cpp:1222-1318, row by row.
int edr_rowmajor(const float* q, int nq, const float* cand, int ncand, double eps) {
  const float *r, *c;
  int nr, nc;
  orient(r, nr, c, nc, q, nq, cand, ncand);
  std::vector<int> prev(nc + 1), cur(nc + 1);
  for (int j = 0; j <= nc; ++j) prev[j] = j;
  for (int i = 1; i <= nr; ++i) {
    cur[0] = i;
    for (int j = 1; j <= nc; ++j) {
      int s = edr_match(r[2 * (i - 1)], r[2 * (i - 1) + 1], c[2 * (j - 1)], c[2 * (j - 1) + 1], eps) ? 0 : 1;
      cur[j] = edr_rule(prev[j - 1] + s, prev[j] + 1, cur[j - 1] + 1);
    }
    std::swap(prev, cur);
  }
  return prev[nc];
}

// The same recurrence walked anti-diagonal by anti-diagonal over a full table,
// the evaluation order of the reference's GPU kernel (kernel.cu:327-406).
int edr_antidiag(const float* q, int nq, const float* cand, int ncand, double eps) {
  const float *r, *c;
  int nr, nc;
  orient(r, nr, c, nc, q, nq, cand, ncand);
  std::vector<int> t(size_t(nr + 1) * (nc + 1));
  auto at = [&](int i, int j) -> int& { return t[size_t(i) * (nc + 1) + j]; };
  for (int i = 0; i <= nr; ++i) at(i, 0) = i;
  for (int j = 0; j <= nc; ++j) at(0, j) = j;
  for (int d = 2; d <= nr + nc; ++d) {
    int ilo = std::max(1, d - nc), ihi = std::min(nr, d - 1);
    for (int i = ihi; i >= ilo; --i) {
      int j = d - i;
      int s = edr_match(r[2 * (i - 1)], r[2 * (i - 1) + 1], c[2 * (j - 1)], c[2 * (j - 1) + 1], eps) ? 0 : 1;
      at(i, j) = edr_rule(at(i - 1, j - 1) + s, at(i - 1, j) + 1, at(i, j - 1) + 1);
    }
  }
  return at(nr, nc);
}

// Textbook EDR (minimum of the three moves) with the same match test; used
// only as a lower bound of the reference's recurrence.
int edr_textbook(const float* q, int nq, const float* cand, int ncand, double eps) {
  const float *r, *c;
  int nr, nc;
  orient(r, nr, c, nc, q, nq, cand, ncand);
  std::vector<int> prev(nc + 1), cur(nc + 1);
  for (int j = 0; j <= nc; ++j) prev[j] = j;
  for (int i = 1; i <= nr; ++i) {
    cur[0] = i;
    for (int j = 1; j <= nc; ++j) {
      int s = edr_match(r[2 * (i - 1)], r[2 * (i - 1) + 1], c[2 * (j - 1)], c[2 * (j - 1) + 1], eps) ? 0 : 1;
      cur[j] = std::min(prev[j - 1] + s, std::min(prev[j] + 1, cur[j - 1] + 1));
    }
    std::swap(prev, cur);
  }
  return prev[nc];
}

// ------------------------------------------------------------------------
// The reference's CPU top-k: frequency vectors over virtual cells, the
// frequency-distance (FD) queue and the rounds of 80 candidates.

// getIdxFromXY (Grid.cpp:25-54) as the reference's x86-64 g++ -O2 build
// evaluates it for the negative coordinates findNeighbor passes at the grid's
// edge (FVTable.cpp:67-73): int(log2(negative)) converts NaN, which x86-64
// turns into INT_MIN (bit length INT_MIN + 1), and 2 * max(lenx, leny) wraps
// modulo 2^32.  Equal to morton() for non-negative coordinates.
int morton_x86(int x, int y) {
  auto bitlen = [](int v) { return v == 0 ? 1 : v < 0 ? INT_MIN + 1 : int(std::log2(v)) + 1; };
  const int total = int(2u * uint32_t(std::max(bitlen(x), bitlen(y))));
  int result = 0, xb = 0, yb = 0;
  for (int i = 1; i <= total; ++i) {
    if (i & 1) {
      result += ((x >> xb) & 1) * (1 << (i - 1));
      ++xb;
    } else {
      result += ((y >> yb) & 1) * (1 << (i - 1));
      ++yb;
    }
  }
  return result;
}

// FVTable::findNeighbor — FVTable.cpp:53-75: decode the cell's int(log2(cellNum))
// bits (even bits x, odd bits y), then the eight cells around it, x-major.
void fv_neighbors(int cell, int cellNum, int out[8]) {
  int x = 0, y = 0;
  const int bits = int(std::log2(cellNum));
  for (int bit = 0; bit < bits; ++bit) {
    if (bit % 2 == 0)
      x += ((cell >> bit) & 1) * (1 << (bit / 2));
    else
      y += ((cell >> bit) & 1) * (1 << (bit / 2));
  }
  int n = 0;
  for (int xx = x - 1; xx <= x + 1; ++xx)
    for (int yy = y - 1; yy <= y + 1; ++yy)
      if (xx != x || yy != y) out[n++] = morton_x86(xx, yy);
}

using FreqVector = std::map<int, int>;  // virtual cell -> points

// One trajectory's frequency distance to the query — the body of
// FVTable::formPriorityQueue, FVTable.cpp:86-198.  The difference t - q is
// split into positive and negative parts (89-123); each positive cell then
// cancels against negative neighbours (129-156) and each negative cell against
// positive neighbours (160-187); FD = max(sum of positives, sum of negatives).
// As in the reference, a cell's value is read once before its neighbour loop
// (132, 163): after a partial cancellation the next neighbour is compared
// with the cell's original value.  Where the reference erases the element its
// loop iterator points to and then advances that iterator (145-146, 150-152,
// 176-177, 181-183: undefined behaviour), the loop continues at the element
// that followed it; nothing but the current element is ever erased from the
// map being walked, so that is the order the loop would have seen.
int fd_value(const FreqVector& t, const FreqVector& q, int cellNum) {
  FreqVector pos, neg;
  for (const auto& e : t) {
    auto it = q.find(e.first);
    if (it == q.end()) {
      pos.insert(e);
    } else {
      const int d = e.second - it->second;
      if (d > 0)
        pos.insert({e.first, d});
      else if (d < 0)
        neg.insert({e.first, -d});
    }
  }
  for (const auto& e : q) {
    if (t.count(e.first)) continue;
    const int d = -e.second;
    if (d > 0)
      pos.insert({e.first, d});
    else if (d < 0)
      neg.insert({e.first, -d});
  }
  int nb[8];
  for (auto it = pos.begin(); it != pos.end();) {
    const auto next = std::next(it);
    const int cid = it->first, posi = it->second;
    fv_neighbors(cid, cellNum, nb);
    for (int i = 0; i < 8; ++i) {
      auto jn = neg.find(nb[i]);
      if (jn == neg.end()) continue;
      const int nega = jn->second;
      if (posi > nega) {
        it->second = posi - nega;
        neg.erase(jn);
      } else if (posi == nega) {
        pos.erase(it);
        neg.erase(jn);
        break;
      } else {
        jn->second = nega - posi;
        pos.erase(it);
        break;
      }
    }
    it = next;
  }
  for (auto it = neg.begin(); it != neg.end();) {
    const auto next = std::next(it);
    const int cid = it->first, nega = it->second;
    fv_neighbors(cid, cellNum, nb);
    for (int i = 0; i < 8; ++i) {
      auto jp = pos.find(nb[i]);
      if (jp == pos.end()) continue;
      const int posi = jp->second;
      if (nega > posi) {
        it->second = nega - posi;
        pos.erase(jp);
      } else if (posi == nega) {
        neg.erase(it);
        pos.erase(jp);
        break;
      } else {
        jp->second = posi - nega;
        neg.erase(it);
        break;
      }
    }
    it = next;
  }
  int sp = 0, sn = 0;
  for (const auto& e : pos) sp += e.second;
  for (const auto& e : neg) sn += e.second;
  return std::max(sp, sn);
}

// FDwithID and the two heap orders of FVTable.h:8-23.
struct FdWithId {
  int traID;
  int FD;
};
struct FdMinFirst {  // cmp: the smallest FD on top
  bool operator()(const FdWithId& a, const FdWithId& b) const { return a.FD > b.FD; }
};
struct FdMaxFirst {  // cmpBig: the largest on top
  bool operator()(const FdWithId& a, const FdWithId& b) const { return a.FD < b.FD; }
};
using FdQueue = std::priority_queue<FdWithId, std::vector<FdWithId>, FdMinFirst>;
using ResultHeap = std::priority_queue<FdWithId, std::vector<FdWithId>, FdMaxFirst>;

// The reference's Grid as its CPU top-k sees it: the cell geometry (for the
// query's cells), each trajectory's frequency vector over virtual cells
// (Morton cell >> VITURAL_CELL_PARAM, Grid.cpp:152-182, VITURAL_CELL_PARAM = 2
// as main.cpp builds the grid) and the trajectories' points (the reference
// reassembles them from the cell-ordered store, Grid.cpp:1056-1071, which
// gives each trajectory's points in order).
struct FvIndex {
  GridIndex grid;
  int shift = 2;
  int cellNum = 0;  // FVTable::cellNum = cellnum >> 2 (Grid.cpp:214)
  std::vector<FreqVector> fv;  // by 1-based trajectory id
  std::vector<int64_t> offs;
  std::vector<float> xy;  // interleaved (lon, lat)

  void build(const Box& r, float cs, const int64_t* o, int ntraj, const float* lon, const float* lat) {
    grid.init(r, cs, 256);
    cellNum = grid.cellnum >> 2;
    fv.assign(size_t(ntraj) + 1, FreqVector{});
    offs.assign(o, o + ntraj + 1);
    xy.resize(size_t(o[ntraj] - o[0]) * 2);
    for (int t = 0; t < ntraj; ++t)
      for (int64_t i = o[t]; i < o[t + 1]; ++i) {
        // addTrajectoryIntoCell adds each same-cell run's point count to its
        // virtual cell (Grid.cpp:128-199): every point once.
        ++fv[size_t(t) + 1][grid.cell_of(lon[i], lat[i]) >> shift];
        xy[size_t(i - o[0]) * 2] = lon[i];
        xy[size_t(i - o[0]) * 2 + 1] = lat[i];
      }
  }
  // The query's frequency vector — Grid.cpp:1141-1157.
  FreqVector query_fv(const float* q, int n) const {
    FreqVector f;
    for (int i = 0; i < n; ++i) ++f[grid.cell_of(q[2 * i], q[2 * i + 1]) >> shift];
    return f;
  }
  // FVTable::formPriorityQueue — FVTable.cpp:77-207: trajectories 1..trajNum
  // pushed in id order.
  void form_queue(const FreqVector& qf, FdQueue* queue) const {
    for (size_t i = 1; i < fv.size(); ++i) queue->push(FdWithId{int(i), fd_value(fv[i], qf, cellNum)});
  }
  const float* traj(int id, int* len) const {
    *len = int(offs[size_t(id)] - offs[size_t(id) - 1]);
    return xy.data() + size_t(offs[size_t(id) - 1] - offs[0]) * 2;
  }
};

struct FdTopkStats {
  int64_t rounds = 0, evaluated = 0;
};

// Grid::SimilarityMultiThreadHandler for one query — Grid.cpp:1021-1125:
// rounds of KSIMILARITY = 80 candidates taken from the FD queue (1046-1051),
// their EDR by SimilarityExecuter (1078), a max-heap of the kValue best
// (1082-1106), until the worst kept distance no longer exceeds the next FD
// (1040).  Where the reference would take top() of an exhausted queue
// (undefined), a round takes what is left and the loop ends with the queue.
void fd_topk_query(const FvIndex& ix, FdQueue& queue, const float* q, int qlen, int kValue, double eps,
                   ResultHeap& calc, FdTopkStats& st) {
  constexpr int kBatch = 80;  // KSIMILARITY, ConstDefine.h:38
  int worstNow = 9999999;
  int held = 0;
  while (!queue.empty() && worstNow > queue.top().FD) {
    int cand[kBatch];
    int m = 0;
    while (m < kBatch && !queue.empty()) {
      cand[m++] = queue.top().traID;
      queue.pop();
    }
    for (int i = 0; i < m; ++i) {
      int len = 0;
      const float* c = ix.traj(cand[i], &len);
      const int d = edr_rowmajor(q, qlen, c, len, eps);
      if (held < kValue) {
        calc.push(FdWithId{cand[i], d});
        ++held;
      } else if (d < calc.top().FD) {
        calc.pop();
        calc.push(FdWithId{cand[i], d});
      }
    }
    worstNow = calc.top().FD;
    ++st.rounds;
    st.evaluated += m;
  }
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

int orc_pinbox(const float* box, float x, float y) { return pin_box(box_from(box), x, y) ? 1 : 0; }

int orc_box_intersect(const float* a, const float* b) { return box_intersect(box_from(a), box_from(b)); }

int orc_morton(int x, int y) { return morton(x, y); }

uint64_t orc_pair_digest(int32_t tid, int32_t pid) { return pair_digest(tid, pid); }

// ---------------------------------------------------------------- Grid (GT-quadtree)
void* orc_grid_new(const float* range, float cell_size, int max_points_in_node) {
  auto* g = new GridIndex();
  g->init(box_from(range), cell_size, max_points_in_node);
  return g;
}

void orc_grid_free(void* g) { delete static_cast<GridIndex*>(g); }

int orc_grid_add(void* g, const int64_t* offs, int ntraj, const float* lon, const float* lat) {
  static_cast<GridIndex*>(g)->add(offs, ntraj, lon, lat);
  return 0;
}

// out: axis, cellnum, total_level, nodes, leaves, points
int orc_grid_info(void* gp, int64_t* out) {
  auto* g = static_cast<GridIndex*>(gp);
  int64_t leaves = 0;
  for (auto& n : g->nodes) leaves += n.leaf;
  out[0] = g->axis;
  out[1] = g->cellnum;
  out[2] = g->total_level;
  out[3] = int64_t(g->nodes.size());
  out[4] = leaves;
  out[5] = int64_t(g->store.x.size());
  return 0;
}

int orc_grid_cell_of(void* g, float lon, float lat) { return static_cast<GridIndex*>(g)->cell_of(lon, lat); }

int orc_grid_cell_box(void* gp, int cell, float* out) {
  auto* g = static_cast<GridIndex*>(gp);
  if (cell < 0 || cell >= g->cellnum) return -1;
  const Box& b = g->cell_box[cell];
  out[0] = b.xmin; out[1] = b.ymin; out[2] = b.xmax; out[3] = b.ymax;
  return 0;
}

// The store in index order: x, y, tid, pid (each `n` long).
int orc_grid_store(void* gp, float* x, float* y, int32_t* tid, int32_t* pid) {
  auto* g = static_cast<GridIndex*>(gp);
  size_t n = g->store.x.size();
  std::memcpy(x, g->store.x.data(), n * 4);
  std::memcpy(y, g->store.y.data(), n * 4);
  std::memcpy(tid, g->store.tid.data(), n * 4);
  std::memcpy(pid, g->store.pid.data(), n * 4);
  return 0;
}

// Leaves selected for one rectangle, in traversal order: (store start, count) pairs.
int64_t orc_grid_leaves(void* gp, const float* box, int64_t* out, int64_t cap) {
  auto* g = static_cast<GridIndex*>(gp);
  std::vector<int> leaves;
  g->match(0, box_from(box), leaves);
  int64_t n = int64_t(leaves.size());
  for (int64_t i = 0; i < n && i < cap; ++i) {
    out[2 * i] = g->leaf_start(g->nodes[leaves[i]]);
    out[2 * i + 1] = g->nodes[leaves[i]].npts;
  }
  return n;
}

int64_t orc_grid_range(void* g, const float* boxes, int nq, int nthreads, uint8_t* bitmap, int64_t stride,
                       int64_t* counts, uint64_t* digests, int32_t* pairs, int64_t cap) {
  return run_range(*static_cast<GridIndex*>(g), boxes, nq, nthreads, bitmap, stride, counts, digests, pairs, cap);
}

// ---------------------------------------------------------------- MortonGrid
void* orc_mgrid_new(const float* range, float cell_size, int max_points_in_node) {
  auto* m = new MortonIndex();
  m->init(box_from(range), cell_size, max_points_in_node);
  return m;
}

void orc_mgrid_free(void* m) { delete static_cast<MortonIndex*>(m); }

int orc_mgrid_add(void* m, const int64_t* offs, int ntraj, const float* lon, const float* lat) {
  static_cast<MortonIndex*>(m)->add(offs, ntraj, lon, lat);
  return 0;
}

int orc_mgrid_bitmap_index(int level, int id) { return MortonIndex::bitmap_index(level, id); }

// out: axis, cellnum, total_level, leaves, points
int orc_mgrid_info(void* mp, int64_t* out) {
  auto* m = static_cast<MortonIndex*>(mp);
  int64_t leaves = 0;
  for (bool b : m->is_leaf) leaves += b;
  out[0] = m->axis;
  out[1] = m->cellnum;
  out[2] = m->total_level;
  out[3] = leaves;
  out[4] = int64_t(m->store.x.size());
  return 0;
}

int orc_mgrid_is_leaf(void* mp, int level, int id) {
  auto* m = static_cast<MortonIndex*>(mp);
  return m->is_leaf[MortonIndex::bitmap_index(level, id)] ? 1 : 0;
}

int orc_mgrid_node_box(void* mp, int level, int id, float* out) {
  const Box b = static_cast<MortonIndex*>(mp)->node_box(level, id);
  out[0] = b.xmin; out[1] = b.ymin; out[2] = b.xmax; out[3] = b.ymax;
  return 0;
}

// Leaves one rectangle selects, in traversal order, as (level, id, store
// start, point count).
int64_t orc_mgrid_leaves(void* mp, const float* box, int64_t* out, int64_t cap) {
  auto* m = static_cast<MortonIndex*>(mp);
  std::vector<std::pair<int, int>> nodes;
  m->match(0, 0, box_from(box), nodes);
  int64_t n = int64_t(nodes.size());
  for (int64_t i = 0; i < n && i < cap; ++i) {
    int span = int(std::pow(4, m->total_level - nodes[i].first) + 0.5);
    int first = nodes[i].second * span;
    int64_t pts = 0;
    for (int c = first; c < first + span; ++c) pts += m->store.count[c];
    out[4 * i] = nodes[i].first;
    out[4 * i + 1] = nodes[i].second;
    out[4 * i + 2] = m->store.start[first];
    out[4 * i + 3] = pts;
  }
  return n;
}

int64_t orc_mgrid_range(void* m, const float* boxes, int nq, int nthreads, uint8_t* bitmap, int64_t stride,
                        int64_t* counts, uint64_t* digests, int32_t* pairs, int64_t cap) {
  return run_range(*static_cast<MortonIndex*>(m), boxes, nq, nthreads, bitmap, stride, counts, digests, pairs, cap);
}

// ---------------------------------------------------------------- FSG
void* orc_fsg_new(const float* range, float cell_size) {
  auto* f = new FsgIndex();
  f->init(box_from(range), cell_size);
  return f;
}
void orc_fsg_free(void* f) { delete static_cast<FsgIndex*>(f); }
int orc_fsg_add(void* f, const int64_t* offs, int ntraj, const float* lon, const float* lat) {
  static_cast<FsgIndex*>(f)->add(offs, ntraj, lon, lat);
  return 0;
}
int orc_fsg_cell_of(void* f, float lon, float lat) { return static_cast<FsgIndex*>(f)->cell_of(lon, lat); }
int64_t orc_fsg_range(void* f, const float* boxes, int nq, int nthreads, uint8_t* bitmap, int64_t stride,
                      int64_t* counts, uint64_t* digests, int32_t* pairs, int64_t cap) {
  return run_range(*static_cast<FsgIndex*>(f), boxes, nq, nthreads, bitmap, stride, counts, digests, pairs, cap);
}

// ---------------------------------------------------------------- STIG
void* orc_stig_new() { return new StigIndex(); }
void orc_stig_free(void* s) { delete static_cast<StigIndex*>(s); }
int orc_stig_add(void* s, const int64_t* offs, int ntraj, const float* lon, const float* lat, int block) {
  static_cast<StigIndex*>(s)->add(offs, ntraj, lon, lat, block);
  return 0;
}
int64_t orc_stig_range(void* s, const float* boxes, int nq, int nthreads, uint8_t* bitmap, int64_t stride,
                       int64_t* counts, uint64_t* digests, int32_t* pairs, int64_t cap) {
  return run_range(*static_cast<StigIndex*>(s), boxes, nq, nthreads, bitmap, stride, counts, digests, pairs, cap);
}

// ---------------------------------------------------------------- brute force
int64_t orc_brute_range(const int64_t* offs, int ntraj, const float* lon, const float* lat, const float* boxes,
                        int nq, int nthreads, uint8_t* bitmap, int64_t stride, int64_t* counts,
                        uint64_t* digests, int32_t* pairs, int64_t cap) {
  BruteIndex b{offs, ntraj, lon, lat};
  return run_range(b, boxes, nq, nthreads, bitmap, stride, counts, digests, pairs, cap);
}

// ---------------------------------------------------------------- EDR
// Points are interleaved (x, y) floats.
int orc_edr(const float* q, int nq, const float* c, int nc, double eps) { return edr_rowmajor(q, nq, c, nc, eps); }
int orc_edr_antidiag(const float* q, int nq, const float* c, int nc, double eps) {
  return edr_antidiag(q, nq, c, nc, eps);
}
int orc_edr_textbook(const float* q, int nq, const float* c, int nc, double eps) {
  return edr_textbook(q, nq, c, nc, eps);
}

// One query against candidates 1..ncand stored back to back (offsets in points,
// ncand+1 entries); out[i] = distance to candidate i+1.
int orc_edr_batch(const float* q, int nq, const int64_t* coffs, int ncand, const float* cxy, double eps,
                  int32_t* out, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([=] {
      for (int i = t; i < ncand; i += nthreads)
        out[i] = edr_rowmajor(q, nq, cxy + 2 * coffs[i], int(coffs[i + 1] - coffs[i]), eps);
    });
  for (auto& x : th) x.join();
  return 0;
}

// The all-mismatch table F[r][c] (r, c in 0..maxlen): the recurrence with every
// substitution costing 1.  A pair with no matching point has distance
// F[min(n,m)][max(n,m)].
int orc_ftable(int maxlen, int32_t* out) {
  int w = maxlen + 1;
  for (int j = 0; j < w; ++j) out[j] = j;
  for (int i = 1; i < w; ++i) {
    out[int64_t(i) * w] = i;
    for (int j = 1; j < w; ++j)
      out[int64_t(i) * w + j] =
          edr_rule(out[int64_t(i - 1) * w + j - 1] + 1, out[int64_t(i - 1) * w + j] + 1, out[int64_t(i) * w + j - 1] + 1);
  }
  return 0;
}

// Exact top-k of n (distance, id) pairs in ascending (distance, id) order.
int orc_topk(const int32_t* dist, const int32_t* ids, int n, int k, int32_t* out_ids, int32_t* out_dist) {
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  auto less = [&](int a, int b) { return dist[a] != dist[b] ? dist[a] < dist[b] : ids[a] < ids[b]; };
  int kk = std::min(k, n);
  std::partial_sort(idx.begin(), idx.begin() + kk, idx.end(), less);
  for (int i = 0; i < kk; ++i) {
    out_ids[i] = ids[idx[i]];
    out_dist[i] = dist[idx[i]];
  }
  return kk;
}

// ---- the reference's CPU top-k (FVTable + 80-candidate rounds), a timed baseline

int orc_morton_x86(int x, int y) { return morton_x86(x, y); }

int orc_fv_neighbors(int cell, int cellNum, int32_t* out) {
  int nb[8];
  fv_neighbors(cell, cellNum, nb);
  for (int i = 0; i < 8; ++i) out[i] = nb[i];
  return 0;
}

// FD of two frequency vectors given as (cell, count) pairs.
int orc_fd_value(const int32_t* t, int nt, const int32_t* q, int nq, int cellNum) {
  FreqVector a, b;
  for (int i = 0; i < nt; ++i) a[t[2 * i]] += t[2 * i + 1];
  for (int i = 0; i < nq; ++i) b[q[2 * i]] += q[2 * i + 1];
  return fd_value(a, b, cellNum);
}

void* orc_fv_new(const float* range, float cell_size, const int64_t* offs, int ntraj, const float* lon,
                 const float* lat) {
  auto* f = new FvIndex();
  f->build(box_from(range), cell_size, offs, ntraj, lon, lat);
  return f;
}

void orc_fv_free(void* f) { delete static_cast<FvIndex*>(f); }

// FD of query (qlen interleaved points) to every trajectory: out[t - 1].
int orc_fv_fd(void* fp, const float* q, int qlen, int32_t* out) {
  const FvIndex& f = *static_cast<FvIndex*>(fp);
  const FreqVector qf = f.query_fv(q, qlen);
  for (size_t i = 1; i < f.fv.size(); ++i) out[i - 1] = fd_value(f.fv[i], qf, f.cellNum);
  return 0;
}

// Grid::SimilarityQueryBatchCPUParallel — Grid.cpp:1133-1220: the query
// frequency vectors (1141-1157), one FD thread per query (1162-1168), then
// batches of THREAD_CPU = 20 query threads, each running the 80-candidate
// rounds for its query (1183-1197); out_ids[q * kValue + i] in the order the
// reference pops its max-heap, worst first (1202-1209), 0 past the heap's
// size.  stats[2q], stats[2q + 1]: rounds and candidates evaluated.
int orc_fv_topk(void* fp, const int64_t* qoffs, int nq, const float* qxy, int kValue, double eps, int32_t* out_ids,
                int32_t* out_dist, int64_t* stats) {
  const FvIndex& f = *static_cast<FvIndex*>(fp);
  std::vector<FdQueue> queues(static_cast<size_t>(nq));
  std::vector<FreqVector> qfv(static_cast<size_t>(nq));
  for (int q = 0; q < nq; ++q) qfv[q] = f.query_fv(qxy + 2 * qoffs[q], int(qoffs[q + 1] - qoffs[q]));
  {
    std::vector<std::thread> th;
    for (int q = 0; q < nq; ++q) th.emplace_back([&, q] { f.form_queue(qfv[q], &queues[q]); });
    for (auto& t : th) t.join();
  }
  std::vector<ResultHeap> calc(static_cast<size_t>(nq));
  std::vector<FdTopkStats> st(static_cast<size_t>(nq));
  constexpr int kThreads = 20;  // THREAD_CPU, Grid.cpp:1183
  for (int q0 = 0; q0 < nq; q0 += kThreads) {
    std::vector<std::thread> th;
    for (int q = q0; q < std::min(nq, q0 + kThreads); ++q)
      th.emplace_back([&, q] {
        fd_topk_query(f, queues[q], qxy + 2 * qoffs[q], int(qoffs[q + 1] - qoffs[q]), kValue, eps, calc[q], st[q]);
      });
    for (auto& t : th) t.join();
  }
  for (int q = 0; q < nq; ++q) {
    for (int i = 0; i < kValue; ++i) {
      const bool has = !calc[q].empty();
      out_ids[int64_t(q) * kValue + i] = has ? calc[q].top().traID : 0;
      if (out_dist) out_dist[int64_t(q) * kValue + i] = has ? calc[q].top().FD : -1;
      if (has) calc[q].pop();
    }
    if (stats) {
      stats[2 * q] = st[q].rounds;
      stats[2 * q + 1] = st[q].evaluated;
    }
  }
  return 0;
}

}  // extern "C"
