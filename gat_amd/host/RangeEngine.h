// The device side of every range index: a resident point store in HBM (the
// index's points in its own order, as (lon, lat) floats and (tid, pid) ints),
// pinned staging for the per-batch task lists, and the refine launches.
//
// It replaces the reference's per-index arenas and leaf caches: the reference
// copies each candidate leaf into a bump-allocated device arena on first use
// (Grid.cpp:715-727, FSG.cpp:268-273) or on every use (STIG.cpp:293) and builds
// a RangeQueryStateTable per (query, leaf chunk) on the host (Grid.cpp:729-747);
// here the whole store is uploaded once when the index is built and a batch
// only uploads its task list.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "DeviceArena.h"
#include "MBB.h"
#include "Trajectory.h"

// Candidate runs longer than this are split across tasks so one wave never
// walks an unbounded run (the reference splits at MAXPOINTINNODE,
// Grid.cpp:729-747).
constexpr uint32_t kTaskChunk = 1024;

// Builds one query's task list: appends candidate runs in order, merging a run
// that starts where the previous one of the same query ended, and splitting at
// kTaskChunk points.
class TaskList {
 public:
  void addRun(uint64_t start, uint64_t count, uint32_t qid);
  void flush();
  std::vector<gat_range_task> tasks;
  std::vector<uint64_t> candidates;  // per query handled by this list (indexed by qid - firstQuery)
  uint32_t firstQuery = 0;
  uint64_t runsAdded = 0;    // addRun calls with points (one per selected leaf), before merging
  uint64_t pointsAdded = 0;  // their points

 private:
  bool open_ = false;
  uint64_t runStart_ = 0, runCount_ = 0;
  uint32_t runQid_ = 0;
};

// Result of one batch in compacted-pair form; the pairs stay in HBM.
struct RangeBatch {
  int queryNum = 0;
  std::vector<uint64_t> base;   // per query: first pair slot in `pairs`
  std::vector<uint32_t> count;  // per query: number of result pairs
  int64_t candidates = 0;       // candidate points scanned by the batch
  int64_t taskNum = 0;
  int64_t pairTotal = 0;
  int64_t leafQueries = 0;      // batched form: selected (leaf, query) entries
  int64_t workItems = 0;        // batched form: leaf-major work items
  bool replayed = false;        // batched form: ran as a replay of a captured graph
  double selectMs = 0;          // host candidate selection (batched form: device selection and scans)
  double uploadMs = 0;          // task-list upload
  float kernelMs = 0;           // refine kernel (device events)
  float refineMs = 0;           // batched form, step-by-step run: the leaf-major refine kernel alone
  DeviceBuffer pairs;           // int32 (tid, pid) pairs
  DeviceBuffer qbase, qcount;   // device copies of base / count
  // Run-indexed layout (leaf-major form with runIndexed): the pairs are
  // compacted over the batch in work-item order, query q's set being the union
  // of its runs (gat_range_run: query, count, first pair); base is unused.
  bool runIndexed = false;
  DeviceBuffer runs;
  int64_t runTotal = 0;   // run records (one per list entry; count 0 when no hits)
  int64_t pairSpan = 0;   // run-indexed: extent of `pairs` the regions cover
  bool hostStale = false;       // base / count not yet copied from the device
  void* stream = nullptr;       // the engine stream that produced the batch (readbacks wait on it only)
  // Copies base / count from the device when they are stale.
  int syncHost();
  // Copies the pairs to the host, grouped by query (query q's pairs at
  // out[2*base[q] ...], count[q] of them).
  int fetch(int32_t* out, int64_t capacityPairs);
  // Copies the pairs to the host query after query, back to back (query q's
  // after those of queries 0..q-1); capacityPairs must hold pairTotal.
  int fetchCompact(int32_t* out, int64_t capacityPairs);
  // Run-indexed layout: the run records (runTotal of them).
  int fetchRuns(gat_range_run* out, int64_t capacity);
  // Per-query order-independent digest (wrapping sum of a 64-bit mix of each
  // (tid, pid) pair), computed on the device.
  int digests(uint64_t* out) const;
  // The same into a caller-owned device buffer of queryNum entries (no
  // allocation), copied to `out` when it is not null.
  int digestsInto(uint64_t* dev, uint64_t* out) const;
};

// Device copy of an index's tree for the batched form: gat_tree_node in
// breadth-first order, nchild children per inner node, leaves at depth <= depth.
// A flat grid (nx, ny > 0) is the same with no inner nodes: nodes[] holds one
// leaf record per cell, row-major from (xmin, ymax), and selection enumerates
// a rectangle's cell range instead of walking a tree.  An implicit quadtree
// (implicitLevel L > 0, gat_morton_select) keeps no inner nodes either:
// nodes[] holds the leaves in Morton order and cellWords a (leaf code, store
// start) pair per cell of the 2^L x 2^L Morton grid, plus one past the end.
struct DeviceTree {
  DeviceBuffer nodes;
  int64_t nnodes = 0;
  int nchild = 0;
  int depth = 0;
  int nx = 0, ny = 0;  // flat grid only
  float xmin = 0, ymax = 0, cellSize = 0;
  DeviceBuffer cellWords;  // implicit quadtree only
  int implicitLevel = 0;
  // Start table of a quadtree walk (gat_tree_start): the owner of each block
  // of level startLevel, blocks of side startBlock from (xmin, ymax).
  DeviceBuffer startOwner;
  int startLevel = 0;
  float startBlock = 0;
  // Bumped by every upload: a captured pipeline bakes the tree's sizes and
  // geometry into its launches and is replayed only for the same generation.
  uint64_t generation = 0;
  bool flat() const { return nx > 0; }
  bool implicit() const { return implicitLevel > 0; }
  // Checks the traversal stack bound and uploads `host`.
  int upload(const std::vector<gat_tree_node>& host, int nchild, int depth, void* stream);
  // Uploads the start table (4^level owners) of the tree upload() holds.
  int uploadStart(const std::vector<uint32_t>& owner, int level, float xmin, float ymax, float block, void* stream);
  // Uploads one leaf record per cell of an nx x ny grid.
  int uploadGrid(const std::vector<gat_tree_node>& cells, int nx, int ny, float xmin, float ymax, float cellSize,
                 void* stream);
  // Uploads an implicit quadtree: the leaf records in Morton order and the
  // 4^level + 1 (code, store start) cell pairs (gat_morton_select).
  int uploadImplicit(const std::vector<gat_tree_node>& leaves, const std::vector<uint32_t>& words, int level,
                     float xmin, float ymax, float cellSize, void* stream);
};

class RangeEngine {
 public:
  RangeEngine() = default;
  RangeEngine(const RangeEngine&) = delete;
  RangeEngine& operator=(const RangeEngine&) = delete;

  // Uploads the index-ordered store: xy holds (lon, lat) pairs, tp (tid, pid)
  // pairs; ids are packed into 32 bits when they fit (gat_point_store).
  int upload(const float* xy, const int32_t* tp, int64_t n);
  // Builds the store on the device from trajectory-ordered points, ordered by
  // the cells of `grid` (gat_cell_store_build), with the same id packing as
  // upload(); *cellStart receives the grid's cellnum + 1 store offsets and
  // *badPoint the first point off the grid (then no store is kept) or -1.
  int buildStore(const gat_cell_grid& grid, const TrajectoryView& db, std::vector<int64_t>* cellStart,
                 int64_t* badPoint);
  // Builds the store on the device in the order of a kd-tree of the given
  // shape (gat_kd_store_build), with the same id packing; *medians gets one
  // value per split node, *leafBoxes four per leaf.
  struct KdSkeleton {
    std::vector<int64_t> splitLo, splitHi;  // split nodes, depth by depth, ascending within a depth
    std::vector<int64_t> depthFirst;        // depths + 1 offsets into splitLo / splitHi
    std::vector<int64_t> leafLo, leafHi;    // inclusive leaf ranges
  };
  int buildKdStore(const TrajectoryView& db, const KdSkeleton& skeleton, std::vector<float>* medians,
                   std::vector<float>* leafBoxes);
  int64_t pointNum() const { return n_; }
  const gat_point_store& store() const { return store_; }
  void* stream() { return stream_.get(); }
  int ensureStream() { return stream_.create(); }

  // Candidate selection callback: appends query q's runs to `list`.
  using Selector = std::function<void(int q, TaskList& list)>;

  // Selects candidates for every query on `threads` host threads, then refines
  // them into compacted pairs.
  int runPairs(const MBB* bounds, int nq, const Selector& select, int threads, RangeBatch& out);
  // Selects candidates, then fills the trajectory bitmap through the
  // reference-signature handler (synchronous, host result).
  int runBitmap(const MBB* bounds, int nq, const Selector& select, int threads, int maxTrajNum,
                uint8_t* resultHost, bool stigHandler);
  // Batched form: device selection against `tree`, grouping by leaf and the
  // leaf-major refine, for the whole batch.  With graph=true the device
  // pipeline is captured once per buffer layout and replayed while the batch
  // fits the captured capacities.
  int runTree(const DeviceTree& tree, const MBB* bounds, int nq, RangeBatch& out, bool graph);
  int queriesPerItem = 32;  // rectangles per leaf-major work item (1..64)
  // Batches of at most this many rectangles take the direct form: every
  // selected (leaf, rectangle) becomes a task of the per-task refine kernel,
  // and nothing sized by the tree is cleared or scanned.
  int directMaxQueries = 256;
  // Leaf-major form: lay the pairs out by work item with a run index
  // (gat_leaf_refine_runs) instead of in per-query regions.  Each work item's
  // pairs go to one contiguous chunk, so far fewer lines are shared between
  // waves; per-query counts, digests and fetches are the same either way.
  bool runIndexed = true;
  // After a batched run: the points of the leaves at least one rectangle
  // selected (what the leaf-major refine must read at least once).
  int uniqueCandidates(const DeviceTree& tree, int64_t* points);
  ~RangeEngine();

  // Pieces of the batched form for a coordinator that captures several
  // engines into one graph.  After a runTree of the same batch size has sized
  // the workspace: stageBatch uploads the rectangles, enqueueBatch queues the
  // whole device pipeline on this engine's stream (no host sync, capturable),
  // finishBatch reads the counters back and reports in *fits whether the batch
  // stayed inside the workspace (if not, rerun it with runTree).
  int stageBatch(const DeviceTree& tree, const MBB* bounds, int nq);
  int enqueueBatch(const DeviceTree& tree, int nq, RangeBatch& out);
  int finishBatch(int nq, RangeBatch& out, bool* fits);

  // Everything a captured batch pipeline has baked into its kernel arguments:
  // every device buffer it reads or writes, the capacities it was sized with,
  // the batch size and the layout switches.  A captured graph may be replayed
  // only while the key it was captured under still equals layoutKey(): any
  // workspace growth reallocates a buffer and changes the key.
  struct GraphKey {
    static constexpr int kPtrs = 20;
    const void* ptr[kPtrs] = {};
    int64_t nq = -1, lqCap = 0, qlistCap = 0, itemsCap = 0, pairsCap = 0, runsCap = 0;
    uint64_t treeGeneration = 0;
    int qb = 0;
    bool runs = false;
    bool operator==(const GraphKey& o) const {
      for (int i = 0; i < kPtrs; ++i)
        if (ptr[i] != o.ptr[i]) return false;
      return nq == o.nq && lqCap == o.lqCap && qlistCap == o.qlistCap && itemsCap == o.itemsCap &&
             pairsCap == o.pairsCap && runsCap == o.runsCap && qb == o.qb && runs == o.runs &&
             treeGeneration == o.treeGeneration;
    }
    bool operator!=(const GraphKey& o) const { return !(*this == o); }
  };
  GraphKey layoutKey(const DeviceTree& tree, int nq, const RangeBatch& out) const;

  // Shares `src`'s resident store (which must outlive this engine): a second
  // engine over the same points with its own stream and batch workspace.
  int attachStore(const RangeEngine& src);
  // Sizes the batched-form workspace for batches of up to maxQueries
  // rectangles and maxCandidates candidate points, so later batches within
  // those bounds allocate nothing.
  int reserveBatched(const DeviceTree& tree, int maxQueries, int64_t maxCandidates, RangeBatch& out);

  // Candidate selection only (host), exposed for tests of the host logic.
  static void select(int nq, const Selector& select, int threads, std::vector<gat_range_task>& tasks,
                     std::vector<uint64_t>& candidates);

 private:
  int stage(const MBB* bounds, int nq, const std::vector<gat_range_task>& tasks, const uint64_t* qbase);
  int stageBuild(const TrajectoryView& db, int pidBits, DeviceBuffer& lon, DeviceBuffer& lat, DeviceBuffer& off);
  struct BoxLayout {
    size_t order = 0, ntasks = 0, tasks = 0, bytes = 0, tasksCap = 0;
  };
  static BoxLayout boxLayout(int nq);
  int uploadBoxes(const DeviceTree* tree, const MBB* bounds, int nq);
  int enqueueSelect(const DeviceTree& tree, int nq, RangeBatch& out);
  // The selection kernel for the tree's kind (nodeCount may be null).  The
  // start-table walk adds to the candidate counts, which it clears first
  // unless the caller already has (candCleared).
  static bool selectAddsCandidates(const DeviceTree& tree) {
    return !tree.flat() && !tree.implicit() && tree.startLevel > 0;
  }
  int enqueueTreeSelect(const DeviceTree& tree, int nq, uint32_t* nodeCount, bool candCleared = false);
  int enqueueJoin(const DeviceTree& tree, int nq, RangeBatch& out);
  int readCounters(int nq, RangeBatch& out);
  bool countersFit() const;
  int finishCounters(RangeBatch& out, bool replay);
  int runDirect(const DeviceTree& tree, int nq, RangeBatch& out);
  void dropGraph();

  DeviceBuffer xy_, id_, pid_;
  gat_point_store store_{};
  int64_t n_ = 0;
  Stream stream_;
  PinnedBuffer hostStage_;
  DeviceBuffer devTasks_, devBoxes_;
  EventTimer timer_;

  // Batched-form workspace.  counters_ (uint32 words): [0,1] selected entries,
  // [2,3] candidate total, [4] status, [5] list total, [6] work items,
  // [8,9] result pairs, [10,11] run-indexed region total.
  DeviceBuffer nodeCount_, lqRank_, nodeOff_, tiles_, cand_, lq_, qlist_, items_, counters_, directTasks_;
  DeviceBuffer itemBase_, itemTiles_;  // run-indexed refine: item regions and their scan scratch
  PinnedBuffer countersHost_;
  int64_t lqCap_ = 0, qlistCap_ = 0, itemsCap_ = 0, pairsCap_ = 0;
  EventTimer selTimer_, refineTimer_;
  bool timeRefine_ = false;
  void* graph_ = nullptr;
  GraphKey graphKey_;
};
