// Grid — the GT-quadtree index of the reference (Grid.h:34-100): a square grid
// of 2^k x 2^k cells numbered in Morton order, a cell-ordered point store, and a
// quadtree over Morton ranges whose leaves hold fewer than maxPointsInNode
// points (or are single cells).  Range queries select the leaves whose box
// overlaps the rectangle on the host and refine their points on the GPU;
// similarity queries run the EDR kernels against a trajectory-ordered copy of
// the points.  The indexes and stores are built once and stay resident in HBM.
#pragma once

#include <cstdint>
#include <vector>

#include "DeviceArena.h"
#include "MBB.h"
#include "RangeEngine.h"
#include "Trajectory.h"

struct QuadtreeNode {  // Grid.h:19-26, flattened: children are indices into Grid::nodes
  int NodeID = 0;      // the node's Morton block at its level
  int level = 0;
  int64_t numPoints = 0;
  int64_t pointStart = 0;  // first store index of the node's cells
  bool isLeaf = true;
  int child[4] = {-1, -1, -1, -1};  // UL, UR, DL, DR
  MBB mbb;
};

// Top-k answer of one similarity query: ids and distances in ascending
// (distance, trajectory id) order.
struct TopK {
  std::vector<int32_t> ids, dist;
};

// Counters of one batched EDR call.
struct EdrBatchStats {
  int64_t pairs = 0;       // pairs whose DP ran
  int64_t cells = 0;       // DP cells of those pairs (sum of n * m)
  int64_t jobs = 0;
  int64_t pruned = 0;      // pairs resolved by the all-mismatch table
  float kernelMs = 0;      // batched DP kernel (device events)
  float filterMs = 0;      // candidate filter kernel (device events)
  double hostMs = 0;       // host job building (and the host top-k merge, k > 1024)
  double uploadMs = 0;     // query store packing and upload
  double mergeMs = 0;      // device top-k merge stage: survivor lists, upload, merge kernel, readback
};

class Grid {
 public:
  Grid() = default;
  // Grid.cpp:97-124.  VITURAL_CELL_PARAM only drove the reference's frequency
  // vectors (Grid.cpp:152) and is kept for signature parity.
  Grid(const MBB& mbb, float val_cell_size, int VITURAL_CELL_PARAM, int maxPointsInNode = MAXPOINTINNODE_DEFAULT);
  Grid(const Grid&) = delete;
  Grid& operator=(const Grid&) = delete;
  virtual ~Grid() = default;

  // Grid.cpp:210-316: builds the cell-ordered store and the quadtree and makes
  // both resident on the device (with upload=false only the host structures
  // are built — used to test the host logic without a GPU).
  int addDatasetToGrid(Trajectory* db, int traNum);
  int addDatasetToGrid(const TrajectoryView& db, bool upload = true);

  int WhichCellPointIn(SamplePoint p) const { return int(cellOf(p.lon, p.lat)); }
  uint32_t cellOf(float lon, float lat) const;
  MBB cellMBB(uint32_t cellId) const;

  // Grid.cpp:450-542: the trajectory-set answer of every rectangle, computed on
  // the GPU (ResultTable[i] is resized to trajNum + 1).  resultSetSize, when
  // non-null, receives each query's number of trajectories.
  int rangeQueryBatchGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                         int device_idx = 0);
  // The same selection refined into compacted (tid, pid) pairs left in HBM.
  int rangeQueryBatchGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out);
  // Batched form (no reference counterpart): the rectangles are selected on the
  // device against a breadth-first copy of the quadtree and refined leaf by
  // leaf (RangeEngine::runTree).  Same result sets as rangeQueryBatchGPUPairs.
  int rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph = false);
  // Points of the leaves the last batched run selected at least once.
  int uniqueCandidates(int64_t* points) { return engine_.uniqueCandidates(tree_, points); }
  // The quadtree as breadth-first gat_tree_nodes (the root untested, every
  // other node tested with BoxIntersect != 0); *depth = deepest level.
  virtual std::vector<gat_tree_node> treeNodes(int* depth) const;
  // The owner of every level-k block for the device walk's start table
  // (gat_tree_start): (level << 27) | breadth-first index; empty when the
  // tree does not qualify.
  std::vector<uint32_t> startOwners(int k) const;
  // The start table the device tree gets for a tree of this depth: its
  // owners, *level and *block (the blocks' side); empty (level 0) when none.
  std::vector<uint32_t> startTable(int depth, int* level, float* block) const;
  // Number of quadtree leaves.
  virtual int64_t leafNum() const;
  // Grid.cpp:428-446: the leaves one rectangle selects, in traversal order.
  void findMatchNodeInQuadTree(const MBB& bound, std::vector<int>* leaves) const;
  // The leaves one rectangle selects (the batched form's device selection
  // applies the same rule) and their points: what its candidate and result
  // buffers need, at most `points` result pairs over `leaves` list entries.
  void selectionCost(const MBB& bound, int64_t* points, int64_t* leaves) const;
  // Host candidate selection for a batch (tasks + per-query candidate counts).
  void selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                        std::vector<uint64_t>& candidates) const;

  // Similarity (Grid.cpp:1321-1755, plain form): one EDR launch per query
  // over every trajectory, then an exact host top-k by (distance, id).
  // topKSimilarityTraj receives queryTrajNum * kValue ids.
  int SimilarityQueryBatchOnGPU(Trajectory* qTra, int queryTrajNum, int* topKSimilarityTraj, int kValue);
  int similarityPlain(const TrajectoryView& queries, int kValue, std::vector<TopK>& out);
  // EDR distances of one query against every trajectory (plain kernel).
  int edrAll(const float* qlon, const float* qlat, int qlen, int32_t* dist);

  // Batched forms (no reference counterpart).  edrBatchedAll: the distance of
  // every (query, trajectory) pair, dist[q * trajNum + t - 1], in one
  // persistent launch.  similarityBatched: exact top-k by (distance, id) for
  // every query: pairs the candidate filter drops take the all-mismatch table
  // value, the rest run the batched DP.
  int edrBatchedAll(const TrajectoryView& queries, std::vector<int32_t>& dist, EdrBatchStats* stats);
  int similarityBatched(const TrajectoryView& queries, int kValue, std::vector<TopK>& out, EdrBatchStats* stats);
  // The all-mismatch table F[r][c], (GAT_MAX_TRAJ_LEN + 1)^2 entries, computed
  // on the device once.
  const std::vector<int32_t>& mismatchTable();

  // Batched-EDR tunables: pairs whose longer trajectory has at most edrSplit
  // points form the short bucket, those up to max(edrSplit, 1024) the long
  // one, and longer pairs a third bucket at the kernel's default occupancy;
  // workgroups per CU of the short and long launches (0: the kernel's default
  // for its LDS class).
  int edrSplit = 512;
  int edrPerCuShort = 0, edrPerCuLong = 0;
  // Waves per pair (1, 2, 4 or 8) for the long pairs of a batch small enough
  // to fit one round of the device's waves (a top-k batch's survivors, a
  // serving batch), which ends when its longest pair does: above 1 the pair's
  // strips run on that many waves in lockstep (edr_pipe_kernel).  Only pairs
  // that one wave would keep busy past the longest pair's pipelined schedule
  // take them, unless edrWavesEveryLongPair is set.
  int edrWavesPerPair = 4;
  bool edrWavesEveryLongPair = false;
  // Top-k batches of k <= 1024 and at most 8192 queries run filter, jobs, DP,
  // survivor lists and merge on the device with no host step between them,
  // and the batched EDR forms build their job lists on the device; off: the
  // jobs (and lists) are built on the host.
  bool deviceChain = true;

  // Batched EDR of given (query index, trajectory index) pairs, both 0-based:
  // dist[i] for pairs[2i], pairs[2i + 1].
  int edrBatchedPairs(const TrajectoryView& queries, std::vector<int32_t>& pairs, std::vector<int32_t>& dist,
                      EdrBatchStats* stats);

  // Run-time EDR threshold (the reference's compile-time EPSILON, ConstDefine.h:28).
  double epsilon = 0.005;
  int hostThreads = 0;  // 0: hardware concurrency (capped at 16)
  // With upload: build the cell-ordered store on the device (gat_cell_store_build)
  // instead of bucketing on the host; the same store either way.
  bool deviceBuild = false;

  // Geometry (Grid.h:69-80).
  MBB range;
  float cell_size = 0;
  int cellNum_axis = 0;
  int cellnum = 0;
  int totalLevel = 0;
  int maxPointsInNode = MAXPOINTINNODE_DEFAULT;
  int64_t totalPointNum = 0;
  int trajNum = 0;
  int nodeNum = 0;
  std::vector<QuadtreeNode> nodes;   // nodes[0] is the root
  std::vector<int64_t> cellStart;    // cellnum + 1 prefix of per-cell point counts

  RangeEngine& engine() { return engine_; }
  const DeviceTree& deviceTree() const { return tree_; }
  double lastSelectMs = 0;
  float lastKernelMs = 0;

 protected:
  // The host tree over cellStart, then (finishBuild) its device form and the
  // trajectory store.
  virtual int buildHostTree();
  virtual int finishBuild(const TrajectoryView& db);
  // Appends the store runs of the leaves `bound` selects, as query q.
  virtual void addLeafRuns(const MBB& bound, uint32_t q, TaskList& list) const;
  int threads() const;
  int uploadTrajectories(const TrajectoryView& db);

  RangeEngine engine_;
  DeviceTree tree_;

 private:
  int buildNode(int level, int id);
  int ensureQueryStore(const TrajectoryView& queries);
  int checkTrajLengths(const char* what) const;
  int runBatched(std::vector<int32_t>& pairs, int queryNum, std::vector<int32_t>& dist, EdrBatchStats* stats);
  // Pinned arena bytes runBatched takes for np pairs (its callers begin() it).
  static size_t runBatchedStaging(int64_t np);
  // Top-k of up to kDeviceTopkMax per query merged on the device; larger k on the host.
  static constexpr int kDeviceTopkMax = 1024;
  int deviceTopk(const std::vector<int32_t>& pairs, const std::vector<int32_t>& dist, int nq, int kValue,
                 std::vector<TopK>& out);
  // The whole top-k batch on the device (filter, jobs, DP, survivor lists,
  // merge) with one readback at the end; *done is false when the batch must
  // take the host-built path instead (a query with more survivors than the
  // device's per-query sort holds).
  int similarityOnDevice(int nq, int kValue, std::vector<TopK>& out, EdrBatchStats* st, bool* done);
  // The DP launches over a device-built job list (windows in win).
  int launchDeviceBuckets(int64_t cap, int nq, const int32_t* win);
  // Batched EDR with the job list built on the device (pairsHost null: every
  // query x trajectory pair); dist in the pairs' order.
  int runBatchedOnDevice(const int32_t* pairsHost, int64_t np, int queryNum, std::vector<int32_t>& dist,
                         EdrBatchStats* st);
  static constexpr int kDeviceListQueries = 8192;
  // Trajectory-ordered copy of the points for EDR (the reference reassembles
  // trajectories from the cell-ordered store per query, Grid.cpp:1500-1508).
  // Stores are NaN-padded (GAT_EDR_PAD points before every trajectory,
  // GAT_EDR_TAIL_PAD after the last) for the batched kernel.
  DeviceBuffer trajXY_, trajOff_, trajLen_, trajBox_, allTasks_, dist_;
  // The candidate grid of the top-k filter (gat_edr_filter_grid), built with
  // the trajectory store.
  DeviceBuffer candStart_, candId_, candBox_, candAlways_;
  gat_edr_cand_grid candGrid_{};
  int buildCandidateGrid(const std::vector<float>& boxes, int T);
  // The top-k filter over the uploaded query boxes: every (query, trajectory)
  // pair whose boxes could hold a matching point pair.
  int filterPairs(int nq, int64_t cap, void* stream);
  DeviceBuffer queryXY_, queryBox_;
  // The uploaded queries: offsets and lengths, then the raw coordinates
  // (packed into queryXY_ on the device); the pointers below view it.
  DeviceBuffer queryUp_;
  const int64_t* querySrcOffDev_ = nullptr;
  const int64_t* queryOff64Dev_ = nullptr;
  const int32_t* queryOffDev_ = nullptr;
  const int32_t* queryLenDev_ = nullptr;
  PinnedBuffer queryStage_;  // query coordinates and offsets, uploaded from here
  PinnedArena stage_;        // the transfers of one batched EDR / top-k call
  DeviceBuffer pairsDev_, jobsDev_, outDev_, counterDev_, ftableDev_;
  // Top-k merge on the device: trajectories grouped by length (0-based ids,
  // ascending) with per-length offsets; per batch, each query's survivors
  // ascending by id with their distances, and the results.
  DeviceBuffer lenIdsDev_, lenOffDev_, survDev_, topDev_;
  // The device-built job list: pairs grouped by bucket, scratch, bucket
  // windows + status + cell count, and the DP's job counters.
  DeviceBuffer sortedDev_, nmDev_, windowsDev_, dpCounterDev_, originDev_, unsortDev_;
  int64_t deviceTopkCap_ = 0;  // survivor capacity of the device path, grown on overflow
  EventTimer edrTimer_, filterTimer_, mergeTimer_;
  SideStream edrSide_;  // the longest EDR bucket runs here, beside the others
  SideStream edrSide2_;  // one-wave long pairs of a batch whose longest run over several waves
  std::vector<int32_t> trajLenHost_, queryLenHost_;
  int badTrajLen_ = 0;  // first trajectory (1-based) whose length the EDR kernels reject, or 0
  std::vector<int32_t> ftable_;
  std::vector<std::vector<int32_t>> lengthIds_;  // per length: trajectory ids, ascending
  // Per query length a: the index's trajectory lengths L ordered by their
  // all-mismatch value F[min(a, L)][max(a, L)] (then by L), built on first use.
  std::vector<std::vector<std::pair<int32_t, int32_t>>> lengthOrder_;
};
