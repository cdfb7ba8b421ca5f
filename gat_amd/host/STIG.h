// STIG — the kd-tree index of the reference (STIG.h:9-66): all points split
// alternately on x and y at the median down to leaves of at most blockSize
// points, each leaf a contiguous range of the index-ordered store with the
// bounding box of its points.
//
// The split is a correct kd-split (the reference's std::sort range leaves the
// last element out, STIG.cpp:101/135, and its left leaf repeats the median
// point, STIG.cpp:107-118): the node's points are partitioned around the
// element at (lo+hi)/2, left [lo, mid-1] holding keys <= the median and right
// [mid, hi] keys >= it.  The search visits the left child when the rectangle
// reaches down to the median and the right child when it reaches up to it
// (the reference bounds those half-planes by a hard-coded [0, 180],
// STIG.cpp:206-230), and keeps a leaf when its box overlaps the rectangle
// inclusively (the reference's BoxIntersect drops touching leaves,
// STIG.cpp:243), so the answer is exactly the inclusive containment set.
#pragma once

#include <cstdint>
#include <vector>

#include "MBB.h"
#include "RangeEngine.h"
#include "Trajectory.h"

struct STIGBlock {  // STIG.h:9-12
  int64_t startIdx;
  int64_t endIdx;  // inclusive
};

class STIG {
 public:
  STIG() = default;
  STIG(const STIG&) = delete;
  STIG& operator=(const STIG&) = delete;

  // STIG.cpp:35-66 (totalDim must be 2).
  int initial(int blockSize, int totalDim, Trajectory* db, int trajNum);
  int initial(int blockSize, const TrajectoryView& db, bool upload = true);
  // STIG.cpp:179-252: the leaves a rectangle selects, in store order.
  int searchTree(const MBB& queryMBB, std::vector<STIGBlock>* allCandBlocks) const;
  // STIG.cpp:254-338: trajectory-set answers on the GPU.
  int rangeQueryGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                    int device_idx = 0);
  int rangeQueryGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out);
  void selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                        std::vector<uint64_t>& candidates) const;
  // Batched form (no reference counterpart), as Grid::rangeQueryBatchedGPU.
  int rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph = false);
  // Points of the leaves the last batched run selected at least once.
  int uniqueCandidates(int64_t* points) { return engine_.uniqueCandidates(tree_, points); }
  // The kd-tree as breadth-first gat_tree_nodes: an inner node's children
  // carry the half-plane of its split (inclusive, as searchTree), a leaf its
  // points' box (inclusive); *depth = deepest level.
  std::vector<gat_tree_node> treeNodes(int* depth) const;

  struct Node {
    bool leaf = true;
    int dim = 0;
    float median = 0;
    int left = -1, right = -1;
    int64_t lo = 0, hi = -1;  // leaf range (inclusive)
    MBB box;
  };
  std::vector<Node> nodes;  // nodes[0] is the root
  int blockSize = 0;
  int64_t totalPointsNum = 0;
  int maxTid = 0;
  int hostThreads = 0;
  // With upload: order the points on the device (gat_kd_store_build) instead
  // of partitioning them on the host.  Same tree shape and medians defined
  // the same way; points with equal coordinates may sit on the other side of
  // a median than in the host build, which changes no answer.
  bool deviceBuild = false;
  RangeEngine& engine() { return engine_; }
  const DeviceTree& deviceTree() const { return tree_; }

 private:
  struct P {
    float x, y;
    int32_t tid, pid;
  };
  int skeleton(int depth, int64_t lo, int64_t hi, bool leaf);
  int buildOnDevice(const TrajectoryView& db);
  void partition(int node, std::vector<P>& pts, int spawnDepth);
  void addLeaves(const MBB& q, uint32_t qid, TaskList& list) const;
  int threads() const;
  RangeEngine engine_;
  DeviceTree tree_;
};
