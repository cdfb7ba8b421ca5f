// Mixed range batches over the three indexes of one data set (no reference
// counterpart: the reference runs each index's test separately,
// SystemTest.cpp:24-102, 251-384).
//
// Each rectangle goes to one index by its area: the flat grid (FSG) for the
// smallest, whose cell range is a handful of cells; the GT-quadtree for the
// middle; STIG for the largest.  The batch runs in groups of at most
// `groupSize` rectangles per index, so the result buffers stay bounded however
// many pairs the whole batch has; each group's counts and order-independent
// digests are kept per rectangle, and its pairs are overwritten by the next
// group.  The first group of a shape is run step by step on every index
// (sizing the workspaces), then the three device pipelines are captured into
// one graph — forked from a coordinating stream onto each index's stream and
// joined back — and every later group of that shape is one graph launch, as
// long as no engine's workspace has moved since the capture.
#pragma once

#include <cstdint>
#include <vector>

#include "FSG.h"
#include "Grid.h"
#include "STIG.h"

struct MixedConfig {
  double fsgMaxArea = 1e-6;   // rectangles smaller than this go to FSG
  double gridMaxArea = 1e-3;  // then to the GT-quadtree; the rest to STIG
  int groupSize = 2048;       // rectangles per index per group
  bool graph = true;
};

struct MixedResult {
  std::vector<uint32_t> count;   // per rectangle: result pairs
  std::vector<uint64_t> digest;  // per rectangle: wrapping sum of the pair mix (gat_pairs_digest)
  std::vector<int8_t> route;     // per rectangle: 0 FSG, 1 GT-quadtree, 2 STIG
  int64_t pairs = 0;
  int64_t candidates = 0;
  int groups = 0, replays = 0;
  double ms = 0;                 // wall time of the whole run
  double digestMs = 0;           // of which spent digesting results
};

class MixedRange {
 public:
  MixedRange(Grid& grid, STIG& stig, FSG& fsg) : grid_(grid), stig_(stig), fsg_(fsg) {}
  ~MixedRange();
  MixedRange(const MixedRange&) = delete;
  MixedRange& operator=(const MixedRange&) = delete;
  int run(const MBB* bounds, int n, const MixedConfig& cfg, MixedResult* out);

 private:
  int capture(const int nq[3]);
  RangeEngine& engine(int i);
  const DeviceTree& tree(int i);
  Grid& grid_;
  STIG& stig_;
  FSG& fsg_;
  RangeBatch batch_[3];
  DeviceBuffer digests_[3];
  Stream coord_;
  void* fork_ = nullptr;
  void* join_[3] = {nullptr, nullptr, nullptr};
  void* graph_ = nullptr;
  int shape_[3] = {-1, -1, -1};
  // Each engine's layout key at capture: the graph replays only while all
  // three still hold (an engine's own range_batched, tune or reserve call can
  // grow and so reallocate its workspace between two runs).
  RangeEngine::GraphKey keys_[3];
  bool keysMatch(const int nq[3]);
 public:
  int captures = 0;  // graphs captured over this object's life
};
