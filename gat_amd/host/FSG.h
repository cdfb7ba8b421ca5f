// FSG — the flat row-major grid index of the reference (FSG.h:9-46): cells
// numbered row * cell_num_x + col from (xmin, ymax), a cell-ordered point
// store, and range queries that scan every cell of the rectangle's row/column
// range.  Because a row of cells is contiguous in the store, each rectangle
// row becomes one candidate run.
#pragma once

#include <vector>

#include "MBB.h"
#include "RangeEngine.h"
#include "Trajectory.h"

class FSG {
 public:
  FSG() = default;
  FSG(const MBB& mbb, float val_cell_size);  // FSG.cpp:7-28
  FSG(const FSG&) = delete;
  FSG& operator=(const FSG&) = delete;

  int addDatasetToGrid(Trajectory* db, int traNum);  // FSG.cpp:95-143
  int addDatasetToGrid(const TrajectoryView& db, bool upload = true);
  int WhichCellPointIn(SamplePoint p) const;  // FSG.cpp:88-93
  int64_t cellOf(float lon, float lat) const;

  // FSG.cpp:145-393: trajectory-set answers on the GPU.
  int rangeQueryBatchGPU(const MBB* bounds, int rangeNum, CPURangeQueryResult* ResultTable, int* resultSetSize,
                         int device_idx = 0);
  int rangeQueryBatchGPUPairs(const MBB* bounds, int rangeNum, RangeBatch& out);
  // The rectangle's column range [g1, g2] and row range [g3, g4] (FSG.cpp:183-186),
  // clamped to the grid; false when the rectangle has no cell.
  bool cellRange(const MBB& q, int* g1, int* g2, int* g3, int* g4) const;
  void selectCandidates(const MBB* bounds, int rangeNum, std::vector<gat_range_task>& tasks,
                        std::vector<uint64_t>& candidates) const;
  // Batched form (no reference counterpart), as Grid::rangeQueryBatchedGPU:
  // the cells are the leaves, each rectangle enumerating its cell range on the
  // device.
  int rangeQueryBatchedGPU(const MBB* bounds, int rangeNum, RangeBatch& out, bool graph = false);
  int uniqueCandidates(int64_t* points) { return engine_.uniqueCandidates(tree_, points); }

  MBB range;
  float cell_size = 0;
  int cell_num_x = 0, cell_num_y = 0;
  int64_t cellnum = 0;
  int64_t totalPointNum = 0;
  int trajNum = 0;
  int hostThreads = 0;
  // With upload: build the cell-ordered store on the device (gat_cell_store_build)
  // instead of bucketing on the host; the same store either way.
  bool deviceBuild = false;
  std::vector<int64_t> cellStart;  // cellnum + 1
  RangeEngine& engine() { return engine_; }
  const DeviceTree& deviceTree() const { return tree_; }

 private:
  void addRows(const MBB& q, uint32_t qid, TaskList& list) const;
  int threads() const;
  int uploadCells();
  RangeEngine engine_;
  DeviceTree tree_;
};
