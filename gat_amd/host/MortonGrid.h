// MortonGrid — the reference's implicit GT-quadtree (MortonGrid.h:27-86,
// MortonGrid.cpp): the same 2^L x 2^L Morton grid, cell-ordered point store
// and leaf rule as Grid (a node is a leaf when its cells hold fewer than
// maxPointsInNode points or it is a single cell), with the tree kept as one
// leaf flag per (level, id) instead of node records, and node boxes computed
// from their first and last cells.  It selects the same leaves as Grid, so its
// range answers are Grid's.
//
// On the device the tree is one word per cell (the leaf that starts there,
// by rank and level) plus the leaf records in Morton order; the batched form
// walks it with gat_morton_select, the rest of the pipeline is Grid's.
// Similarity queries are Grid's (the reference's MortonGrid stubs them out,
// MortonGrid.cpp:614-646).
#pragma once

#include <cstdint>
#include <vector>

#include "Grid.h"

struct MortonNode {  // MortonGrid.h:21-24
  int level = 0;
  int nid = 0;
};

class MortonGrid : public Grid {
 public:
  // MortonGrid.cpp:23-49 (the grid geometry is Grid's, Grid.cpp:97-124).
  MortonGrid(const MBB& mbb, float val_cell_size, int VITURAL_CELL_PARAM,
             int maxPointsInNode = MAXPOINTINNODE_DEFAULT);

  // MortonGrid.cpp:8-13: the flag of node (level, id) in isLeaf.
  static int64_t findIdxInBitmap(int level, int id);
  // MortonGrid.cpp:238-261: flags the leaves at or below (level, id).
  int buildQuadTree(int level, int id);
  // MortonGrid.cpp:263-268: first cell's xmin and ymax, last cell's xmax and ymin.
  MBB generateMBBfromNode(int level, int id) const;
  bool nodeIsLeaf(int level, int id) const { return isLeaf[size_t(findIdxInBitmap(level, id))]; }
  // MortonGrid.cpp:355-393: the leaves `bound` selects below `node`, children
  // in UL, UR, DL, DR order, each visited when BoxIntersect != 0 (the node
  // itself is not tested).
  int findMatchNodeInQuadTree(MortonNode node, const MBB& bound, std::vector<MortonNode>* cells) const;
  // First store index and point count of a node's cells.
  int64_t nodeStart(const MortonNode& n) const;
  int64_t nodePoints(const MortonNode& n) const;

  // The leaves in Morton order as gat_tree_nodes (an implicit tree has no
  // inner records); *depth = totalLevel.
  std::vector<gat_tree_node> treeNodes(int* depth) const override;
  int64_t leafNum() const override { return int64_t(leaves.size()); }

  std::vector<bool> isLeaf;        // (4^(L+1) - 1) / 3 flags, levels 0..L
  std::vector<MortonNode> leaves;  // in Morton order (depth-first, UL, UR, DL, DR)

 protected:
  int buildHostTree() override;
  int finishBuild(const TrajectoryView& db) override;
  void addLeafRuns(const MBB& bound, uint32_t q, TaskList& list) const override;
};
