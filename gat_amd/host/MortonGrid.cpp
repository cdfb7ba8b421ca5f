#include "MortonGrid.h"

#include <algorithm>

MortonGrid::MortonGrid(const MBB& mbb, float val_cell_size, int VITURAL_CELL_PARAM, int maxPoints)
    : Grid(mbb, val_cell_size, VITURAL_CELL_PARAM, maxPoints) {}

// (4^level - 1) / 3 nodes sit above level `level`.
int64_t MortonGrid::findIdxInBitmap(int level, int id) {
  if (level <= 0) return 0;
  return ((int64_t(1) << (2 * level)) - 1) / 3 + id;
}

int64_t MortonGrid::nodeStart(const MortonNode& n) const {
  const int64_t first = int64_t(n.nid) << (2 * (totalLevel - n.level));
  return cellStart[size_t(first)];
}

int64_t MortonGrid::nodePoints(const MortonNode& n) const {
  const int shift = 2 * (totalLevel - n.level);
  const int64_t first = int64_t(n.nid) << shift, end = first + (int64_t(1) << shift);
  return cellStart[size_t(end)] - cellStart[size_t(first)];
}

int MortonGrid::buildQuadTree(int level, int id) {
  const MortonNode n{level, id};
  if (nodePoints(n) < maxPointsInNode || level == totalLevel) {
    isLeaf[size_t(findIdxInBitmap(level, id))] = true;
    leaves.push_back(n);
    return 0;
  }
  for (int k = 0; k < 4; ++k) buildQuadTree(level + 1, (id << 2) + k);
  return 0;
}

MBB MortonGrid::generateMBBfromNode(int level, int id) const {
  const int shift = 2 * (totalLevel - level);
  const uint32_t first = uint32_t(id) << shift, last = first + (uint32_t(1) << shift) - 1u;
  const MBB a = cellMBB(first), b = cellMBB(last);
  return MBB(a.xmin, b.ymin, b.xmax, a.ymax);
}

int MortonGrid::findMatchNodeInQuadTree(MortonNode node, const MBB& bound, std::vector<MortonNode>* cells) const {
  if (nodeIsLeaf(node.level, node.nid)) {
    cells->push_back(node);
    return 0;
  }
  for (int k = 0; k < 4; ++k) {
    const MortonNode child{node.level + 1, (node.nid << 2) + k};
    if (bound.intersect(generateMBBfromNode(child.level, child.nid))) findMatchNodeInQuadTree(child, bound, cells);
  }
  return 0;
}

int MortonGrid::buildHostTree() {
  // The implicit device tree packs a first cell into 28 bits of a word.
  if (totalLevel > 14) {
    gath_set_error("MortonGrid: %d levels exceed the implicit tree's 14", totalLevel);
    return GAT_ERR_ARG;
  }
  isLeaf.assign(size_t(findIdxInBitmap(totalLevel + 1, 0)), false);
  leaves.clear();
  buildQuadTree(0, 0);
  return GAT_OK;
}

std::vector<gat_tree_node> MortonGrid::treeNodes(int* depth) const {
  std::vector<gat_tree_node> out(leaves.size());
  for (size_t k = 0; k < leaves.size(); ++k) {
    const MBB b = generateMBBfromNode(leaves[k].level, leaves[k].nid);
    gat_tree_node& d = out[k];
    d.box[0] = b.xmin;
    d.box[1] = b.ymin;
    d.box[2] = b.xmax;
    d.box[3] = b.ymax;
    d.first_child = -1;
    d.start = uint32_t(nodeStart(leaves[k]));
    d.count = uint32_t(nodePoints(leaves[k]));
    d.flags = leaves[k].level == 0 ? GAT_NODE_UNTESTED : 0u;
  }
  *depth = totalLevel;
  return out;
}

int MortonGrid::finishBuild(const TrajectoryView& db) {
  if (int rc = buildHostTree()) return rc;
  int depth = 0;
  const std::vector<gat_tree_node> recs = treeNodes(&depth);
  // Per cell (code, store start); the code says "no leaf starts here" except
  // at the leaves' first cells.  One more pair holds the total.
  std::vector<uint32_t> words(2 * (size_t(cellnum) + 1), 0xFFFFFFFFu);
  for (size_t c = 0; c <= size_t(cellnum); ++c) words[2 * c + 1] = uint32_t(cellStart[c]);
  for (size_t k = 0; k < leaves.size(); ++k) {
    const uint32_t first = uint32_t(leaves[k].nid) << (2 * (totalLevel - leaves[k].level));
    words[2 * size_t(first)] = (uint32_t(k) << 4) | uint32_t(leaves[k].level);
  }
  if (int rc = tree_.uploadImplicit(recs, words, totalLevel, range.xmin, range.ymax, cell_size, engine_.stream()))
    return rc;
  return uploadTrajectories(db);
}

void MortonGrid::addLeafRuns(const MBB& bound, uint32_t q, TaskList& list) const {
  std::vector<MortonNode> found;
  findMatchNodeInQuadTree(MortonNode{0, 0}, bound, &found);
  for (const MortonNode& n : found) list.addRun(uint64_t(nodeStart(n)), uint64_t(nodePoints(n)), q);
}
