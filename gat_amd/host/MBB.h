// Geometry and shared host types of the trajectory-query path: the rectangle
// class MBB (MBB.h:2-12), the reference's point and result types
// (ConstDefine.h:63-116, QueryResult.h:4) and the Morton cell numbering of the
// GT-quadtree (Grid.cpp:25-54).
#pragma once

#include <cstdint>
#include <vector>

// Leaf capacity of the GT-quadtree (ConstDefine.h:32 uses 20000; the workloads
// here use 256 points per leaf).
constexpr int MAXPOINTINNODE_DEFAULT = 256;

class MBB {
 public:
  float xmin = 0, ymin = 0, xmax = 0, ymax = 0;
  MBB() = default;
  MBB(float val_xmin, float val_ymin, float val_xmax, float val_ymax)
      : xmin(val_xmin), ymin(val_ymin), xmax(val_xmax), ymax(val_ymax) {}
  // Inclusive containment (MBB.cpp:21-26).
  bool pInBox(float x, float y) const { return x >= xmin && x <= xmax && y >= ymin && y <= ymax; }
  // BoxIntersect(*this, b) (MBB.cpp:80-82).
  int intersect(const MBB& b) const;
};

// MBB.cpp:28-73: 0 when the boxes are disjoint or only touch, 1 when they
// overlap, 2 when a1 contains b1, 3 when b1 contains a1.
int BoxIntersect(const MBB& a1, const MBB& b1);

// Cell id of (column x, row y): x's bits interleaved into the even bit
// positions and y's into the odd ones (Grid.cpp:25-54).
uint32_t getIdxFromXY(uint32_t x, uint32_t y);
// Inverse of getIdxFromXY.
void xyFromIdx(uint32_t idx, uint32_t* x, uint32_t* y);

// The reference's trajectory-set answer of one range query: entry tid is true
// when trajectory tid has a point inside the rectangle (QueryResult.h:4).
typedef std::vector<bool> CPURangeQueryResult;
