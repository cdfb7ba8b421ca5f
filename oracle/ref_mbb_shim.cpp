// TEST INFRASTRUCTURE ONLY.  C entry points over the reference's own MBB class
// (MBB.h / MBB.cpp, compiled from /root/reference by oracle/Makefile), so the
// tests can check the oracle's containment and box-intersection restatements
// against the reference's compiled code.
#include "MBB.h"

int BoxIntersect(MBB& a1, MBB& b1);  // MBB.cpp:28

extern "C" int ref_pinbox(const float* box, float x, float y) {
  MBB b(box[0], box[1], box[2], box[3]);
  return b.pInBox(x, y) ? 1 : 0;
}

extern "C" int ref_box_intersect(const float* a, const float* b) {
  MBB ma(a[0], a[1], a[2], a[3]);
  MBB mb(b[0], b[1], b[2], b[3]);
  return ma.intersect(mb);
}
