#include "MBB.h"

int BoxIntersect(const MBB& a1, const MBB& b1) {
  // The reference orders the pair so that the first box starts no further right
  // than the second; on equal xmin the box reaching further right comes first,
  // and on a full tie the second argument does (MBB.cpp:31-52).
  const bool a1_first = a1.xmin < b1.xmin || (a1.xmin == b1.xmin && a1.xmax > b1.xmax);
  const MBB& a = a1_first ? a1 : b1;
  const MBB& b = a1_first ? b1 : a1;
  // Disjoint or touching along x, or along y (MBB.cpp:53-60).
  if (b.xmin >= a.xmax || b.ymax <= a.ymin || b.ymin >= a.ymax) return 0;
  // Containment of the second box by the first (MBB.cpp:62-69).
  const bool inside = a.pInBox(b.xmin, b.ymin) && a.pInBox(b.xmin, b.ymax) && a.pInBox(b.xmax, b.ymin) &&
                      a.pInBox(b.xmax, b.ymax);
  if (!inside) return 1;
  return a1_first ? 2 : 3;
}

int MBB::intersect(const MBB& b) const { return BoxIntersect(*this, b); }

namespace {
// Spread the low 16 bits of v to the even bit positions.
inline uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
inline uint32_t gather16(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0F0F0F0Fu;
  v = (v | (v >> 4)) & 0x00FF00FFu;
  v = (v | (v >> 8)) & 0x0000FFFFu;
  return v;
}
}  // namespace

uint32_t getIdxFromXY(uint32_t x, uint32_t y) { return spread16(x) | (spread16(y) << 1); }

void xyFromIdx(uint32_t idx, uint32_t* x, uint32_t* y) {
  *x = gather16(idx);
  *y = gather16(idx >> 1);
}
