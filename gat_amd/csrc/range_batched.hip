// Batched range form for gfx950: device-side candidate selection over a
// breadth-first copy of the GT-quadtree (Grid.cpp:58-94, 428-446) or the STIG
// kd-tree (STIG.cpp:71-168, 179-252), and a leaf-major join.
//
// Pipeline for one batch of rectangles (all kernels, no host round trip):
//   tree_select_kernel   one wave per rectangle walks the tree with a stack in
//                        LDS, 64 nodes per step; emits (leaf, query) entries,
//                        each with its rank among its leaf's entries, and the
//                        candidate count of every rectangle;
//   scan kernels         per-query output regions (exclusive sum of candidate
//                        counts) and per-leaf list / work-item offsets;
//   leaf_scatter_kernel  groups the entries by leaf (each to its rank's slot);
//   leaf_refine_kernel   per work item (one leaf, up to qb rectangles) the
//                        leaf's points and ids are loaded into registers once
//                        and tested against each rectangle; hits are compacted
//                        with ballots and written as (tid, pid) pairs.
// A point of a leaf selected by r rectangles is read from HBM once instead of
// r times, which is what takes the batch below the per-query scan's traffic.
#include <cmath>

#include "gat_common.hpp"

namespace {

constexpr int kWave = 64;

struct Node {
  float4 box;
  int32_t first_child;
  uint32_t start, count, flags;
};
static_assert(sizeof(Node) == sizeof(gat_tree_node), "node layout");

struct Store {
  const float2* xy;
  const uint32_t* id;
  const uint16_t* pid;
  uint64_t npoints;
  uint32_t pid_bits;
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  return uint64_t(uniform(uint32_t(v))) | (uint64_t(uniform(uint32_t(v >> 32))) << 32);
}

// BoxIntersect(a1, b1) != 0 (MBB.cpp:28-60): the pair is ordered so the box
// starting further left (on a tie, reaching further right; on a full tie, b1)
// comes first, and boxes that only touch do not intersect.
__device__ __forceinline__ bool box_intersect(const float4& a1, const float4& b1) {
  const bool a1_first = a1.x < b1.x || (a1.x == b1.x && a1.z > b1.z);
  const float4 a = a1_first ? a1 : b1;
  const float4 b = a1_first ? b1 : a1;
  return !(b.x >= a.z || b.w <= a.y || b.y >= a.w);
}

__device__ __forceinline__ bool node_selected(const Node& n, const float4& q) {
  if (n.flags & GAT_NODE_UNTESTED) return true;
  if (n.flags & GAT_NODE_INCLUSIVE) return n.box.x <= q.z && n.box.z >= q.x && n.box.y <= q.w && n.box.w >= q.y;
  return box_intersect(q, n.box);
}

// MBB::pInBox (MBB.cpp:21-26): inclusive on all four edges.
__device__ __forceinline__ bool in_box(const float4& b, float x, float y) {
  return x <= b.z && x >= b.x && y <= b.w && y >= b.y;
}

// ---------------------------------------------------------------- fill
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Several spans cleared by one launch: a batch's resets are a few small
// buffers beside the per-node counts, and each launch of its own costs a
// graph node.
constexpr int kFillSpans = 4;
struct FillSpans {
  uint32_t* p[kFillSpans];
  uint64_t end[kFillSpans];  // running totals of the spans' lengths
};

__global__ void fill_spans_kernel(FillSpans f, uint32_t v) {
  const uint64_t n = f.end[kFillSpans - 1];
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    int k = 0;
    while (i >= f.end[k]) ++k;  // i < end[kFillSpans - 1], so k stays in range
    f.p[k][i - (k ? f.end[k - 1] : 0)] = v;
  }
}

// ---------------------------------------------------------------- selection
// The stack holds node indices still to be tested.  Each step pops up to 64
// entries from the top (one per lane), tests them, records selected leaves and
// pushes the children of selected inner nodes in lane order.  Depth never
// decreases from the bottom of the stack to its top, so each depth keeps at
// most nchild * 64 entries and a tree of depth D needs (D + 1) * nchild * 64
// slots; the step size is also cut so a push can never pass the end.
// Selected leaves collect in an LDS buffer that is flushed to the global
// (leaf, query) list with one atomic per flush, so a step costs one dependent
// memory access (its node loads).
constexpr uint32_t kEmitCap = 1024;
static_assert(kEmitCap % kWave == 0, "the flush covers the buffer in whole waves");

// Per-wave buffer of selected leaves in LDS, flushed to the global (leaf,
// query) list with one atomic per flush.  Every lane calls every member.
struct LeafEmitter {
  uint32_t* buf;
  uint32_t n;
  uint32_t q;
  uint32_t lane;
  uint2* lq;
  uint64_t cap;
  unsigned long long* total;
  uint32_t* node_qcount;
  uint32_t* rank;  // per list entry: its rank among its leaf's entries (or null)
  uint64_t cand;   // this lane's share of the rectangle's candidate points

  // Records the lanes with emit set (their leaf idx holds cnt points); flushes
  // when the buffer could not take another full wave.  The per-leaf counts are
  // added at the flush: an atomic here would hold the walk's next node loads
  // behind it (gfx9 waits for stores and atomics with the loads in vmcnt).
  __device__ void push(bool emit, uint32_t idx, uint32_t cnt) {
    const uint64_t em = __ballot(emit);
    if (emit) {
      buf[n + lanes_below(em)] = idx;
      cand += cnt;
    }
    n += uint32_t(__popcll(em));
  }
  __device__ void maybe_flush() {
    if (n > kEmitCap - kWave) flush();
  }
  __device__ void flush() {
    if (n == 0) return;
    __syncthreads();  // the buffer writes of every lane are visible
    if (node_qcount && rank) {
      // The count atomics return each entry's rank among its leaf's entries,
      // which places it in the leaf's list without another atomic.  They and
      // the list atomic are all issued before the first result is used, so
      // the flush waits on one round trip.
      constexpr uint32_t kPer = kEmitCap / kWave;
      uint32_t r[kPer];
#pragma unroll
      for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t e = lane + k * kWave;
        r[k] = e < n ? atomicAdd(node_qcount + buf[e], 1u) : 0u;
      }
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(total, (unsigned long long)n);
      base = uniform64(base);
#pragma unroll
      for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t e = lane + k * kWave;
        const uint64_t at = base + e;
        if (e < n && at < cap) {
          lq[at] = make_uint2(buf[e], q);
          rank[at] = r[k];
        }
      }
    } else {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(total, (unsigned long long)n);
      base = uniform64(base);
      for (uint32_t e = lane; e < n; e += kWave) {
        const uint64_t at = base + e;
        const uint32_t leaf = buf[e];
        if (at < cap) lq[at] = make_uint2(leaf, q);
        if (node_qcount) atomicAdd(node_qcount + leaf, 1u);
      }
    }
    n = 0;
    __syncthreads();  // the buffer is reused only after every lane has read it
  }
  __device__ uint64_t wave_cand() {
    uint64_t s = cand;
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    return s;
  }
};

// The stack holds the first-child index of inner nodes already selected.
// Each step pops up to 64 of them (one per lane); a lane loads its node's
// nchild children (contiguous in breadth-first order), tests each, records the
// selected leaves and pushes the first-child index of each selected inner
// child, lane by lane in order.  A step therefore costs one dependent memory
// access and handles up to 64 * nchild nodes.  Depth never decreases from the
// bottom of the stack to its top, so each depth keeps at most nchild * 64
// entries and a tree of depth D needs (D + 1) * nchild * 64 slots; the step
// size is also cut so a push can never pass the end.
//
// With a start table (quadtrees, gat_tree_start) the walk skips the levels
// above the table: the candidate blocks (those within one block of the
// rectangle's extent) are taken 64 at a time, each lane loading its block's
// owner and testing it once (an owner covering several blocks is taken at the
// first of them inside the candidate range); selected leaf owners are emitted
// and selected inner owners pushed, at most 64 entries of one depth, and the
// walk runs until the stack is empty before the next 64 blocks.  On 1M
// trajectories / 10k rectangles a level-7 table takes the walk from 125k to
// 67k steps.
struct StartTable {
  const uint32_t* owner;
  uint32_t level;
  float xmin, ymax, block;
  const uint2* tasks;  // (query, chunk | last << 31) per walk task, or null
  const uint32_t* ntasks;  // tasks, and the order entries they stand for
  uint32_t tasks_cap;
};

// Column (or row) bits of a Morton index: v's bits go to the even positions.
__device__ __forceinline__ uint32_t morton_spread(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Block of a coordinate offset v (in blocks), moved by d and clamped to 0..nb-1.
__device__ __forceinline__ uint32_t block_of(float v, int d, uint32_t nb) {
  const float t = floorf(v) + float(d);
  return uint32_t(fminf(fmaxf(t, 0.0f), float(nb - 1)));
}

template <int NCH>
__global__ __launch_bounds__(kWave) void tree_select_kernel(
    const Node* __restrict__ nodes, uint32_t nnodes, StartTable st, const float4* __restrict__ boxes,
    const uint32_t* __restrict__ order, uint32_t q0, uint32_t nq, uint32_t* __restrict__ node_qcount,
    uint64_t* __restrict__ cand, uint2* __restrict__ lq, uint32_t* __restrict__ lq_rank, uint64_t lq_cap,
    unsigned long long* __restrict__ lq_total,
    uint32_t* __restrict__ status) {
  __shared__ uint32_t stack[GAT_TREE_STACK];
  __shared__ uint32_t emitted[kEmitCap];
  const uint32_t lane = threadIdx.x;
  const uint32_t ntask = st.tasks ? min(st.ntasks[0], st.tasks_cap) : 0u;
  const uint32_t nskip = st.tasks ? min(st.ntasks[1], nq) : 0u;
  const uint32_t nwork = ntask + (nq - nskip);
  for (uint32_t i = blockIdx.x; i < nwork; i += gridDim.x) {
    uint32_t q, chunk = 0;
    bool last = true;
    if (i < ntask) {
      const uint2 t = st.tasks[i];
      q = uniform(t.x);
      chunk = uniform(t.y & 0x7FFFFFFFu);
      last = (uniform(t.y) >> 31) != 0u;
    } else {
      const uint32_t j = nskip + (i - ntask);
      q = order ? uniform(order[j]) : q0 + j;
    }
    if (q - q0 >= nq) continue;  // an entry outside the batch (uniform)
    const float4 box = boxes[q];
    LeafEmitter out{emitted, 0, q, lane, lq, lq_cap, lq_total, node_qcount, lq_rank, 0};
    // NaN and inverted rectangles take the walk from the root, in their
    // chunk-0 task.
    const bool tabled = st.owner != nullptr && box.x <= box.z && box.y <= box.w;
    if (!tabled && chunk != 0) continue;
    uint32_t top = 0, c0 = 0, r0 = 0, bw = 1, ncand = 0, tfirst = 0, tend = 0;
    if (tabled) {
      const uint32_t nb = 1u << st.level;
      const float inv = 1.0f / st.block;
      c0 = block_of((box.x - st.xmin) * inv, -1, nb);
      const uint32_t c1 = block_of((box.z - st.xmin) * inv, 1, nb);
      r0 = block_of((st.ymax - box.w) * inv, -1, nb);
      const uint32_t r1 = block_of((st.ymax - box.y) * inv, 1, nb);
      bw = c1 - c0 + 1;
      ncand = bw * (r1 - r0 + 1);
      // This task's blocks: its chunk of 64, or every block from it on.
      const uint32_t first = chunk < ncand / kWave + 1u ? chunk * kWave : ncand;
      const uint32_t end = last ? ncand : min(ncand, first + kWave);
      tfirst = first;
      tend = end;
    } else {
      const Node root = nodes[0];
      const bool sel = node_selected(root, box);
      if (root.first_child < 0) {
        out.push(sel && root.count > 0 && lane == 0, 0u, root.count);
      } else if (sel && uint32_t(root.first_child) + NCH <= nnodes) {
        if (lane == 0) stack[0] = uint32_t(root.first_child);
        top = 1;
      }
    }
    __syncthreads();
    bool overflow = false;
    for (uint32_t tb = tfirst;; tb += kWave) {
      if (tabled) {
        if (tb >= tend || overflow) break;
        const uint32_t b = tb + lane;
        bool emit = false, push = false;
        uint32_t o = 0, cnt = 0, first = 0;
        if (b < tend) {
          const uint32_t r = r0 + b / bw, c = c0 + b % bw;
          const uint32_t w = st.owner[morton_spread(c) | (morton_spread(r) << 1)];
          o = w & ((1u << 27) - 1u);
          const uint32_t lv = w >> 27;
          const uint32_t span = lv <= st.level ? 1u << (st.level - lv) : 1u;
          const uint32_t fr = max(r & ~(span - 1u), r0), fcol = max(c & ~(span - 1u), c0);
          if (fr == r && fcol == c && o < nnodes) {
            const Node n = nodes[o];
            if (node_selected(n, box)) {
              if (n.first_child < 0) {
                emit = n.count > 0;
                cnt = n.count;
              } else if (uint32_t(n.first_child) + NCH <= nnodes) {
                push = true;
                first = uint32_t(n.first_child);
              }
            }
          }
        }
        out.push(emit, o, cnt);
        out.maybe_flush();
        const uint64_t pm = __ballot(push);
        if (push) stack[lanes_below(pm)] = first;  // the stack is empty here
        top = uint32_t(__popcll(pm));
        __syncthreads();
      }
      while (top > 0) {
        uint32_t m = top < kWave ? top : kWave;
        const uint32_t room = (GAT_TREE_STACK - top) / (NCH - 1);
        if (m > room) m = room;
        if (m == 0) {
          if (lane == 0) atomicOr(status, GAT_STATUS_STACK);
          overflow = true;
          break;
        }
        top -= m;
        const bool live = lane < m;
        const uint32_t fc = live ? stack[top + lane] : 0u;
        __syncthreads();  // every lane has read its entry before pushes reuse the slots
        // Every child's record is loaded before the first push: a push may
        // flush (barrier, atomics, stores), which the later loads would
        // otherwise wait behind, one memory round trip per child.
        Node kid[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const uint32_t idx = fc + uint32_t(c);
          kid[c].first_child = -1;
          kid[c].flags = 0;
          if (live && idx < nnodes) kid[c] = nodes[idx];
        }
        uint32_t grand[NCH];
        bool inner[NCH];
        uint32_t below = 0, pushed = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          const uint32_t idx = fc + uint32_t(c);
          bool emit = false;
          uint32_t cnt = 0;
          inner[c] = false;
          grand[c] = 0;
          if (live && idx < nnodes) {
            const Node& n = kid[c];
            if (node_selected(n, box)) {
              if (n.first_child < 0) {
                emit = n.count > 0;
                cnt = n.count;
              } else if (uint32_t(n.first_child) + NCH <= nnodes) {
                inner[c] = true;
                grand[c] = uint32_t(n.first_child);
              }
            }
          }
          out.push(emit, idx, cnt);
          out.maybe_flush();
          const uint64_t dm = __ballot(inner[c]);
          below += lanes_below(dm);
          pushed += uint32_t(__popcll(dm));
        }
        // Lane-major order: a lane's selected inner children follow every lower lane's.
        uint32_t at = top + below;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
          if (inner[c]) stack[at++] = grand[c];
        top += pushed;
        __syncthreads();
      }
      if (!tabled) break;
    }
    out.flush();
    const uint64_t total = out.wave_cand();
    if (lane == 0) {
      if (st.tasks)
        atomicAdd(reinterpret_cast<unsigned long long*>(cand + q), (unsigned long long)total);
      else
        cand[q] = total;
    }
  }
}

// FSG::cellRange (FSG.cpp:183-186): the monotone float formula, clamped to the grid.
__device__ __forceinline__ int clamp_cell(float v, int hi) {
  if (!(v > -1.0f)) return 0;  // also NaN
  if (!(v < float(hi))) return hi - 1;
  const int c = int(v);
  return c < 0 ? 0 : (c > hi - 1 ? hi - 1 : c);
}

// Flat-grid selection (FSG.cpp:204-312): a rectangle visits every cell of its
// column range g1..g2 and row range g3..g4; cells[] holds one leaf record per
// cell (row-major), 64 cells per step.
__global__ __launch_bounds__(kWave) void grid_select_kernel(
    const Node* __restrict__ cells, int nx, int ny, float xmin, float ymax, float cs,
    const float4* __restrict__ boxes, const uint32_t* __restrict__ order, uint32_t q0, uint32_t nq,
    uint32_t* __restrict__ node_qcount, uint64_t* __restrict__ cand, uint2* __restrict__ lq,
    uint32_t* __restrict__ lq_rank, uint64_t lq_cap,
    unsigned long long* __restrict__ lq_total) {
  __shared__ uint32_t emitted[kEmitCap];
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = blockIdx.x; i < nq; i += gridDim.x) {
    const uint32_t q = order ? uniform(order[i]) : q0 + i;
    if (q - q0 >= nq) continue;
    const float4 box = boxes[q];
    LeafEmitter out{emitted, 0, q, lane, lq, lq_cap, lq_total, node_qcount, lq_rank, 0};
    const int g1 = clamp_cell((box.x - xmin) / cs, nx), g2 = clamp_cell((box.z - xmin) / cs, nx);
    const int g3 = clamp_cell((ymax - box.w) / cs, ny), g4 = clamp_cell((ymax - box.y) / cs, ny);
    if (g1 <= g2 && g3 <= g4) {
      const uint32_t w = uint32_t(g2 - g1 + 1);
      const uint32_t ncell = w * uint32_t(g4 - g3 + 1);  // <= nx * ny <= 2^30
      // The next step's counts load before this step's push (which may
      // flush: barrier, atomics, stores, that a later load would wait behind).
      auto cell_of = [&](uint32_t c) { return uint32_t(g3 + int(c / w)) * uint32_t(nx) + uint32_t(g1 + int(c % w)); };
      uint32_t idx = lane < ncell ? cell_of(lane) : 0u;
      uint32_t cnt = lane < ncell ? cells[idx].count : 0u;
      for (uint32_t base = 0; base < ncell; base += kWave) {
        const uint32_t c = base + lane, cn = c + kWave;
        const uint32_t nidx = cn < ncell ? cell_of(cn) : 0u;
        const uint32_t ncnt = cn < ncell ? cells[nidx].count : 0u;
        out.push(c < ncell && cnt > 0, idx, cnt);
        out.maybe_flush();
        idx = nidx;
        cnt = ncnt;
      }
    }
    out.flush();
    const uint64_t total = out.wave_cand();
    if (lane == 0) cand[q] = total;
  }
}

// ------------------------------------------------- implicit quadtree selection
// MortonGrid (MortonGrid.cpp:8-13, 238-268, 355-393): the same quadtree over
// Morton ranges as the GT-quadtree, kept without node records.  Node (level l,
// id) covers cells [id * 4^(L-l), (id+1) * 4^(L-l)) and its box runs from its
// first cell's corner to its last cell's (generateMBBfromNode), computed here
// from the cells' Morton-decoded columns and rows with the float formula of
// the cell boxes (MortonGrid.cpp:41-46).  Each cell c has a word pair
// cells[c] = (code, store start of c), and cells[4^L] = (-, total points).
// code = (rank << 4) | level for the leaf whose first cell is c (rank: its
// position among the leaves in Morton order), and a value naming no level
// where no leaf starts; a node at level l with first cell c is a leaf exactly
// when code names level l (leaves are aligned Morton blocks, so the leaf
// holding a node's first cell starts there).  A leaf's points are the store
// starts of its first cell and of the cell after it, two loads issued
// together.  Selected leaves are emitted by rank, so the leaf-major kernels
// read their store runs from the leaf records leaves[rank].
//
// Stack entries are (first cell << 4) | level of selected inner nodes; the
// traversal otherwise follows tree_select_kernel: the root is taken as it is,
// every other node when BoxIntersect(rectangle, box) != 0, 64 entries a step.
__device__ __forceinline__ uint32_t morton_compact(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0F0F0F0Fu;
  v = (v | (v >> 4)) & 0x00FF00FFu;
  v = (v | (v >> 8)) & 0x0000FFFFu;
  return v;
}

struct CellGeometry {
  float xmin, ymax, cs;
  // Box of the Morton block [first, first + span): the first cell's xmin and
  // ymax, the last cell's xmax and ymin (rows count down from ymax).
  __device__ __forceinline__ float4 block_box(uint32_t first, uint32_t span) const {
    const uint32_t last = first + span - 1u;
    const uint32_t c0 = morton_compact(first), r0 = morton_compact(first >> 1);
    const uint32_t c1 = morton_compact(last), r1 = morton_compact(last >> 1);
    return make_float4(__fadd_rn(xmin, __fmul_rn(cs, float(c0))), __fsub_rn(ymax, __fmul_rn(cs, float(r1 + 1u))),
                       __fadd_rn(xmin, __fmul_rn(cs, float(c1 + 1u))), __fsub_rn(ymax, __fmul_rn(cs, float(r0))));
  }
};

constexpr uint32_t kMaxImplicitLevel = 14;  // first cell < 2^28 fits a stack entry

__global__ __launch_bounds__(kWave) void morton_select_kernel(
    const uint2* __restrict__ cells, uint32_t total_level, uint32_t nleaves,
    CellGeometry geo, const float4* __restrict__ boxes, const uint32_t* __restrict__ order, uint32_t q0, uint32_t nq,
    uint32_t* __restrict__ node_qcount, uint64_t* __restrict__ cand, uint2* __restrict__ lq,
    uint32_t* __restrict__ lq_rank, uint64_t lq_cap,
    unsigned long long* __restrict__ lq_total, uint32_t* __restrict__ status) {
  __shared__ uint32_t stack[GAT_TREE_STACK];
  __shared__ uint32_t emitted[kEmitCap];
  constexpr int NCH = 4;
  const uint32_t lane = threadIdx.x;
  const uint32_t L = total_level;
  for (uint32_t i = blockIdx.x; i < nq; i += gridDim.x) {
    const uint32_t q = order ? uniform(order[i]) : q0 + i;
    if (q - q0 >= nq) continue;  // an order entry outside the batch (uniform)
    const float4 box = boxes[q];
    LeafEmitter out{emitted, 0, q, lane, lq, lq_cap, lq_total, node_qcount, lq_rank, 0};
    uint32_t top = 0;
    {
      const uint2 w = cells[0];
      if ((w.x & 15u) == 0u) {  // the root is a leaf: taken untested
        const uint32_t rank = w.x >> 4;
        const uint32_t cnt = rank < nleaves ? cells[1u << (2 * L)].y - w.y : 0u;
        out.push(lane == 0 && cnt > 0, rank, cnt);
      } else {
        if (lane == 0) stack[0] = 0u;  // first cell 0, level 0
        top = 1;
      }
    }
    __syncthreads();
    while (top > 0) {
      uint32_t m = top < kWave ? top : kWave;
      const uint32_t room = (GAT_TREE_STACK - top) / (NCH - 1);
      if (m > room) m = room;
      if (m == 0) {
        if (lane == 0) atomicOr(status, GAT_STATUS_STACK);
        break;
      }
      top -= m;
      const bool live = lane < m;
      const uint32_t e = live ? stack[top + lane] : 0u;
      __syncthreads();  // every lane has read its entry before pushes reuse the slots
      const uint32_t first = e >> 4, cl = (e & 15u) + 1u;  // children's level
      const bool deeper = live && cl <= L;
      const uint32_t shift = deeper ? 2u * (L - cl) : 0u;
      // Every selected child's words are loaded before the first push (a
      // push may flush: barrier, atomics, stores, which later loads would
      // wait behind).
      bool hit[NCH];
      uint2 cw[NCH];
      uint32_t cend[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const uint32_t fk = first + (uint32_t(c) << shift);
        hit[c] = deeper && box_intersect(box, geo.block_box(fk, 1u << shift));
        cw[c] = make_uint2(0xFFFFFFFFu, 0u);
        cend[c] = 0;
        if (hit[c]) {
          cw[c] = cells[fk];
          cend[c] = cells[fk + (1u << shift)].y;
        }
      }
      uint32_t grand[NCH];
      bool inner[NCH];
      uint32_t below = 0, pushed = 0;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const uint32_t fk = first + (uint32_t(c) << shift);
        bool emit = false;
        uint32_t cnt = 0, rank = 0;
        inner[c] = false;
        grand[c] = 0;
        if (hit[c]) {
          const uint2 w = cw[c];
          const uint32_t end = cend[c];
          if ((w.x & 15u) == cl) {
            rank = w.x >> 4;
            if (rank < nleaves) {
              cnt = end - w.y;
              emit = cnt > 0;
            }
          } else if (cl < L) {
            inner[c] = true;
            grand[c] = (fk << 4) | cl;
          }
        }
        out.push(emit, rank, cnt);
        out.maybe_flush();
        const uint64_t dm = __ballot(inner[c]);
        below += lanes_below(dm);
        pushed += uint32_t(__popcll(dm));
      }
      uint32_t at = top + below;
#pragma unroll
      for (int c = 0; c < NCH; ++c)
        if (inner[c]) stack[at++] = grand[c];
      top += pushed;
      __syncthreads();
    }
    out.flush();
    const uint64_t total = out.wave_cand();
    if (lane == 0) cand[q] = total;
  }
}

// ---------------------------------------------------------------- scans
// Reduce-then-scan over tiles of 4096 elements: per-tile sums, then a
// per-tile scan that takes its carry from the sums of the tiles before it.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

struct V2 {
  uint32_t x, y;
};
__device__ __forceinline__ uint64_t vadd(uint64_t a, uint64_t b) { return a + b; }
__device__ __forceinline__ V2 vadd(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ uint64_t vzero(uint64_t) { return 0; }
__device__ __forceinline__ V2 vzero(V2) { return V2{0, 0}; }
__device__ __forceinline__ uint64_t vshfl_up(uint64_t v, int d) { return __shfl_up(v, d); }
__device__ __forceinline__ V2 vshfl_up(V2 v, int d) { return V2{__shfl_up(v.x, d), __shfl_up(v.y, d)}; }

struct LoadU64 {
  const uint64_t* in;
  using T = uint64_t;
  __device__ T operator()(uint64_t i) const { return in[i]; }
};
struct LoadLists {
  const uint32_t* count;
  uint32_t qb;
  using T = V2;
  __device__ T operator()(uint64_t i) const {
    const uint32_t c = count[i];
    return V2{c, (c + qb - 1) / qb};
  }
};

// Exclusive scan of one value per thread across the workgroup; returns the
// thread's prefix and sets *sum to the workgroup total.
template <class T>
__device__ T block_exclusive(T v, T* sum) {
  __shared__ T wave_sums[kScanBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  T inc = v;
  for (int d = 1; d < kWave; d <<= 1) {
    const T o = vshfl_up(inc, d);
    if (lane >= d) inc = vadd(inc, o);
  }
  if (lane == kWave - 1) wave_sums[w] = inc;
  __syncthreads();
  T before = vzero(v), all = vzero(v);
  for (int k = 0; k < kScanBlock / kWave; ++k) {
    if (k < w) before = vadd(before, wave_sums[k]);
    all = vadd(all, wave_sums[k]);
  }
  __syncthreads();
  *sum = all;
  T excl = vshfl_up(inc, 1);
  if (lane == 0) excl = vzero(v);
  return vadd(before, excl);
}

template <class L>
__global__ __launch_bounds__(kScanBlock) void scan_tiles_kernel(L load, uint64_t n, typename L::T* __restrict__ tiles) {
  using T = typename L::T;
  const uint64_t b = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  T s = vzero(T{});
  for (int k = 0; k < kScanItems; ++k)
    if (b + k < n) s = vadd(s, load(b + k));
  T all;
  block_exclusive(s, &all);
  if (threadIdx.x == 0) tiles[blockIdx.x] = all;
}

// Each workgroup sums the tile totals before its own for its carry (a few
// hundred L2-resident values), and the first also all of them for the
// total, so no separate launch scans the tile totals.
template <class L>
__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(L load, uint64_t n,
                                                                const typename L::T* __restrict__ tiles,
                                                                uint64_t ntiles, typename L::T* __restrict__ out,
                                                                typename L::T* __restrict__ total) {
  using T = typename L::T;
  T c = vzero(T{});
  for (uint64_t t = threadIdx.x; t < blockIdx.x; t += kScanBlock) c = vadd(c, tiles[t]);
  T carry;
  block_exclusive(c, &carry);
  if (blockIdx.x == 0) {
    T g = vzero(T{});
    for (uint64_t t = threadIdx.x; t < ntiles; t += kScanBlock) g = vadd(g, tiles[t]);
    T all;
    block_exclusive(g, &all);
    if (threadIdx.x == 0) *total = all;
  }
  const uint64_t b = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  T v[kScanItems];
  T s = vzero(T{});
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = b + k < n ? load(b + k) : vzero(T{});
    s = vadd(s, v[k]);
  }
  T all;
  T run = vadd(carry, block_exclusive(s, &all));
  for (int k = 0; k < kScanItems; ++k)
    if (b + k < n) {
      out[b + k] = run;
      run = vadd(run, v[k]);
    }
}

template <class L>
int launch_scan(L load, uint64_t n, typename L::T* tiles, typename L::T* out, typename L::T* total, hipStream_t s) {
  using T = typename L::T;
  static_assert(sizeof(T) == 8, "scan totals are two words");
  const uint64_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) return gat_fill_u32(reinterpret_cast<uint32_t*>(total), 0u, 2, s);
  scan_tiles_kernel<L><<<dim3{uint32_t(ntiles)}, dim3{kScanBlock}, 0, s>>>(load, n, tiles);
  GAT_LAUNCH_CHECK("scan_tiles_kernel");
  scan_apply_kernel<L><<<dim3{uint32_t(ntiles)}, dim3{kScanBlock}, 0, s>>>(load, n, tiles, ntiles, out, total);
  GAT_LAUNCH_CHECK("scan_apply_kernel");
  return GAT_OK;
}

// ---------------------------------------------------------------- grouping
// Entries are grouped by leaf: a list entry is the rectangle's id (4 bytes;
// the refine reads the rectangle and its output region from the batch's
// tables, a few hundred KB that stay in L2), a work item carries the leaf's
// store run and its slice of the leaf's list.
struct Item {
  uint32_t start, count, list, n;
};
static_assert(sizeof(Item) == 16, "item layout");

// Output bound of a work item for the run-indexed refine: every rectangle
// could take every point, rounded up to whole 128-byte lines (16 pairs);
// items past the total bound nothing.
struct LoadItemBounds {
  const Item* items;
  const uint32_t* item_total;
  uint64_t cap;
  using T = uint64_t;
  __device__ T operator()(uint64_t i) const {
    const uint64_t total = *item_total;
    if (i >= (total < cap ? total : cap)) return 0;
    const Item it = items[i];
    return (uint64_t(it.n) * it.count + 15) & ~uint64_t(15);
  }
};

__global__ void leaf_scatter_kernel(const uint2* __restrict__ lq, const unsigned long long* __restrict__ lq_total,
                                    uint64_t lq_cap, const Node* __restrict__ nodes, uint32_t nnodes,
                                    const uint2* __restrict__ node_off, const uint32_t* __restrict__ node_qcount,
                                    const uint32_t* __restrict__ lq_rank, uint32_t qb, uint32_t nqueries,
                                    uint32_t* __restrict__ qlist, uint64_t qlist_cap, Item* __restrict__ items,
                                    uint64_t items_cap) {
  const uint64_t total = *lq_total;
  const uint64_t n = total < lq_cap ? total : lq_cap;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint2 e = lq[i];
    if (e.x >= nnodes || e.y >= nqueries) continue;
    const uint32_t r = lq_rank[i];  // the entry's rank among its leaf's, from the selection
    const uint2 off = node_off[e.x];
    const uint64_t at = uint64_t(off.x) + r;
    if (at < qlist_cap) qlist[at] = e.y;
    if (r % qb == 0) {
      const uint64_t it = uint64_t(off.y) + r / qb;
      const uint32_t len = node_qcount[e.x];
      if (it < items_cap && r < len) {
        const uint32_t left = len - r;
        items[it] = Item{nodes[e.x].start, nodes[e.x].count, uint32_t(at), left < qb ? left : qb};
      }
    }
  }
}

// Direct form for small batches: each selected (leaf, query) entry becomes a
// refine task of the per-task kernel (range.hip), so nothing sized by the
// tree is touched.
__global__ void leaf_tasks_kernel(const uint2* __restrict__ lq, const unsigned long long* __restrict__ lq_total,
                                  uint64_t lq_cap, const Node* __restrict__ nodes, uint32_t nnodes,
                                  gat_range_task* __restrict__ tasks) {
  const uint64_t total = *lq_total;
  const uint64_t n = total < lq_cap ? total : lq_cap;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint2 e = lq[i];
    // An entry naming no leaf becomes an empty task.
    const bool ok = e.x < nnodes;
    tasks[i] = gat_range_task{ok ? nodes[e.x].start : 0u, ok ? nodes[e.x].count : 0u, e.y};
  }
}

// ---------------------------------------------------------------- refine
// One wave per work item, chunk by chunk (up to 256 leaf points, NS = 1..4
// slots of 64, the code specialised per NS so no per-slot branch runs): lane
// j holds list entry j of the item (rectangle, output region, id).
//   pass 1: for each rectangle, the hit masks of the chunk's slots straight
//           from the compare instructions (four compares ANDed per slot);
//           lane j keeps rectangle j's masks and hit count;
//   one vector atomic reserves every rectangle's output slots at once;
//   pass 2: for each rectangle with hits, its cached masks place every hit
//           (rank = set bits below the lane) and the pair is written.
// The work per (rectangle, slot) is a handful of VALU/SALU instructions; the
// kernel is bound by that instruction stream as much as by memory.
constexpr int kRefBlock = 256;
constexpr int kSlots = 4;
constexpr int kSpan = kWave * kSlots;
constexpr int kMaxQb = kWave;

__device__ __forceinline__ float readlane_f(float v, uint32_t j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(j)));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t j) {
  return uint32_t(__builtin_amdgcn_readlane(int(v), int(j)));
}

// Lanes whose point (x, y) lies in box, inclusive on every edge
// (MBB::pInBox, MBB.cpp:21-26); NaN points are in no box.
__device__ __forceinline__ uint64_t box_mask(const float4& b, float x, float y) {
  constexpr int kOLE = 5, kOGE = 3;  // ordered <=, >=
  return __builtin_amdgcn_fcmpf(x, b.z, kOLE) & __builtin_amdgcn_fcmpf(x, b.x, kOGE) &
         __builtin_amdgcn_fcmpf(y, b.w, kOLE) & __builtin_amdgcn_fcmpf(y, b.y, kOGE);
}

template <int NS, bool kPacked>
__device__ __forceinline__ void refine_chunk(const Store& store, uint64_t at0, uint32_t left, uint32_t nql,
                                             const float4& bj, uint32_t qj, uint64_t basej,
                                             uint32_t* __restrict__ qcount, int2* __restrict__ pairs,
                                             uint64_t capacity, uint32_t lane) {
  const float nan = __builtin_nanf("");
  float2 v[NS];
  uint32_t id[NS];
  uint32_t pid[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const uint32_t i = k * kWave + lane;
    v[k] = make_float2(nan, 0.0f);
    id[k] = 0;
    pid[k] = 0;
    if (i < left) {
      v[k] = store.xy[at0 + i];
      id[k] = store.id[at0 + i];
      if constexpr (!kPacked) pid[k] = store.pid[at0 + i];
    }
  }
  // Pass 1: lane j keeps rectangle j's masks (as 32-bit halves) and hits.
  uint32_t mlo[NS], mhi[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) mlo[k] = mhi[k] = 0;
  uint32_t myhits = 0;
  for (uint32_t j = 0; j < nql; ++j) {
    const float4 box = make_float4(readlane_f(bj.x, j), readlane_f(bj.y, j), readlane_f(bj.z, j),
                                   readlane_f(bj.w, j));
    uint32_t h = 0;
    const bool mine = lane == j;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint64_t m = box_mask(box, v[k].x, v[k].y);
      h += uint32_t(__popcll(m));
      mlo[k] = mine ? uint32_t(m) : mlo[k];
      mhi[k] = mine ? uint32_t(m >> 32) : mhi[k];
    }
    myhits = mine ? h : myhits;
  }
  // One atomic per rectangle with hits, all issued together.
  uint32_t slot = 0;
  if (myhits > 0) slot = atomicAdd(qcount + qj, myhits);
  const uint64_t firstj = basej + slot;
  const uint32_t pid_mask = (1u << store.pid_bits) - 1u;
  // Pass 2: place and write every hit.
  for (uint32_t j = 0; j < nql; ++j) {
    if (readlane_u(myhits, j) == 0) continue;
    const uint64_t first =
        uint64_t(readlane_u(uint32_t(firstj), j)) | (uint64_t(readlane_u(uint32_t(firstj >> 32), j)) << 32);
    uint32_t before = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint32_t lo = readlane_u(mlo[k], j), hi = readlane_u(mhi[k], j);
      const uint64_t m = uint64_t(lo) | (uint64_t(hi) << 32);
      const uint64_t at = first + before + lanes_below(m);
      if (((m >> lane) & 1) && at < capacity) {
        int2 p;
        if constexpr (kPacked) {
          p = make_int2(int(id[k] >> store.pid_bits), int(id[k] & pid_mask));
        } else {
          p = make_int2(int(id[k]), int(pid[k]));
        }
        // A cached store: runs of different items that share a line can
        // meet in the L2 before it is written back (faster here than the
        // streaming store the per-query scan uses).
        gat::store_pair(pairs + at, p);
      }
      before += uint32_t(__popcll(m));
    }
  }
}

template <bool kPacked>
__global__ __launch_bounds__(kRefBlock) void leaf_refine_kernel(
    Store store, const Item* __restrict__ items, const uint32_t* __restrict__ item_total, uint64_t items_cap,
    const uint32_t* __restrict__ qlist, uint64_t qlist_cap, const float4* __restrict__ boxes,
    const uint64_t* __restrict__ qbase, uint32_t nqueries, uint32_t* __restrict__ qcount, int2* __restrict__ pairs,
    uint64_t capacity) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = uniform((blockIdx.x * kRefBlock + threadIdx.x) / kWave);
  const uint64_t nwaves = uint64_t(gridDim.x) * (kRefBlock / kWave);
  const uint64_t total = *item_total;
  const uint64_t nitems = total < items_cap ? total : items_cap;
  const float nan = __builtin_nanf("");
  Item next = wave < nitems ? items[wave] : Item{0, 0, 0, 0};
  for (uint64_t t = wave; t < nitems; t += nwaves) {
    const uint32_t start = uniform(next.start), count = uniform(next.count);
    const uint32_t list = uniform(next.list), nql = uniform(next.n);
    if (t + nwaves < nitems) next = items[t + nwaves];
    if (uint64_t(start) + count > store.npoints || nql == 0 || nql > kMaxQb || uint64_t(list) + nql > qlist_cap)
      continue;
    // Lane j: rectangle j of the item.  Invalid ones get a NaN box (no hits).
    uint32_t qj = 0xFFFFFFFFu;
    float4 bj = make_float4(nan, nan, nan, nan);
    uint64_t basej = 0;
    if (lane < nql) {
      const uint32_t q = qlist[uint64_t(list) + lane];
      if (q < nqueries) {
        qj = q;
        bj = boxes[q];
        basej = qbase[q];
      }
    }
    for (uint32_t base = 0; base < count; base += kSpan) {
      const uint32_t left = count - base;
      const uint64_t at0 = uint64_t(start) + base;
      const uint32_t ns = left >= uint32_t(kSpan) ? uint32_t(kSlots) : (left + kWave - 1) / kWave;
      switch (ns) {
        case 1:
          refine_chunk<1, kPacked>(store, at0, left, nql, bj, qj, basej, qcount, pairs, capacity, lane);
          break;
        case 2:
          refine_chunk<2, kPacked>(store, at0, left, nql, bj, qj, basej, qcount, pairs, capacity, lane);
          break;
        case 3:
          refine_chunk<3, kPacked>(store, at0, left, nql, bj, qj, basej, qcount, pairs, capacity, lane);
          break;
        default:
          refine_chunk<4, kPacked>(store, at0, left, nql, bj, qj, basej, qcount, pairs, capacity, lane);
      }
    }
  }
}

// ------------------------------------------------------- run-indexed refine
// The same join with the output laid out by work item instead of by query:
// work item t owns the region [item_base[t], item_base[t + 1]) of `pairs`
// (an exclusive scan of its bound, rectangles x leaf points rounded up to a
// 128-byte line), its hits fill the region from the start, rectangle after
// rectangle, and list entry e of the item records its run (query, hit count,
// first pair) in runs[e].  A region is written by one wave and starts on a
// line, so no line is shared between waves; in per-query regions every run's
// end lines are.  Nothing is reserved with atomics (one counter for every item
// would serialise the batch).  An item whose leaf spans several chunks first
// counts each rectangle's hits over all of them, so it still makes one run per
// rectangle.
struct Run {
  uint32_t q, count;
  uint64_t first;
};
static_assert(sizeof(Run) == sizeof(gat_range_run), "run record layout");

template <int NS, bool kPacked>
__device__ __forceinline__ void load_chunk(const Store& store, uint64_t at0, uint32_t left, uint32_t lane,
                                           float2 (&v)[NS], uint32_t (&id)[NS], uint32_t (&pid)[NS]) {
  const float nan = __builtin_nanf("");
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const uint32_t i = k * kWave + lane;
    v[k] = make_float2(nan, 0.0f);
    id[k] = 0;
    pid[k] = 0;
    if (i < left) {
      v[k] = store.xy[at0 + i];
      id[k] = store.id[at0 + i];
      if constexpr (!kPacked) pid[k] = store.pid[at0 + i];
    }
  }
}

// Lane j receives rectangle j's hit masks (32-bit halves) and hit count.
template <int NS>
__device__ __forceinline__ uint32_t mask_pass(const float2 (&v)[NS], uint32_t nql, const float4& bj, uint32_t lane,
                                              uint32_t (&mlo)[NS], uint32_t (&mhi)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) mlo[k] = mhi[k] = 0;
  uint32_t myhits = 0;
  for (uint32_t j = 0; j < nql; ++j) {
    const float4 box = make_float4(readlane_f(bj.x, j), readlane_f(bj.y, j), readlane_f(bj.z, j),
                                   readlane_f(bj.w, j));
    uint32_t h = 0;
    const bool mine = lane == j;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint64_t m = box_mask(box, v[k].x, v[k].y);
      h += uint32_t(__popcll(m));
      mlo[k] = mine ? uint32_t(m) : mlo[k];
      mhi[k] = mine ? uint32_t(m >> 32) : mhi[k];
    }
    myhits = mine ? h : myhits;
  }
  return myhits;
}

// Writes every hit; rectangle j's hits of this chunk start at lane j's `first`.
template <int NS, bool kPacked>
__device__ __forceinline__ void write_pass(const Store& store, const uint32_t (&id)[NS], const uint32_t (&pid)[NS],
                                           uint32_t nql, uint32_t myhits, uint64_t first_j,
                                           const uint32_t (&mlo)[NS], const uint32_t (&mhi)[NS],
                                           int2* __restrict__ pairs, uint64_t capacity, uint32_t lane) {
  const uint32_t pid_mask = (1u << store.pid_bits) - 1u;
  for (uint32_t j = 0; j < nql; ++j) {
    if (readlane_u(myhits, j) == 0) continue;
    const uint64_t first =
        uint64_t(readlane_u(uint32_t(first_j), j)) | (uint64_t(readlane_u(uint32_t(first_j >> 32), j)) << 32);
    uint32_t before = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint64_t m = uint64_t(readlane_u(mlo[k], j)) | (uint64_t(readlane_u(mhi[k], j)) << 32);
      const uint64_t at = first + before + lanes_below(m);
      if (((m >> lane) & 1) && at < capacity) {
        int2 p;
        if constexpr (kPacked) {
          p = make_int2(int(id[k] >> store.pid_bits), int(id[k] & pid_mask));
        } else {
          p = make_int2(int(id[k]), int(pid[k]));
        }
        gat::store_pair(pairs + at, p);
      }
      before += uint32_t(__popcll(m));
    }
  }
}

// Per-wave staging ring of the run-indexed refine: an item's pairs are
// appended here in output order and leave in whole blocks of 128 pairs
// (1 KB, one 16-byte store per lane, line-aligned because the item's region
// is), so a block costs one store instruction instead of one per rectangle
// and slot.  The stores are nontemporal: nothing in the launch reads the
// pairs back (1M trajectories x 10k rectangles: 1.44 ms per refine, 1.61 ms
// with plain stores).
constexpr uint32_t kRing = 512;   // pairs per wave (4 KB)
constexpr uint32_t kFlush = 128;  // pairs per flush: 2 per lane

// Orders one lane's LDS writes before other lanes' reads of them and the
// reads before later writes to the same slots (LDS executes a wave's
// instructions in order; this keeps the compiler from reordering them).
__device__ __forceinline__ void lds_order() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// write_pass through the ring: the item's pairs fill [base, base + total)
// back to back, rectangle j's from base + (first_j - base).
template <int NS, bool kPacked>
__device__ __forceinline__ void write_pass_staged(const Store& store, const uint32_t (&id)[NS],
                                                  const uint32_t (&pid)[NS], uint32_t nql, uint32_t myhits,
                                                  uint64_t base, const uint32_t (&mlo)[NS],
                                                  const uint32_t (&mhi)[NS], int2* __restrict__ pairs,
                                                  uint64_t capacity, uint32_t lane, int2* ring) {
  const uint32_t pid_mask = (1u << store.pid_bits) - 1u;
  uint32_t fill = 0, flushed = 0;  // pairs appended / stored, relative to base
  auto flush_block = [&]() __attribute__((always_inline)) {
    lds_order();
    const int4 v = reinterpret_cast<const int4*>(ring)[((flushed & (kRing - 1)) >> 1) + lane];
    const uint64_t at = base + flushed + 2 * lane;
    if (at + 1 < capacity) gat::store_pair2(pairs + at, v);
    flushed += kFlush;
    lds_order();
  };
  for (uint32_t j = 0; j < nql; ++j) {
    const uint32_t h = readlane_u(myhits, j);
    if (h == 0) continue;
    uint32_t before = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const uint64_t m = uint64_t(readlane_u(mlo[k], j)) | (uint64_t(readlane_u(mhi[k], j)) << 32);
      if ((m >> lane) & 1) {
        int2 p;
        if constexpr (kPacked) {
          p = make_int2(int(id[k] >> store.pid_bits), int(id[k] & pid_mask));
        } else {
          p = make_int2(int(id[k]), int(pid[k]));
        }
        ring[(fill + before + lanes_below(m)) & (kRing - 1)] = p;
      }
      before += uint32_t(__popcll(m));
    }
    fill += h;
    // Unstored pairs stay under kFlush after each rectangle and a rectangle
    // adds at most 256, so appends never reach a slot still waiting.
    while (fill - flushed >= kFlush) flush_block();
  }
  // The tail: fewer than 128 pairs, two per lane, padded with (-1, -1) to
  // the end of its 128-byte line.  The item's region starts on a line and
  // holds its bound rounded up to whole lines, so the padding stays inside
  // it; every line of the region is then written whole, and no partial line
  // reaches HBM (a partial line costs a read of the line there).
  lds_order();
  const uint32_t rem = fill - flushed;
  const uint32_t padded = (rem + 15u) & ~15u;
  const uint32_t i = 2 * lane;
  const uint64_t at = base + flushed + i;
  if (i < padded) {
    int4 v = reinterpret_cast<const int4*>(ring)[((flushed & (kRing - 1)) >> 1) + lane];
    if (i >= rem) v.x = v.y = -1;
    if (i + 1 >= rem) v.z = v.w = -1;
    if (at + 1 < capacity) gat::store_pair2(pairs + at, v);
  }
  lds_order();
}

template <int NS, bool kPacked>
__device__ __forceinline__ uint32_t count_chunk(const Store& store, uint64_t at0, uint32_t left, uint32_t nql,
                                                const float4& bj, uint32_t lane) {
  float2 v[NS];
  uint32_t id[NS], pid[NS], mlo[NS], mhi[NS];
  load_chunk<NS, kPacked>(store, at0, left, lane, v, id, pid);
  return mask_pass<NS>(v, nql, bj, lane, mlo, mhi);
}

template <int NS, bool kPacked>
__device__ __forceinline__ uint32_t refine_run_chunk(const Store& store, uint64_t at0, uint32_t left, uint32_t nql,
                                                     const float4& bj, uint64_t first_j, int2* __restrict__ pairs,
                                                     uint64_t capacity, uint32_t lane) {
  float2 v[NS];
  uint32_t id[NS], pid[NS], mlo[NS], mhi[NS];
  load_chunk<NS, kPacked>(store, at0, left, lane, v, id, pid);
  const uint32_t myhits = mask_pass<NS>(v, nql, bj, lane, mlo, mhi);
  write_pass<NS, kPacked>(store, id, pid, nql, myhits, first_j, mlo, mhi, pairs, capacity, lane);
  return myhits;
}

// Slots of 64 points a chunk of `left` (>= 1) remaining points spans, at most kSlots.
__device__ __forceinline__ uint32_t slots_for(uint32_t left) {
  return left >= uint32_t(kSpan) ? uint32_t(kSlots) : (left + kWave - 1) / kWave;
}

template <bool kPacked>
__device__ __forceinline__ uint32_t count_any(const Store& store, uint64_t at0, uint32_t left, uint32_t nql,
                                              const float4& bj, uint32_t lane) {
  switch (slots_for(left)) {
    case 1: return count_chunk<1, kPacked>(store, at0, left, nql, bj, lane);
    case 2: return count_chunk<2, kPacked>(store, at0, left, nql, bj, lane);
    case 3: return count_chunk<3, kPacked>(store, at0, left, nql, bj, lane);
    default: return count_chunk<4, kPacked>(store, at0, left, nql, bj, lane);
  }
}

template <bool kPacked>
__device__ __forceinline__ uint32_t refine_run_any(const Store& store, uint64_t at0, uint32_t left, uint32_t nql,
                                                   const float4& bj, uint64_t first_j, int2* __restrict__ pairs,
                                                   uint64_t capacity, uint32_t lane) {
  switch (slots_for(left)) {
    case 1: return refine_run_chunk<1, kPacked>(store, at0, left, nql, bj, first_j, pairs, capacity, lane);
    case 2: return refine_run_chunk<2, kPacked>(store, at0, left, nql, bj, first_j, pairs, capacity, lane);
    case 3: return refine_run_chunk<3, kPacked>(store, at0, left, nql, bj, first_j, pairs, capacity, lane);
    default: return refine_run_chunk<4, kPacked>(store, at0, left, nql, bj, first_j, pairs, capacity, lane);
  }
}

// Inclusive prefix sum over the wave's lanes.
__device__ __forceinline__ uint32_t wave_prefix(uint32_t x, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = uint32_t(__shfl_up(int(x), d, kWave));
    x += lane >= uint32_t(d) ? y : 0u;
  }
  return x;
}

struct RunOut {
  Run* runs;
  uint64_t runs_cap;
  const uint64_t* item_base;
};

// Places the item's runs (lane j: `hits` of rectangle j, query qj, list entry
// list + j) one after another from the start of its region and records them;
// false when the item has no hits.  *first_j: where lane j's pairs start.
// The per-query counts are summed from the records afterwards
// (run_counts_kernel): an atomic per rectangle here, 10k counters shared by
// every wave, held each wave's next loads behind it (gfx9 counts stores and
// atomics in vmcnt with the loads) and cost 0.13 ms of a 1.73 ms refine.
__device__ __forceinline__ bool place_item(uint32_t hits, uint32_t qj, uint32_t lane, uint32_t nql, uint32_t list,
                                           uint64_t base, const RunOut& o, uint64_t* first_j) {
  const uint32_t incl = wave_prefix(hits, lane);
  *first_j = base + (incl - hits);
  if (lane < nql && uint64_t(list) + lane < o.runs_cap) o.runs[uint64_t(list) + lane] = Run{qj, hits, *first_j};
  return readlane_u(incl, kWave - 1) > 0;
}

// An item whose leaf fits one chunk: masks once, then the writes.
template <int NS, bool kPacked>
__device__ __forceinline__ void single_chunk_item(const Store& store, uint32_t start, uint32_t count, uint32_t nql,
                                                  uint32_t list, uint64_t base, const float4& bj, uint32_t qj,
                                                  const RunOut& o, int2* __restrict__ pairs, uint64_t capacity,
                                                  uint32_t lane, int2* ring) {
  float2 v[NS];
  uint32_t id[NS], pid[NS], mlo[NS], mhi[NS];
  load_chunk<NS, kPacked>(store, start, count, lane, v, id, pid);
  const uint32_t hits = mask_pass<NS>(v, nql, bj, lane, mlo, mhi);
  uint64_t first_j = 0;
  if (!place_item(hits, qj, lane, nql, list, base, o, &first_j)) return;
  write_pass_staged<NS, kPacked>(store, id, pid, nql, hits, base, mlo, mhi, pairs, capacity, lane, ring);
}

// Eight waves per SIMD (at most 64 VGPRs, no spills): the kernel waits on
// load and store latency, and the eighth wave shortened the cfg2 refine from
// 1.81 to 1.74 ms; loading the next item's inputs ahead of the current item's
// stores instead needs about 100 VGPRs and ran slower (1.93 ms).
constexpr int kRunsBlock = 256;

template <bool kPacked>
__global__ __launch_bounds__(kRunsBlock, 8) void leaf_refine_runs_kernel(
    Store store, const Item* __restrict__ items, const uint32_t* __restrict__ item_total, uint64_t items_cap,
    const uint32_t* __restrict__ qlist, uint64_t qlist_cap, const float4* __restrict__ boxes, uint32_t nqueries,
    int2* __restrict__ pairs, uint64_t capacity, RunOut o) {
  __shared__ int4 rings[kRunsBlock / kWave][kRing / 2];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  int2* ring = reinterpret_cast<int2*>(rings[threadIdx.x / kWave]);
  const uint64_t wave = uniform((blockIdx.x * kRunsBlock + threadIdx.x) / kWave);
  const uint64_t nwaves = uint64_t(gridDim.x) * (kRunsBlock / kWave);
  const uint64_t total = *item_total;
  const uint64_t nitems = total < items_cap ? total : items_cap;
  const float nan = __builtin_nanf("");
  Item next = wave < nitems ? items[wave] : Item{0, 0, 0, 0};
  for (uint64_t t = wave; t < nitems; t += nwaves) {
    const uint32_t start = uniform(next.start), count = uniform(next.count);
    const uint32_t list = uniform(next.list), nql = uniform(next.n);
    if (t + nwaves < nitems) next = items[t + nwaves];
    if (uint64_t(start) + count > store.npoints || count == 0 || nql == 0 || nql > kMaxQb ||
        uint64_t(list) + nql > qlist_cap) {
      // Its records read as empty to the per-query counts.
      const uint32_t nrec = nql < uint32_t(kMaxQb) ? nql : uint32_t(kMaxQb);
      if (lane < nrec && uint64_t(list) + lane < o.runs_cap) o.runs[uint64_t(list) + lane] = Run{0xFFFFFFFFu, 0u, 0u};
      continue;
    }
    uint32_t qj = 0xFFFFFFFFu;
    float4 bj = make_float4(nan, nan, nan, nan);
    if (lane < nql) {
      const uint32_t q = qlist[uint64_t(list) + lane];
      if (q < nqueries) {
        qj = q;
        bj = boxes[q];
      }
    }
    const uint64_t base = o.item_base[t];
    if (count <= uint32_t(kSpan)) {
      switch (slots_for(count)) {
        case 1:
          single_chunk_item<1, kPacked>(store, start, count, nql, list, base, bj, qj, o, pairs, capacity, lane,
                                           ring);
          break;
        case 2:
          single_chunk_item<2, kPacked>(store, start, count, nql, list, base, bj, qj, o, pairs, capacity, lane,
                                           ring);
          break;
        case 3:
          single_chunk_item<3, kPacked>(store, start, count, nql, list, base, bj, qj, o, pairs, capacity, lane,
                                           ring);
          break;
        default:
          single_chunk_item<4, kPacked>(store, start, count, nql, list, base, bj, qj, o, pairs, capacity, lane,
                                           ring);
      }
      continue;
    }
    // A leaf of several chunks: every rectangle's hits over all of them first,
    // so each rectangle gets one run; then chunk by chunk into it.
    uint32_t hits = 0;
    for (uint32_t c = 0; c < count; c += kSpan)
      hits += count_any<kPacked>(store, uint64_t(start) + c, count - c, nql, bj, lane);
    uint64_t first_j = 0;
    if (!place_item(hits, qj, lane, nql, list, base, o, &first_j)) continue;
    for (uint32_t c = 0; c < count; c += kSpan)
      first_j += refine_run_any<kPacked>(store, uint64_t(start) + c, count - c, nql, bj, first_j, pairs, capacity,
                                         lane);
  }
}

// qcount[q] += the hits of every run record of query q.  The records cover
// the list entries [0, L) in order (L = the last work item's list end; the
// refine writes empty records for items it skips), so each workgroup sums a
// contiguous range of them, one 16-byte record per thread, into LDS counters
// for all nqueries rectangles (neighbouring leaves share few rectangles) and
// adds the non-zero ones to qcount; each wave adds its records' sum to
// *pair_total.  Walking the records item by item, a thread per item, took 76
// us on 1M trajectories / 10k rectangles; adding every record to qcount
// directly put 5.77M atomics on 10k counters (0.6 ms).
constexpr uint32_t kCountLdsQueries = 16384;  // 64 KB of LDS counters

template <bool kLds>
__global__ __launch_bounds__(256) void run_counts_kernel(const Item* __restrict__ items,
                                                         const uint32_t* __restrict__ item_total, uint64_t items_cap,
                                                         uint64_t list_cap, const Run* __restrict__ runs,
                                                         uint32_t nqueries, uint32_t* __restrict__ qcount,
                                                         unsigned long long* __restrict__ pair_total) {
  extern __shared__ uint32_t acc[];  // nqueries counters when kLds
  const uint64_t total = *item_total;
  const uint64_t nitems = total < items_cap ? total : items_cap;
  uint64_t nrec = 0;
  if (nitems > 0) {
    const Item last = items[nitems - 1];
    nrec = uint64_t(last.list) + last.n;
    if (nrec > list_cap) nrec = list_cap;
  }
  if constexpr (kLds) {
    for (uint32_t q = threadIdx.x; q < nqueries; q += blockDim.x) acc[q] = 0;
    __syncthreads();
  }
  const uint64_t per = (nrec + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = uint64_t(blockIdx.x) * per;
  const uint64_t hi = lo + per < nrec ? lo + per : nrec;
  uint64_t mine = 0;
  auto add = [&](const Run& r) {
    if (r.count == 0 || r.q >= nqueries) return;
    mine += r.count;
    if constexpr (kLds) {
      atomicAdd(acc + r.q, r.count);
    } else {
      atomicAdd(qcount + r.q, r.count);
    }
  };
  // Four records in flight per thread: the loop otherwise waits out one
  // load's latency per record.
  uint64_t e = lo + threadIdx.x;
  for (; e + 3 * blockDim.x < hi; e += 4 * blockDim.x) {
    const Run r0 = runs[e], r1 = runs[e + blockDim.x], r2 = runs[e + 2 * blockDim.x], r3 = runs[e + 3 * blockDim.x];
    add(r0);
    add(r1);
    add(r2);
    add(r3);
  }
  for (; e < hi; e += blockDim.x) add(runs[e]);
  if (pair_total) {
    for (int off = kWave / 2; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if ((threadIdx.x & (kWave - 1)) == 0 && mine) atomicAdd(pair_total, (unsigned long long)mine);
  }
  if constexpr (kLds) {
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nqueries; q += blockDim.x)
      if (acc[q]) atomicAdd(qcount + q, acc[q]);
  }
}

// Per-query digest of run-indexed pairs: one wave per run adds the run's
// wrapping sum of pair mixes to its query's entry (out must start at zero).
__global__ __launch_bounds__(256) void runs_digest_kernel(const int2* __restrict__ pairs, uint64_t npairs,
                                                          const Run* __restrict__ runs, uint64_t nruns,
                                                          uint32_t nqueries, unsigned long long* __restrict__ out) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t nw = uint64_t(gridDim.x) * (blockDim.x / kWave);
  for (uint64_t r = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave; r < nruns; r += nw) {
    const Run run = runs[r];
    if (run.q >= nqueries || run.first > npairs || run.count > npairs - run.first) continue;
    uint64_t sum = 0;
    for (uint64_t i = lane; i < run.count; i += kWave) sum += gat::pair_mix(pairs[run.first + i]);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(sum)), d, kWave));
      const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(sum >> 32)), d, kWave));
      sum += uint64_t(lo) | (uint64_t(hi) << 32);
    }
    if (lane == 0) atomicAdd(out + run.q, (unsigned long long)sum);
  }
}

// Sum of n uint32 values into *out (which must start at zero).
__global__ void sum_u32_kernel(const uint32_t* __restrict__ in, uint64_t n, unsigned long long* __restrict__ out) {
  uint64_t s = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    s += in[i];
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & (kWave - 1)) == 0 && s) atomicAdd(out, (unsigned long long)s);
}

// *total = in[0] + ... + in[n-1], one workgroup.
constexpr int kReduceBlock = 1024;
__global__ __launch_bounds__(kReduceBlock) void reduce_u64_kernel(const uint64_t* __restrict__ in, uint64_t n,
                                                                  uint64_t* __restrict__ total) {
  __shared__ uint64_t part[kReduceBlock / kWave];
  uint64_t s = 0;
  for (uint64_t i = threadIdx.x; i < n; i += kReduceBlock) s += in[i];
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kReduceBlock / kWave; ++w) t += part[w];
    *total = t;
  }
}

int grid_for(uint64_t work, int per_block, int cap_per_cu) {
  uint64_t blocks = (work + per_block - 1) / per_block;
  const uint64_t cap = uint64_t(gat::cu_count()) * cap_per_cu;
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  return int(blocks);
}

}  // namespace

extern "C" {

int gat_fill_u32(uint32_t* dst, uint32_t value, int64_t n, void* stream) {
  GAT_REQUIRE(n >= 0 && (n == 0 || dst), "gat_fill_u32: bad arguments");
  if (n == 0) return GAT_OK;
  fill_u32_kernel<<<dim3{uint32_t(grid_for(uint64_t(n), 256, 8))}, dim3{256}, 0, gat::as_stream(stream)>>>(
      dst, value, uint64_t(n));
  GAT_LAUNCH_CHECK("fill_u32_kernel");
  return GAT_OK;
}

int gat_fill_u32_spans(int32_t nspans, uint32_t* const* dst, const int64_t* n, uint32_t value, void* stream) {
  GAT_REQUIRE(nspans >= 0 && nspans <= kFillSpans && (nspans == 0 || (dst && n)), "gat_fill_u32_spans: bad arguments");
  FillSpans f{};
  uint64_t total = 0;
  for (int k = 0; k < kFillSpans; ++k) {
    if (k < nspans) {
      GAT_REQUIRE(n[k] >= 0 && (n[k] == 0 || dst[k]), "gat_fill_u32_spans: bad span");
      f.p[k] = dst[k];
      total += uint64_t(n[k]);
    }
    f.end[k] = total;
  }
  if (total == 0) return GAT_OK;
  fill_spans_kernel<<<dim3{uint32_t(grid_for(total, 256, 8))}, dim3{256}, 0, gat::as_stream(stream)>>>(f, value);
  GAT_LAUNCH_CHECK("fill_spans_kernel");
  return GAT_OK;
}

int64_t gat_scan_tiles(int64_t n) { return n <= 0 ? 1 : (n + kScanTile - 1) / kScanTile; }

int gat_scan_u64(const uint64_t* in, int64_t n, uint64_t* out, uint64_t* total, uint64_t* tiles, void* stream) {
  GAT_REQUIRE(n >= 0 && total && (n == 0 || in), "gat_scan_u64: bad arguments");
  GAT_REQUIRE(uint64_t(n) / kScanTile < UINT32_MAX, "gat_scan_u64: too many elements");
  if (!out) {
    // The total alone: one workgroup sums the input, one launch instead of three.
    reduce_u64_kernel<<<dim3{1}, dim3{kReduceBlock}, 0, gat::as_stream(stream)>>>(in, uint64_t(n), total);
    GAT_LAUNCH_CHECK("reduce_u64_kernel");
    return GAT_OK;
  }
  GAT_REQUIRE(tiles, "gat_scan_u64: bad arguments");
  return launch_scan(LoadU64{in}, uint64_t(n), tiles, out, total, gat::as_stream(stream));
}

int gat_scan_leaf_lists(const uint32_t* count, int64_t nnodes, int32_t qb, uint32_t* node_off, uint32_t* totals,
                        uint32_t* tiles, void* stream) {
  GAT_REQUIRE(nnodes >= 0 && qb > 0 && totals && tiles && (nnodes == 0 || (count && node_off)),
              "gat_scan_leaf_lists: bad arguments");
  GAT_REQUIRE(uint64_t(nnodes) / kScanTile < UINT32_MAX, "gat_scan_leaf_lists: too many nodes");
  return launch_scan(LoadLists{count, uint32_t(qb)}, uint64_t(nnodes), reinterpret_cast<V2*>(tiles),
                     reinterpret_cast<V2*>(node_off), reinterpret_cast<V2*>(totals), gat::as_stream(stream));
}

int gat_tree_select(const gat_tree_node* nodes, int64_t nnodes, int32_t nchild, const gat_tree_start* start,
                    const float* boxes, const uint32_t* order, int64_t q0, int64_t nq, uint32_t* node_qcount,
                    uint64_t* cand, uint32_t* leaf_query, uint32_t* lq_rank, int64_t lq_capacity,
                    uint64_t* lq_total, uint32_t* status, void* stream) {
  GAT_REQUIRE(nnodes >= 1 && nnodes < int64_t(INT32_MAX) && nodes, "gat_tree_select: bad tree");
  GAT_REQUIRE(nchild == 2 || nchild == 4, "gat_tree_select: nchild must be 2 or 4");
  StartTable st{nullptr, 0, 0.0f, 0.0f, 1.0f, nullptr, nullptr, 0};
  int64_t nwork = nq;
  if (start) {
    GAT_REQUIRE(nchild == 4 && start->owner_dev && start->level >= 1 && start->level <= 15 &&
                    start->block > 0.0f && std::isfinite(start->block) && std::isfinite(start->xmin) &&
                    std::isfinite(start->ymax),
                "gat_tree_select: bad start table");
    GAT_REQUIRE(!start->tasks_dev || (start->ntasks_dev && start->tasks_capacity >= 0 &&
                                      start->tasks_capacity <= int64_t(UINT32_MAX) - nq),
                "gat_tree_select: bad walk tasks");
    st = StartTable{start->owner_dev,
                    uint32_t(start->level),
                    start->xmin,
                    start->ymax,
                    start->block,
                    reinterpret_cast<const uint2*>(start->tasks_dev),
                    start->ntasks_dev,
                    uint32_t(start->tasks_capacity)};
    if (start->tasks_dev) nwork = nq + start->tasks_capacity;
  }
  GAT_REQUIRE(q0 >= 0 && nq >= 0 && q0 + nq <= int64_t(UINT32_MAX), "gat_tree_select: bad query range");
  GAT_REQUIRE(lq_capacity >= 0 && lq_total && status, "gat_tree_select: bad outputs");
  if (nq == 0 || nwork == 0) return GAT_OK;
  GAT_REQUIRE(boxes && cand && (lq_capacity == 0 || leaf_query), "gat_tree_select: null pointer");
  const int blocks = grid_for(uint64_t(nwork), 1, 40);
  const dim3 grid{uint32_t(blocks)}, block{kWave};
  hipStream_t s = gat::as_stream(stream);
  const auto* N = reinterpret_cast<const Node*>(nodes);
  const auto* B = reinterpret_cast<const float4*>(boxes);
  auto* L = reinterpret_cast<uint2*>(leaf_query);
  auto* T = reinterpret_cast<unsigned long long*>(lq_total);
  if (nchild == 4)
    tree_select_kernel<4><<<grid, block, 0, s>>>(N, uint32_t(nnodes), st, B, order, uint32_t(q0), uint32_t(nq),
                                                 node_qcount, cand, L, lq_rank, uint64_t(lq_capacity), T, status);
  else
    tree_select_kernel<2><<<grid, block, 0, s>>>(N, uint32_t(nnodes), st, B, order, uint32_t(q0), uint32_t(nq),
                                                 node_qcount, cand, L, lq_rank, uint64_t(lq_capacity), T, status);
  GAT_LAUNCH_CHECK("tree_select_kernel");
  return GAT_OK;
}

int gat_morton_select(const uint32_t* cell_words, int32_t total_level, int64_t nleaves, float xmin, float ymax,
                      float cell_size, const float* boxes, const uint32_t* order, int64_t q0, int64_t nq,
                      uint32_t* node_qcount, uint64_t* cand, uint32_t* leaf_query, uint32_t* lq_rank,
                      int64_t lq_capacity, uint64_t* lq_total, uint32_t* status, void* stream) {
  GAT_REQUIRE(total_level >= 1 && uint32_t(total_level) <= kMaxImplicitLevel && cell_words && nleaves >= 1 &&
                  nleaves <= (int64_t(1) << (2 * total_level)) && cell_size > 0.0f,
              "gat_morton_select: bad implicit tree");
  GAT_REQUIRE(q0 >= 0 && nq >= 0 && q0 + nq <= int64_t(UINT32_MAX), "gat_morton_select: bad query range");
  GAT_REQUIRE(lq_capacity >= 0 && lq_total && status, "gat_morton_select: bad outputs");
  if (nq == 0) return GAT_OK;
  GAT_REQUIRE(boxes && cand && (lq_capacity == 0 || leaf_query), "gat_morton_select: null pointer");
  const int blocks = grid_for(uint64_t(nq), 1, 40);
  morton_select_kernel<<<dim3{uint32_t(blocks)}, dim3{kWave}, 0, gat::as_stream(stream)>>>(
      reinterpret_cast<const uint2*>(cell_words), uint32_t(total_level), uint32_t(nleaves),
      CellGeometry{xmin, ymax, cell_size}, reinterpret_cast<const float4*>(boxes), order, uint32_t(q0), uint32_t(nq),
      node_qcount, cand, reinterpret_cast<uint2*>(leaf_query), lq_rank, uint64_t(lq_capacity),
      reinterpret_cast<unsigned long long*>(lq_total), status);
  GAT_LAUNCH_CHECK("morton_select_kernel");
  return GAT_OK;
}

int gat_grid_select(const gat_tree_node* cells, int32_t nx, int32_t ny, float xmin, float ymax, float cell_size,
                    const float* boxes, const uint32_t* order, int64_t q0, int64_t nq, uint32_t* node_qcount,
                    uint64_t* cand, uint32_t* leaf_query, uint32_t* lq_rank, int64_t lq_capacity, uint64_t* lq_total,
                    void* stream) {
  GAT_REQUIRE(cells && nx > 0 && ny > 0 && int64_t(nx) * ny <= (int64_t(1) << 30) && cell_size > 0.0f,
              "gat_grid_select: bad grid");
  GAT_REQUIRE(q0 >= 0 && nq >= 0 && q0 + nq <= int64_t(UINT32_MAX), "gat_grid_select: bad query range");
  GAT_REQUIRE(lq_capacity >= 0 && lq_total, "gat_grid_select: bad outputs");
  if (nq == 0) return GAT_OK;
  GAT_REQUIRE(boxes && cand && (lq_capacity == 0 || leaf_query), "gat_grid_select: null pointer");
  const int blocks = grid_for(uint64_t(nq), 1, 40);
  grid_select_kernel<<<dim3{uint32_t(blocks)}, dim3{kWave}, 0, gat::as_stream(stream)>>>(
      reinterpret_cast<const Node*>(cells), nx, ny, xmin, ymax, cell_size, reinterpret_cast<const float4*>(boxes),
      order, uint32_t(q0), uint32_t(nq), node_qcount, cand, reinterpret_cast<uint2*>(leaf_query), lq_rank,
      uint64_t(lq_capacity), reinterpret_cast<unsigned long long*>(lq_total));
  GAT_LAUNCH_CHECK("grid_select_kernel");
  return GAT_OK;
}

int gat_leaf_tasks(const uint32_t* leaf_query, const uint64_t* lq_total, int64_t lq_capacity,
                   const gat_tree_node* nodes, int64_t nnodes, gat_range_task* tasks, void* stream) {
  GAT_REQUIRE(lq_total && nodes && tasks && nnodes >= 1 && nnodes < int64_t(INT32_MAX) && lq_capacity >= 0,
              "gat_leaf_tasks: bad arguments");
  if (lq_capacity == 0) return GAT_OK;
  GAT_REQUIRE(leaf_query, "gat_leaf_tasks: null entries");
  leaf_tasks_kernel<<<dim3{uint32_t(grid_for(uint64_t(lq_capacity), 256, 8))}, dim3{256}, 0,
                      gat::as_stream(stream)>>>(reinterpret_cast<const uint2*>(leaf_query),
                                                reinterpret_cast<const unsigned long long*>(lq_total),
                                                uint64_t(lq_capacity), reinterpret_cast<const Node*>(nodes),
                                                uint32_t(nnodes), tasks);
  GAT_LAUNCH_CHECK("leaf_tasks_kernel");
  return GAT_OK;
}

int gat_leaf_scatter(const uint32_t* leaf_query, const uint64_t* lq_total, int64_t lq_capacity,
                     const gat_tree_node* nodes, int64_t nnodes, const uint32_t* node_off,
                     const uint32_t* node_qcount, const uint32_t* lq_rank, int32_t qb, int64_t nqueries,
                     uint32_t* qlist, int64_t qlist_capacity, uint32_t* items, int64_t items_capacity, void* stream) {
  GAT_REQUIRE(lq_total && nodes && node_off && node_qcount && qlist && items &&
                  qb > 0 && qb <= kMaxQb && nnodes >= 1 && nnodes < int64_t(INT32_MAX),
              "gat_leaf_scatter: bad arguments");
  GAT_REQUIRE(lq_capacity >= 0 && qlist_capacity >= 0 && qlist_capacity <= int64_t(UINT32_MAX) &&
                  items_capacity >= 0 && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX),
              "gat_leaf_scatter: bad capacity");
  if (lq_capacity == 0) return GAT_OK;
  GAT_REQUIRE(leaf_query && lq_rank, "gat_leaf_scatter: null entries");
  leaf_scatter_kernel<<<dim3{uint32_t(grid_for(uint64_t(lq_capacity), 256, 8))}, dim3{256}, 0,
                        gat::as_stream(stream)>>>(
      reinterpret_cast<const uint2*>(leaf_query), reinterpret_cast<const unsigned long long*>(lq_total),
      uint64_t(lq_capacity), reinterpret_cast<const Node*>(nodes), uint32_t(nnodes),
      reinterpret_cast<const uint2*>(node_off), node_qcount, lq_rank, uint32_t(qb), uint32_t(nqueries), qlist,
      uint64_t(qlist_capacity), reinterpret_cast<Item*>(items), uint64_t(items_capacity));
  GAT_LAUNCH_CHECK("leaf_scatter_kernel");
  return GAT_OK;
}

int gat_leaf_refine(const gat_point_store* sp, const uint32_t* items, const uint32_t* item_total,
                    int64_t items_capacity, const uint32_t* qlist, int64_t qlist_capacity, const float* boxes,
                    const uint64_t* qbase, int64_t nqueries, uint32_t* qcount, int32_t* pairs, int64_t capacity,
                    uint64_t* pair_total, void* stream) {
  GAT_REQUIRE(sp && sp->npoints >= 0 && sp->pid_bits >= 0 && sp->pid_bits < 32, "gat_leaf_refine: bad store");
  GAT_REQUIRE(sp->npoints == 0 || (sp->xy_dev && sp->id_dev && (sp->pid_bits > 0 || sp->pid_dev)),
              "gat_leaf_refine: null store array");
  GAT_REQUIRE(item_total && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX) && items_capacity >= 0 &&
                  qlist_capacity >= 0 && capacity >= 0,
              "gat_leaf_refine: bad sizes");
  hipStream_t s = gat::as_stream(stream);
  if (items_capacity > 0 && nqueries > 0) {
    GAT_REQUIRE(items && qlist && boxes && qbase && qcount && pairs, "gat_leaf_refine: null pointer");
    const Store st{reinterpret_cast<const float2*>(sp->xy_dev), sp->id_dev, sp->pid_dev, uint64_t(sp->npoints),
                   uint32_t(sp->pid_bits)};
    const int blocks = grid_for(uint64_t(items_capacity), kRefBlock / kWave, 16);
    const dim3 grid{uint32_t(blocks)}, block{kRefBlock};
    auto* I = reinterpret_cast<const Item*>(items);
    auto* B = reinterpret_cast<const float4*>(boxes);
    auto* P = reinterpret_cast<int2*>(pairs);
    if (sp->pid_bits > 0)
      leaf_refine_kernel<true><<<grid, block, 0, s>>>(st, I, item_total, uint64_t(items_capacity), qlist,
                                                      uint64_t(qlist_capacity), B, qbase, uint32_t(nqueries), qcount,
                                                      P, uint64_t(capacity));
    else
      leaf_refine_kernel<false><<<grid, block, 0, s>>>(st, I, item_total, uint64_t(items_capacity), qlist,
                                                       uint64_t(qlist_capacity), B, qbase, uint32_t(nqueries), qcount,
                                                       P, uint64_t(capacity));
    GAT_LAUNCH_CHECK("leaf_refine_kernel");
  }
  if (pair_total && nqueries > 0) {
    GAT_REQUIRE(qcount, "gat_leaf_refine: null counts");
    sum_u32_kernel<<<dim3{uint32_t(grid_for(uint64_t(nqueries), 256, 1))}, dim3{256}, 0, s>>>(
        qcount, uint64_t(nqueries), reinterpret_cast<unsigned long long*>(pair_total));
    GAT_LAUNCH_CHECK("sum_u32_kernel");
  }
  return GAT_OK;
}

int gat_leaf_item_regions(const uint32_t* items, const uint32_t* item_total, int64_t items_capacity,
                          uint64_t* item_base, uint64_t* tiles, uint64_t* bound_total, void* stream) {
  GAT_REQUIRE(item_total && bound_total && items_capacity >= 0, "gat_leaf_item_regions: bad arguments");
  if (items_capacity == 0) return GAT_OK;
  GAT_REQUIRE(items && item_base && tiles, "gat_leaf_item_regions: null pointer");
  // The exclusive scan of the items' bounds; *bound_total = their sum.
  return launch_scan(LoadItemBounds{reinterpret_cast<const Item*>(items), item_total, uint64_t(items_capacity)},
                     uint64_t(items_capacity), tiles, item_base, bound_total, gat::as_stream(stream));
}

int gat_leaf_refine_runs(const gat_point_store* sp, const uint32_t* items, const uint32_t* item_total,
                         int64_t items_capacity, const uint32_t* qlist, int64_t qlist_capacity, const float* boxes,
                         int64_t nqueries, int32_t* pairs, int64_t capacity, gat_range_run* runs,
                         int64_t runs_capacity, const uint64_t* item_base, void* stream) {
  GAT_REQUIRE(sp && sp->npoints >= 0 && sp->pid_bits >= 0 && sp->pid_bits < 32, "gat_leaf_refine_runs: bad store");
  GAT_REQUIRE(sp->npoints == 0 || (sp->xy_dev && sp->id_dev && (sp->pid_bits > 0 || sp->pid_dev)),
              "gat_leaf_refine_runs: null store array");
  GAT_REQUIRE(item_total && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX) && items_capacity >= 0 &&
                  qlist_capacity >= 0 && capacity >= 0 && runs_capacity >= 0,
              "gat_leaf_refine_runs: bad sizes");
  if (items_capacity == 0 || nqueries == 0) return GAT_OK;
  GAT_REQUIRE(items && qlist && boxes && pairs && runs && item_base, "gat_leaf_refine_runs: null pointer");
  hipStream_t s = gat::as_stream(stream);
  const Store st{reinterpret_cast<const float2*>(sp->xy_dev), sp->id_dev, sp->pid_dev, uint64_t(sp->npoints),
                 uint32_t(sp->pid_bits)};
  // Up to 256 workgroups per CU (about one wave per item on 10k rectangles
  // over 1M trajectories): items differ widely in work, and the dispatcher
  // hands a finished workgroup's slot to the next one, where a grid of 16 per
  // CU left each wave a fixed share of items and the launch waited on the
  // slowest shares (1.40 ms per refine against 1.26 ms; 64 per CU: 1.32,
  // 128: 1.28, 4096: 1.26).
  const int blocks = grid_for(uint64_t(items_capacity), kRunsBlock / kWave, 256 * 256 / kRunsBlock);
  const dim3 grid{uint32_t(blocks)}, block{kRunsBlock};
  auto* I = reinterpret_cast<const Item*>(items);
  auto* B = reinterpret_cast<const float4*>(boxes);
  auto* P = reinterpret_cast<int2*>(pairs);
  const RunOut o{reinterpret_cast<Run*>(runs), uint64_t(runs_capacity), item_base};
  if (sp->pid_bits > 0)
    leaf_refine_runs_kernel<true><<<grid, block, 0, s>>>(st, I, item_total, uint64_t(items_capacity), qlist,
                                                         uint64_t(qlist_capacity), B, uint32_t(nqueries), P,
                                                         uint64_t(capacity), o);
  else
    leaf_refine_runs_kernel<false><<<grid, block, 0, s>>>(st, I, item_total, uint64_t(items_capacity), qlist,
                                                          uint64_t(qlist_capacity), B, uint32_t(nqueries), P,
                                                          uint64_t(capacity), o);
  GAT_LAUNCH_CHECK("leaf_refine_runs_kernel");
  return GAT_OK;
}

int gat_run_counts(const gat_point_store* sp, const uint32_t* items, const uint32_t* item_total,
                   int64_t items_capacity, int64_t qlist_capacity, const gat_range_run* runs, int64_t runs_capacity,
                   int64_t nqueries, uint32_t* qcount, uint64_t* pair_total, void* stream) {
  GAT_REQUIRE(sp && sp->npoints >= 0 && item_total && items_capacity >= 0 && qlist_capacity >= 0 &&
                  runs_capacity >= 0 && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX),
              "gat_run_counts: bad arguments");
  if (nqueries == 0) return GAT_OK;
  GAT_REQUIRE(qcount, "gat_run_counts: null counts");
  hipStream_t s = gat::as_stream(stream);
  if (items_capacity > 0) {
    GAT_REQUIRE(items && runs, "gat_run_counts: null pointer");
    const uint64_t list_cap = uint64_t(qlist_capacity < runs_capacity ? qlist_capacity : runs_capacity);
    auto* I = reinterpret_cast<const Item*>(items);
    const auto* R = reinterpret_cast<const Run*>(runs);
    auto* T = reinterpret_cast<unsigned long long*>(pair_total);
    // Up to 2 workgroups per CU: the LDS counters take 40 KB on 10k rectangles.
    // With LDS counters one workgroup per CU: each zeroes and flushes a full
    // set of counters, so fewer, longer ranges cost less (cfg2: 37 us, against
    // 54 us with two per CU and one record in flight per thread, 46 us with
    // four).
    if (uint64_t(nqueries) <= kCountLdsQueries)
      run_counts_kernel<true><<<dim3{uint32_t(grid_for(uint64_t(items_capacity) * 8, 256, 1))}, dim3{256},
                                size_t(nqueries) * 4, s>>>(I, item_total, uint64_t(items_capacity), list_cap, R,
                                                           uint32_t(nqueries), qcount, T);
    else
      run_counts_kernel<false><<<dim3{uint32_t(grid_for(uint64_t(items_capacity) * 8, 256, 2))}, dim3{256}, 0, s>>>(
          I, item_total, uint64_t(items_capacity), list_cap, R, uint32_t(nqueries), qcount, T);
    GAT_LAUNCH_CHECK("run_counts_kernel");
  }
  return GAT_OK;
}

int gat_runs_digest(const int32_t* pairs, int64_t npairs, const gat_range_run* runs, int64_t nruns,
                    int64_t nqueries, uint64_t* out, void* stream) {
  GAT_REQUIRE(npairs >= 0 && nruns >= 0 && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX),
              "gat_runs_digest: bad sizes");
  if (nqueries == 0) return GAT_OK;
  GAT_REQUIRE(out && (nruns == 0 || (pairs && runs)), "gat_runs_digest: null pointer");
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(out), 0u, 2 * nqueries, stream)) return rc;
  if (nruns == 0) return GAT_OK;
  runs_digest_kernel<<<dim3{uint32_t(grid_for(uint64_t(nruns) * kWave, 256, 16))}, dim3{256}, 0,
                       gat::as_stream(stream)>>>(reinterpret_cast<const int2*>(pairs), uint64_t(npairs),
                                                 reinterpret_cast<const Run*>(runs), uint64_t(nruns),
                                                 uint32_t(nqueries), reinterpret_cast<unsigned long long*>(out));
  GAT_LAUNCH_CHECK("runs_digest_kernel");
  return GAT_OK;
}

int gat_graph_begin(void* stream) {
  GAT_HIP_CHECK(hipStreamBeginCapture(gat::as_stream(stream), hipStreamCaptureModeThreadLocal));
  return GAT_OK;
}

int gat_graph_end(void* stream, void** graph_exec) {
  GAT_REQUIRE(graph_exec, "gat_graph_end: null output");
  hipGraph_t g = nullptr;
  GAT_HIP_CHECK(hipStreamEndCapture(gat::as_stream(stream), &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  const hipError_t err_destroy = hipGraphDestroy(g);
  if (err != hipSuccess) return gat::hip_fail(err, "hipGraphInstantiate", __FILE__, __LINE__);
  if (err_destroy != hipSuccess) {
    (void)hipGraphExecDestroy(e);
    return gat::hip_fail(err_destroy, "hipGraphDestroy", __FILE__, __LINE__);
  }
  *graph_exec = e;
  return GAT_OK;
}

int gat_graph_launch(void* graph_exec, void* stream) {
  GAT_REQUIRE(graph_exec, "gat_graph_launch: null graph");
  GAT_HIP_CHECK(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), gat::as_stream(stream)));
  return GAT_OK;
}

int gat_graph_destroy(void* graph_exec) {
  if (graph_exec) GAT_HIP_CHECK(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return GAT_OK;
}

}  // extern "C"
