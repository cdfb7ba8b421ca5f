// Device construction of a cell-ordered point store: the GT-quadtree's Morton
// cells (Grid::addDatasetToGrid, Grid.cpp:202-261) and the flat grid's
// row-major cells (FSG::addDatasetToGrid, FSG.cpp:88-143).
//
//   cell_kernel        the cell of every point by the reference's float
//                      formula: row = (int)((ymax - lat) / cs), column =
//                      (int)((lon - xmin) / cs) (Grid.cpp:205-207), then the
//                      Morton interleave (Grid.cpp:25-54) or row * columns +
//                      column (FSG.cpp:90-92);
//   tid_kernel         trajectory id of every point (one wave per trajectory);
//   radix sort         stable, by cell, of the point indices: points
//                      of one cell keep input order, i.e. trajectory id, then
//                      position in the trajectory — the reference's order of
//                      sub-trajectories within a cell (Grid.cpp:242-261);
//   gather_kernel      the store arrays in that order: (lon, lat) pairs and
//                      the packed (trajectory id, point id);
//   cell_start_kernel  the first store index of every cell (cellnum + 1
//                      entries, the last = npoints).
// The trajectory store is read once per pass and every pass is a streaming
// kernel, so the build is bound by HBM bandwidth and the sort's passes over
// 8 bytes of key and value per point.
//
// Device construction of STIG's kd-tree store (STIG::initial / createIndex,
// STIG.cpp:35-168, with the corrected split): the tree's shape depends only
// on node sizes, so the host lays out the skeleton and the device orders the
// points, one depth at a time:
//   kd_key_kernel     the k-th node split at this depth covers positions
//                     [lo_k, hi_k]; a position inside it gets the key
//                     (2(k+1), coordinate of its point on the depth's axis as
//                     sortable bits), a position in the gap after the r-th
//                     node (2r + 1, 0) — non-decreasing along the positions
//                     outside the split nodes;
//   radix sort        stable, 64-bit keys with the point indices: every split
//                     node's points come out ordered by the axis (so the left
//                     half is <= the median point <= the right half, what
//                     nth_element guarantees) and everything else stays put;
//   kd_median_kernel  the median of node k: the axis coordinate of the point
//                     now at (lo_k + hi_k) / 2;
// then kd_box_kernel the leaves' bounding boxes (STIG.cpp:77-90) and
// gather_kernel the store in the final order.
#include <algorithm>
#include <utility>

#include "gat_common.hpp"

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Cell of (lon, lat), or UINT32_MAX when its row or column is off the grid.
__device__ __forceinline__ uint32_t cell_of(const gat_cell_grid& g, float lon, float lat) {
  const float fr = (g.ymax - lat) / g.cell_size;
  const float fc = (lon - g.xmin) / g.cell_size;
  if (!(fr > -1.0f && fr < float(g.cells_y)) || !(fc > -1.0f && fc < float(g.cells_x))) return UINT32_MAX;
  const uint32_t row = uint32_t(int(fr)), col = uint32_t(int(fc));
  return g.morton ? (spread16(col) | (spread16(row) << 1)) : row * uint32_t(g.cells_x) + col;
}

__global__ __launch_bounds__(kBlock) void cell_kernel(gat_cell_grid g, const float* __restrict__ lon,
                                                      const float* __restrict__ lat, uint64_t n,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                      unsigned long long* __restrict__ bad) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t c = cell_of(g, lon[i], lat[i]);
    if (c == UINT32_MAX) {
      atomicMin(bad, (unsigned long long)i);
      c = 0;
    }
    keys[i] = c;
    vals[i] = uint32_t(i);
  }
}

// Trajectory t (1-based) holds points [offsets[t-1], offsets[t]); entries
// outside [0, n) are not written.
__global__ __launch_bounds__(kBlock) void tid_kernel(const int64_t* __restrict__ offsets, uint32_t ntraj, uint64_t n,
                                                     uint32_t* __restrict__ tid) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = uint64_t(gridDim.x) * (kBlock / 64);
  for (uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64; w < ntraj; w += nw) {
    const int64_t b = offsets[w] < 0 ? 0 : offsets[w];
    const int64_t e = offsets[w + 1] > int64_t(n) ? int64_t(n) : offsets[w + 1];
    for (int64_t i = b + lane; i < e; i += 64) tid[i] = uint32_t(w + 1);
  }
}

__global__ __launch_bounds__(kBlock) void gather_kernel(const float* __restrict__ lon, const float* __restrict__ lat,
                                                        const int64_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ tid,
                                                        const uint32_t* __restrict__ order, uint64_t n, uint32_t ntraj,
                                                        uint32_t pid_bits, float2* __restrict__ xy,
                                                        uint32_t* __restrict__ id, uint16_t* __restrict__ pid) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; d < n; d += stride) {
    const uint32_t i = order[d];
    // Points no trajectory covers (inconsistent offsets) keep id 0.
    const uint32_t t = tid[i];
    const uint32_t p = (t >= 1 && t <= ntraj) ? uint32_t(int64_t(i) - offsets[t - 1]) : 0u;
    xy[d] = make_float2(lon[i], lat[i]);
    if (pid_bits > 0) {
      id[d] = (t << pid_bits) | p;
    } else {
      id[d] = t;
      pid[d] = uint16_t(p);
    }
  }
}

// starts[c] = first store index whose cell is >= c, for c in [0, cellnum].
__global__ __launch_bounds__(kBlock) void cell_start_kernel(const uint32_t* __restrict__ keys, uint64_t n,
                                                            uint64_t cellnum, int64_t* __restrict__ starts) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t d = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; d < n; d += stride) {
    const uint64_t k = keys[d];
    const uint64_t prev = d == 0 ? 0 : uint64_t(keys[d - 1]) + 1;
    for (uint64_t c = prev; c <= k; ++c) starts[c] = int64_t(d);
    if (d == n - 1)
      for (uint64_t c = k + 1; c <= cellnum; ++c) starts[c] = int64_t(n);
  }
}

// Unsigned order of the result = float order of f (-0 before +0; no NaNs in
// a trajectory store).
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(kBlock) void kd_key_kernel(const float* __restrict__ lon, const float* __restrict__ lat,
                                                        const uint32_t* __restrict__ perm, uint64_t n,
                                                        const int64_t* __restrict__ lo, const int64_t* __restrict__ hi,
                                                        uint32_t nsplit, uint32_t dim, uint64_t* __restrict__ keys) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    // r = number of split nodes starting at or before i.
    uint32_t a = 0, b = nsplit;
    while (a < b) {
      const uint32_t m = (a + b) >> 1;
      if (lo[m] <= int64_t(i)) a = m + 1;
      else b = m;
    }
    uint64_t key = uint64_t(2 * a + 1) << 32;
    if (a > 0 && int64_t(i) <= hi[a - 1]) {
      const uint32_t j = perm[i];
      key = (uint64_t(2 * a) << 32) | float_key(dim ? lat[j] : lon[j]);
    }
    keys[i] = key;
  }
}

__global__ __launch_bounds__(kBlock) void kd_median_kernel(const float* __restrict__ lon,
                                                           const float* __restrict__ lat,
                                                           const uint32_t* __restrict__ perm,
                                                           const int64_t* __restrict__ lo,
                                                           const int64_t* __restrict__ hi, uint32_t nsplit,
                                                           uint32_t dim, float* __restrict__ median) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsplit) return;
  const uint32_t j = perm[(lo[k] + hi[k]) / 2];
  median[k] = dim ? lat[j] : lon[j];
}

// One wave per leaf [lo, hi]: (xmin, ymin, xmax, ymax) of its points, from
// (999, 999, -999, -999) as STIG.cpp:77-90.
__global__ __launch_bounds__(kBlock) void kd_box_kernel(const float* __restrict__ lon, const float* __restrict__ lat,
                                                        const uint32_t* __restrict__ perm,
                                                        const int64_t* __restrict__ lo, const int64_t* __restrict__ hi,
                                                        uint64_t nleaves, float4* __restrict__ box) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = uint64_t(gridDim.x) * (kBlock / 64);
  for (uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64; w < nleaves; w += nw) {
    float xl = 999.0f, yl = 999.0f, xh = -999.0f, yh = -999.0f;
    for (int64_t i = lo[w] + lane; i <= hi[w]; i += 64) {
      const uint32_t j = perm[i];
      xl = fminf(xl, lon[j]);
      xh = fmaxf(xh, lon[j]);
      yl = fminf(yl, lat[j]);
      yh = fmaxf(yh, lat[j]);
    }
    for (int off = 32; off > 0; off >>= 1) {
      xl = fminf(xl, __shfl_xor(xl, off));
      xh = fmaxf(xh, __shfl_xor(xh, off));
      yl = fminf(yl, __shfl_xor(yl, off));
      yh = fmaxf(yh, __shfl_xor(yh, off));
    }
    if (lane == 0) box[w] = make_float4(xl, yl, xh, yh);
  }
}

__global__ __launch_bounds__(kBlock) void iota_kernel(uint32_t* __restrict__ v, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = uint32_t(i);
}

struct Scratch {
  void* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
};

int blocks_for(uint64_t n, int per_cu) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  const uint64_t cap = uint64_t(gat::cu_count()) * per_cu;
  if (b > cap) b = cap;
  return int(b == 0 ? 1 : b);
}


// Stable LSD radix sort of (key, value) pairs, eight bits of the key per pass
// from bit 0 up to end_bit.  A pass is three steps over tiles of 4,096 pairs:
//   radix_count_kernel    each tile's digit counts (LDS histogram), stored
//                         digit-major: counts[d * ntiles + tile];
//   scan                  their exclusive scan in that order, so offsets[d *
//                         ntiles + tile] is where the tile's pairs of digit d
//                         start (every smaller digit first, then the earlier
//                         tiles' pairs of digit d);
//   radix_scatter_kernel  the tile again, in 16 rounds of 256 pairs: a pair's
//                         rank among the round's pairs of its digit comes from
//                         eight ballots (lanes of the same digit) and the
//                         earlier waves' and rounds' counts, which orders the
//                         pairs of one digit by input position — the sort is
//                         stable; the tile is staged in LDS grouped by digit
//                         and written out in that order, so each digit's pairs
//                         go out as one contiguous run.
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kRadix = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 8;
constexpr int kScanSpan = kScanThreads * kScanPer;
constexpr int kScanTopPer = 32;  // the block sums of up to 2^28 counts

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, int bits) {
  return uint32_t(key >> shift) & ((1u << bits) - 1u);
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void radix_count_kernel(const K* __restrict__ keys, uint64_t n, int shift,
                                                                   int bits, uint32_t ntiles,
                                                                   uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[kRadix];
  const uint32_t t = threadIdx.x, tile = blockIdx.x;
  h[t] = 0;
  __syncthreads();
  const uint64_t base = uint64_t(tile) * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortRounds; ++r) {
    const uint64_t i = base + uint64_t(r) * kSortThreads + t;
    if (i < n) atomicAdd(&h[digit_of(keys[i], shift, bits)], 1u);
  }
  __syncthreads();
  counts[uint64_t(t) * ntiles + tile] = h[t];
}

// Exclusive scan of the workgroup's values (PER per thread, in thread order)
// in place in a; returns the total.  red: one word per wave plus one.
template <int THREADS, int PER>
__device__ uint32_t block_scan(uint32_t (&v)[PER], uint32_t* red) {
  const uint32_t t = threadIdx.x, lane = t & 63;
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) sum += v[k];
  uint32_t incl = sum;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = uint32_t(__shfl_up(int(incl), d, 64));
    if (lane >= uint32_t(d)) incl += y;
  }
  if (lane == 63) red[t / 64] = incl;
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (int w = 0; w < THREADS / 64; ++w) {
      const uint32_t x = red[w];
      red[w] = run;
      run += x;
    }
    red[THREADS / 64] = run;
  }
  __syncthreads();
  uint32_t at = red[t / 64] + incl - sum;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t x = v[k];
    v[k] = at;
    at += x;
  }
  const uint32_t total = red[THREADS / 64];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanThreads) void radix_scan_sums_kernel(const uint32_t* __restrict__ in, uint64_t m,
                                                                       uint32_t* __restrict__ sums) {
  __shared__ uint32_t red[kScanThreads / 64 + 1];
  const uint64_t first = uint64_t(blockIdx.x) * kScanSpan + uint64_t(threadIdx.x) * kScanPer;
  uint32_t v[kScanPer];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) v[k] = first + k < m ? in[first + k] : 0u;
  const uint32_t total = block_scan<kScanThreads, kScanPer>(v, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void radix_scan_top_kernel(uint32_t* __restrict__ sums, uint32_t nblocks) {
  __shared__ uint32_t red[kScanThreads / 64 + 1];
  const uint32_t first = threadIdx.x * kScanTopPer;
  uint32_t v[kScanTopPer];
#pragma unroll
  for (int k = 0; k < kScanTopPer; ++k) v[k] = first + k < nblocks ? sums[first + k] : 0u;
  block_scan<kScanThreads, kScanTopPer>(v, red);
#pragma unroll
  for (int k = 0; k < kScanTopPer; ++k)
    if (first + k < nblocks) sums[first + k] = v[k];
}

__global__ __launch_bounds__(kScanThreads) void radix_scan_apply_kernel(const uint32_t* __restrict__ in, uint64_t m,
                                                                        const uint32_t* __restrict__ sums,
                                                                        uint32_t* __restrict__ out) {
  __shared__ uint32_t red[kScanThreads / 64 + 1];
  const uint64_t first = uint64_t(blockIdx.x) * kScanSpan + uint64_t(threadIdx.x) * kScanPer;
  uint32_t v[kScanPer];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) v[k] = first + k < m ? in[first + k] : 0u;
  block_scan<kScanThreads, kScanPer>(v, red);
  const uint32_t carry = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (first + k < m) out[first + k] = carry + v[k];
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint64_t n, int shift, int bits, uint32_t ntiles,
    const uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets) {
  __shared__ K skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint32_t start[kRadix], gofs[kRadix], run[kRadix];
  __shared__ uint32_t wcnt[kSortWaves][kRadix], wpre[kSortWaves][kRadix];
  __shared__ uint32_t red[kSortThreads / 64 + 1];
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6, tile = blockIdx.x;
  const uint64_t base = uint64_t(tile) * kSortTile;
  const uint32_t tile_n = uint32_t(n - base < uint64_t(kSortTile) ? n - base : uint64_t(kSortTile));
  uint32_t c[1] = {counts[uint64_t(t) * ntiles + tile]};
  gofs[t] = offsets[uint64_t(t) * ntiles + tile];
  block_scan<kSortThreads, 1>(c, red);
  start[t] = c[0];
  run[t] = 0;
  for (int k = 0; k < kSortWaves; ++k) wcnt[k][t] = 0;
  __syncthreads();
  const uint64_t lt = (uint64_t(1) << lane) - 1u;
  // The whole tile is loaded before the first round: a round's loads would
  // otherwise wait behind the previous round's barriers, one memory round
  // trip per round.
  K kr[kSortRounds];
  uint32_t vr[kSortRounds];
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t j = uint32_t(r) * kSortThreads + t;
    kr[r] = j < tile_n ? keys_in[base + j] : K(0);
    vr[r] = j < tile_n ? vals_in[base + j] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const uint32_t j = uint32_t(r) * kSortThreads + t;
    const bool valid = j < tile_n;
    const K key = kr[r];
    const uint32_t val = vr[r];
    const uint32_t d = digit_of(key, shift, bits);
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool set = (d >> b) & 1u;
      const uint64_t bb = __ballot(set);
      m &= set ? bb : ~bb;
    }
    const uint32_t below = uint32_t(__popcll(m & lt));
    if (valid && below == 0) wcnt[w][d] = uint32_t(__popcll(m));
    __syncthreads();
    // Thread t = digit t: where each wave's pairs of digit t start among this group of 256 pairs.
    uint32_t acc = run[t];
#pragma unroll
    for (int k = 0; k < kSortWaves; ++k) {
      const uint32_t x = wcnt[k][t];
      wpre[k][t] = acc;
      wcnt[k][t] = 0;
      acc += x;
    }
    run[t] = acc;
    __syncthreads();
    if (valid) {
      const uint32_t pos = start[d] + wpre[w][d] + below;
      skey[pos] = key;
      sval[pos] = val;
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < tile_n; j += kSortThreads) {
    const K key = skey[j];
    const uint32_t d = digit_of(key, shift, bits);
    const uint64_t at = uint64_t(gofs[d]) + (j - start[d]);
    keys_out[at] = key;
    vals_out[at] = sval[j];
  }
}

// Device bytes radix_sort_pairs needs beside the two key and value buffers.
size_t radix_scratch_bytes(uint64_t n) {
  const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t m = ntiles * kRadix;
  const uint64_t nblocks = (m + kScanSpan - 1) / kScanSpan;
  return size_t((2 * m + nblocks) * 4 + 256);
}

// Sorts n pairs by bits [0, end_bit) of the key, stably, ping-ponging between
// (*keys, *vals) and (*keys_alt, *vals_alt); on return *keys / *vals point to
// the sorted pairs (the other pair of buffers holds whatever the last pass
// left).  Asynchronous on s.
template <typename K>
int radix_sort_pairs(K** keys, K** keys_alt, uint32_t** vals, uint32_t** vals_alt, uint64_t n, int end_bit,
                     void* scratch, hipStream_t s) {
  GAT_REQUIRE(n < (uint64_t(1) << 32), "radix sort: 2^32 or more pairs");
  GAT_REQUIRE(end_bit >= 0 && end_bit <= int(8 * sizeof(K)), "radix sort: bad bit count");
  if (n == 0 || end_bit == 0) return GAT_OK;
  const uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const uint64_t m = ntiles * kRadix;
  const uint64_t nblocks = (m + kScanSpan - 1) / kScanSpan;
  GAT_REQUIRE(nblocks <= uint64_t(kScanThreads) * kScanTopPer, "radix sort: too many pairs for the scan");
  auto* counts = static_cast<uint32_t*>(scratch);
  uint32_t* offsets = counts + m;
  uint32_t* sums = offsets + m;
  for (int shift = 0; shift < end_bit; shift += 8) {
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    radix_count_kernel<K><<<dim3(uint32_t(ntiles)), dim3(kSortThreads), 0, s>>>(*keys, n, shift, bits,
                                                                                uint32_t(ntiles), counts);
    GAT_LAUNCH_CHECK("radix_count_kernel");
    radix_scan_sums_kernel<<<dim3(uint32_t(nblocks)), dim3(kScanThreads), 0, s>>>(counts, m, sums);
    GAT_LAUNCH_CHECK("radix_scan_sums_kernel");
    radix_scan_top_kernel<<<dim3(1), dim3(kScanThreads), 0, s>>>(sums, uint32_t(nblocks));
    GAT_LAUNCH_CHECK("radix_scan_top_kernel");
    radix_scan_apply_kernel<<<dim3(uint32_t(nblocks)), dim3(kScanThreads), 0, s>>>(counts, m, sums, offsets);
    GAT_LAUNCH_CHECK("radix_scan_apply_kernel");
    radix_scatter_kernel<K><<<dim3(uint32_t(ntiles)), dim3(kSortThreads), 0, s>>>(
        *keys, *vals, *keys_alt, *vals_alt, n, shift, bits, uint32_t(ntiles), counts, offsets);
    GAT_LAUNCH_CHECK("radix_scatter_kernel");
    std::swap(*keys, *keys_alt);
    std::swap(*vals, *vals_alt);
  }
  return GAT_OK;
}

}  // namespace

extern "C" int gat_cell_store_build(const gat_cell_grid* grid, const float* lon, const float* lat,
                                    const int64_t* offsets, int32_t ntraj, int64_t npoints, int32_t pid_bits,
                                    float* xy_out, uint32_t* id_out, uint16_t* pid_out, int64_t* cell_start_out,
                                    int64_t* bad_point, void* stream) {
  GAT_REQUIRE(grid && bad_point, "cell store build: null argument");
  const gat_cell_grid g = *grid;
  GAT_REQUIRE(g.cells_x > 0 && g.cells_y > 0 && g.cell_size > 0.0f, "cell store build: bad grid");
  const uint64_t cellnum = uint64_t(g.cells_x) * uint64_t(g.cells_y);
  GAT_REQUIRE(!g.morton || (g.cells_x == g.cells_y && g.cells_x <= 32768 && (g.cells_x & (g.cells_x - 1)) == 0),
              "cell store build: Morton grids are square with a power-of-two side up to 32768");
  GAT_REQUIRE(cellnum <= (uint64_t(1) << 31), "cell store build: more than 2^31 cells");
  GAT_REQUIRE(ntraj >= 0 && npoints >= 0 && npoints < int64_t(UINT32_MAX), "cell store build: bad sizes");
  GAT_REQUIRE(pid_bits >= 0 && pid_bits < 32, "cell store build: pid_bits outside 0..31");
  GAT_REQUIRE(cell_start_out && offsets, "cell store build: null output");
  GAT_REQUIRE(npoints == 0 || (lon && lat && xy_out && id_out && (pid_bits > 0 || pid_out)),
              "cell store build: null point array");
  *bad_point = -1;
  hipStream_t s = gat::as_stream(stream);
  const uint64_t n = uint64_t(npoints);
  if (n == 0) {
    GAT_HIP_CHECK(hipMemsetAsync(cell_start_out, 0, (cellnum + 1) * sizeof(int64_t), s));
    GAT_HIP_CHECK(hipStreamSynchronize(s));
    return GAT_OK;
  }
  // One allocation: keys and values twice (sort input and output), the
  // trajectory id of every point, the off-grid flag, then the sort's storage.
  int end_bit = 1;
  while (end_bit < 32 && (uint64_t(1) << end_bit) < cellnum) ++end_bit;
  const size_t sort_bytes = radix_scratch_bytes(n);
  const size_t arr = (n * 4 + 255) & ~size_t(255);
  const size_t total = 5 * arr + 256 + sort_bytes;
  Scratch scratch;
  GAT_HIP_CHECK(hipMalloc(&scratch.p, total));
  char* base = static_cast<char*>(scratch.p);
  auto* keys = reinterpret_cast<uint32_t*>(base);
  auto* keys_alt = reinterpret_cast<uint32_t*>(base + arr);
  auto* vals = reinterpret_cast<uint32_t*>(base + 2 * arr);
  auto* vals_alt = reinterpret_cast<uint32_t*>(base + 3 * arr);
  auto* tid = reinterpret_cast<uint32_t*>(base + 4 * arr);
  auto* bad = reinterpret_cast<unsigned long long*>(base + 5 * arr);
  void* sort_tmp = base + 5 * arr + 256;

  GAT_HIP_CHECK(hipMemsetAsync(bad, 0xFF, sizeof(unsigned long long), s));
  cell_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(g, lon, lat, n, keys, vals, bad);
  GAT_LAUNCH_CHECK("cell_kernel");
  GAT_HIP_CHECK(hipMemsetAsync(tid, 0, n * 4, s));
  tid_kernel<<<dim3(uint32_t(blocks_for(uint64_t(ntraj) * 64, 16))), dim3(kBlock), 0, s>>>(offsets, uint32_t(ntraj),
                                                                                          n, tid);
  GAT_LAUNCH_CHECK("tid_kernel");
  unsigned long long first_bad = 0;
  GAT_HIP_CHECK(hipMemcpyAsync(&first_bad, bad, sizeof(first_bad), hipMemcpyDeviceToHost, s));
  GAT_HIP_CHECK(hipStreamSynchronize(s));
  if (first_bad != ~0ull) {
    *bad_point = int64_t(first_bad);
    return GAT_OK;
  }
  if (int rc = radix_sort_pairs<uint32_t>(&keys, &keys_alt, &vals, &vals_alt, n, end_bit, sort_tmp, s)) return rc;
  gather_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(
      lon, lat, offsets, tid, vals, n, uint32_t(ntraj), uint32_t(pid_bits), reinterpret_cast<float2*>(xy_out),
      id_out, pid_out);
  GAT_LAUNCH_CHECK("gather_kernel");
  cell_start_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(keys, n, cellnum,
                                                                               cell_start_out);
  GAT_LAUNCH_CHECK("cell_start_kernel");
  // The scratch block is freed on return, after the stream has drained.
  GAT_HIP_CHECK(hipStreamSynchronize(s));
  return GAT_OK;
}

extern "C" int gat_kd_store_build(const float* lon, const float* lat, const int64_t* offsets, int32_t ntraj,
                                  int64_t npoints, int32_t pid_bits, const int64_t* split_lo, const int64_t* split_hi,
                                  const int64_t* depth_first, int32_t depths, const int64_t* leaf_lo,
                                  const int64_t* leaf_hi, int64_t nleaves, float* xy_out, uint32_t* id_out,
                                  uint16_t* pid_out, float* median_out, float* leaf_box_out, void* stream) {
  GAT_REQUIRE(ntraj >= 0 && npoints >= 1 && npoints < int64_t(UINT32_MAX) && depths >= 0 && nleaves >= 1,
              "kd store build: bad sizes");
  GAT_REQUIRE(pid_bits >= 0 && pid_bits < 32, "kd store build: pid_bits outside 0..31");
  GAT_REQUIRE(lon && lat && offsets && xy_out && id_out && (pid_bits > 0 || pid_out) && leaf_lo && leaf_hi &&
                  leaf_box_out && depth_first && (depths == 0 || (split_lo && split_hi && median_out)),
              "kd store build: null argument");
  const uint64_t n = uint64_t(npoints);
  const int64_t nsplit = depth_first[depths];
  GAT_REQUIRE(depth_first[0] == 0 && nsplit >= 0 && nsplit < int64_t(UINT32_MAX / 2), "kd store build: bad depths");
  // The host's skeleton: split nodes per depth disjoint and ascending, leaves
  // inside [0, n).  A bad layout is rejected here, before any kernel reads it.
  for (int32_t d = 0; d < depths; ++d) {
    GAT_REQUIRE(depth_first[d] <= depth_first[d + 1], "kd store build: depths out of order");
    for (int64_t k = depth_first[d]; k < depth_first[d + 1]; ++k) {
      GAT_REQUIRE(split_lo[k] >= 0 && split_lo[k] < split_hi[k] && split_hi[k] < npoints,
                  "kd store build: split node outside the points");
      GAT_REQUIRE(k == depth_first[d] || split_lo[k] > split_hi[k - 1], "kd store build: split nodes overlap");
    }
  }
  for (int64_t k = 0; k < nleaves; ++k)
    GAT_REQUIRE(leaf_lo[k] >= 0 && leaf_lo[k] <= leaf_hi[k] && leaf_hi[k] < npoints,
                "kd store build: leaf outside the points");
  hipStream_t s = gat::as_stream(stream);
  const size_t sort_bytes = radix_scratch_bytes(n);
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t k8 = al(n * 8), k4 = al(n * 4);
  const size_t sp = al(size_t(std::max<int64_t>(nsplit, 1)) * 8), lf = al(size_t(nleaves) * 8);
  const size_t total = 2 * k8 + 3 * k4 + 2 * sp + al(size_t(std::max<int64_t>(nsplit, 1)) * 4) + 2 * lf +
                       al(size_t(nleaves) * 16) + sort_bytes;
  Scratch scratch;
  GAT_HIP_CHECK(hipMalloc(&scratch.p, total));
  char* at = static_cast<char*>(scratch.p);
  auto take = [&](size_t b) {
    char* p = at;
    at += b;
    return p;
  };
  auto* keys_a = reinterpret_cast<uint64_t*>(take(k8));
  auto* keys_b = reinterpret_cast<uint64_t*>(take(k8));
  auto* perm_a = reinterpret_cast<uint32_t*>(take(k4));
  auto* perm_b = reinterpret_cast<uint32_t*>(take(k4));
  auto* tid = reinterpret_cast<uint32_t*>(take(k4));
  auto* dlo = reinterpret_cast<int64_t*>(take(sp));
  auto* dhi = reinterpret_cast<int64_t*>(take(sp));
  auto* dmed = reinterpret_cast<float*>(take(al(size_t(std::max<int64_t>(nsplit, 1)) * 4)));
  auto* llo = reinterpret_cast<int64_t*>(take(lf));
  auto* lhi = reinterpret_cast<int64_t*>(take(lf));
  auto* lbox = reinterpret_cast<float4*>(take(al(size_t(nleaves) * 16)));
  void* sort_tmp = take(sort_bytes);
  if (nsplit > 0) {
    GAT_HIP_CHECK(hipMemcpyAsync(dlo, split_lo, size_t(nsplit) * 8, hipMemcpyHostToDevice, s));
    GAT_HIP_CHECK(hipMemcpyAsync(dhi, split_hi, size_t(nsplit) * 8, hipMemcpyHostToDevice, s));
  }
  GAT_HIP_CHECK(hipMemcpyAsync(llo, leaf_lo, size_t(nleaves) * 8, hipMemcpyHostToDevice, s));
  GAT_HIP_CHECK(hipMemcpyAsync(lhi, leaf_hi, size_t(nleaves) * 8, hipMemcpyHostToDevice, s));
  const dim3 grid(uint32_t(blocks_for(n, 16))), block(kBlock);
  iota_kernel<<<grid, block, 0, s>>>(perm_a, n);
  GAT_LAUNCH_CHECK("iota_kernel");
  for (int32_t d = 0; d < depths; ++d) {
    const int64_t first = depth_first[d], cnt = depth_first[d + 1] - first;
    if (cnt == 0) continue;
    const uint32_t dim = uint32_t(d & 1);  // even depths split on x, odd on y (STIG.cpp:94-140)
    kd_key_kernel<<<grid, block, 0, s>>>(lon, lat, perm_a, n, dlo + first, dhi + first, uint32_t(cnt), dim, keys_a);
    GAT_LAUNCH_CHECK("kd_key_kernel");
    int bits = 1;
    while (bits < 31 && (uint64_t(1) << bits) <= uint64_t(2 * cnt + 1)) ++bits;
    if (int rc = radix_sort_pairs<uint64_t>(&keys_a, &keys_b, &perm_a, &perm_b, n, 32 + bits, sort_tmp, s))
      return rc;
    kd_median_kernel<<<dim3(uint32_t((cnt + kBlock - 1) / kBlock)), block, 0, s>>>(
        lon, lat, perm_a, dlo + first, dhi + first, uint32_t(cnt), dim, dmed + first);
    GAT_LAUNCH_CHECK("kd_median_kernel");
  }
  kd_box_kernel<<<dim3(uint32_t(blocks_for(uint64_t(nleaves) * 64, 16))), block, 0, s>>>(lon, lat, perm_a, llo, lhi,
                                                                                         uint64_t(nleaves), lbox);
  GAT_LAUNCH_CHECK("kd_box_kernel");
  GAT_HIP_CHECK(hipMemsetAsync(tid, 0, n * 4, s));
  tid_kernel<<<dim3(uint32_t(blocks_for(uint64_t(ntraj) * 64, 16))), block, 0, s>>>(offsets, uint32_t(ntraj), n,
                                                                                    tid);
  GAT_LAUNCH_CHECK("tid_kernel");
  gather_kernel<<<grid, block, 0, s>>>(lon, lat, offsets, tid, perm_a, n, uint32_t(ntraj), uint32_t(pid_bits),
                                       reinterpret_cast<float2*>(xy_out), id_out, pid_out);
  GAT_LAUNCH_CHECK("gather_kernel");
  if (nsplit > 0) GAT_HIP_CHECK(hipMemcpyAsync(median_out, dmed, size_t(nsplit) * 4, hipMemcpyDeviceToHost, s));
  GAT_HIP_CHECK(hipMemcpyAsync(leaf_box_out, lbox, size_t(nleaves) * 16, hipMemcpyDeviceToHost, s));
  GAT_HIP_CHECK(hipStreamSynchronize(s));
  return GAT_OK;
}
