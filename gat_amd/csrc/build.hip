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
//   radix sort         stable, by cell, of the point indices (rocPRIM): points
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
#include <rocprim/device/device_radix_sort.hpp>

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
  size_t sort_bytes = 0;
  GAT_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<const uint32_t*>(nullptr),
                                          static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                          static_cast<uint32_t*>(nullptr), size_t(n), 0, end_bit, s));
  const size_t arr = (n * 4 + 255) & ~size_t(255);
  const size_t total = 5 * arr + 256 + sort_bytes;
  Scratch scratch;
  GAT_HIP_CHECK(hipMalloc(&scratch.p, total));
  char* base = static_cast<char*>(scratch.p);
  auto* keys_in = reinterpret_cast<uint32_t*>(base);
  auto* keys_out = reinterpret_cast<uint32_t*>(base + arr);
  auto* vals_in = reinterpret_cast<uint32_t*>(base + 2 * arr);
  auto* vals_out = reinterpret_cast<uint32_t*>(base + 3 * arr);
  auto* tid = reinterpret_cast<uint32_t*>(base + 4 * arr);
  auto* bad = reinterpret_cast<unsigned long long*>(base + 5 * arr);
  void* sort_tmp = base + 5 * arr + 256;

  GAT_HIP_CHECK(hipMemsetAsync(bad, 0xFF, sizeof(unsigned long long), s));
  cell_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(g, lon, lat, n, keys_in, vals_in, bad);
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
  GAT_HIP_CHECK(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys_in, keys_out, vals_in, vals_out, size_t(n), 0,
                                          end_bit, s));
  gather_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(
      lon, lat, offsets, tid, vals_out, n, uint32_t(ntraj), uint32_t(pid_bits), reinterpret_cast<float2*>(xy_out),
      id_out, pid_out);
  GAT_LAUNCH_CHECK("gather_kernel");
  cell_start_kernel<<<dim3(uint32_t(blocks_for(n, 16))), dim3(kBlock), 0, s>>>(keys_out, n, cellnum,
                                                                               cell_start_out);
  GAT_LAUNCH_CHECK("cell_start_kernel");
  // The scratch block is freed on return, after the stream has drained.
  GAT_HIP_CHECK(hipStreamSynchronize(s));
  return GAT_OK;
}
