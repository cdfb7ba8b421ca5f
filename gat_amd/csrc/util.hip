// Calibration kernels.  A plain streaming copy gives the stream-copy HBM
// bandwidth the range scan is compared with, and the known-size reference that
// rocprofv3's FETCH_SIZE / WRITE_SIZE readings of other kernels are scaled by.
// Write-only / read-only streams and packed runs of 8-byte pairs give the
// ceilings of the access patterns the result writes of the range kernels have.
#include "gat_common.hpp"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// Each thread keeps four 16-byte loads in flight before storing them;
// nontemporal accesses keep the streamed data out of the caches.
__global__ __launch_bounds__(256) void stream_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                                          uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const v4f a = __builtin_nontemporal_load(src + i);
    const v4f b = __builtin_nontemporal_load(src + i + stride);
    const v4f c = __builtin_nontemporal_load(src + i + 2 * stride);
    const v4f d = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + stride);
    __builtin_nontemporal_store(c, dst + i + 2 * stride);
    __builtin_nontemporal_store(d, dst + i + 3 * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

typedef int v2i __attribute__((ext_vector_type(2)));

// Write-only and read-only streams, 16 bytes per lane.
__global__ __launch_bounds__(256) void stream_write_kernel(v4f* __restrict__ dst, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  const v4f z = {1.0f, 2.0f, 3.0f, 4.0f};
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(z, dst + i);
}

__global__ __launch_bounds__(256) void stream_read_kernel(const v4f* __restrict__ src, uint64_t n,
                                                          float* __restrict__ sink) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  float acc = 0.0f;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const v4f v = __builtin_nontemporal_load(src + i);
    acc += v.x + v.y + v.z + v.w;
  }
  if (acc == -1.0f) sink[0] = acc;  // keeps the loads; never true for the filled buffer
}

// Runs of `run` 8-byte pairs packed back to back (run k starts at pair k * run,
// any alignment), one run per wave step in scattered order (k times an odd
// constant modulo the power-of-two run count): the layout a batch of result
// runs from many waves leaves in per-query regions.  aligned = false: 64-lane
// stores starting at the run's first pair; true: 64-lane windows aligned to
// 16 pairs (128-byte lines), so only the run's two end lines are partial.
// Streaming (nontemporal) or cached stores.
template <bool kStreaming>
__global__ __launch_bounds__(256) void packed_runs_kernel(v2i* __restrict__ dst, uint64_t nruns, uint32_t run,
                                                          bool aligned) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
  const uint64_t nw = uint64_t(gridDim.x) * blockDim.x / 64;
  for (uint64_t k = w; k < nruns; k += nw) {
    const uint64_t r = (k * 0x9E3779B97F4A7C15ull) & (nruns - 1);
    const uint64_t s = r * run, e = s + run;
    const uint64_t a = aligned ? (s & ~uint64_t(15)) : s;
    for (uint64_t b = a; b < e; b += 64) {
      const uint64_t i = b + lane;
      if (i >= s && i < e) {
        const v2i p = {int(r), int(lane)};
        if constexpr (kStreaming)
          __builtin_nontemporal_store(p, dst + i);
        else
          dst[i] = p;
      }
    }
  }
}

}  // namespace

extern "C" int gat_hbm_pattern(int kind, void* buf, size_t bytes, uint32_t run, void* stream) {
  GAT_REQUIRE(kind >= 0 && kind <= 5, "HBM pattern: kind outside 0..5");
  GAT_REQUIRE(bytes % 16 == 0 && bytes >= 16, "HBM pattern: size must be a positive multiple of 16 bytes");
  GAT_REQUIRE(buf, "HBM pattern: null buffer");
  hipStream_t s = gat::as_stream(stream);
  // Writes run fastest from a wide grid (64 workgroups per CU: 5.5 TB/s
  // against 4.1-5.0 from 16), reads from a narrower one (16 per CU: 6.3-7.1
  // TB/s against 5.7 from 64).
  const uint64_t cap = uint64_t(gat::cu_count()) * 64;
  if (kind <= 1) {
    const uint64_t n = bytes / 16;
    uint64_t blocks = (n + 255) / 256;
    const uint64_t stream_cap = kind == 0 ? cap : uint64_t(gat::cu_count()) * 16;
    if (blocks > stream_cap) blocks = stream_cap;
    if (kind == 0)
      stream_write_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, s>>>(static_cast<v4f*>(buf), n);
    else
      stream_read_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, s>>>(static_cast<const v4f*>(buf), n,
                                                                       static_cast<float*>(buf));
  } else {
    GAT_REQUIRE(run >= 1 && run <= 4096 && uint64_t(run) * 8 <= bytes, "HBM pattern: run outside 1..4096 pairs");
    uint64_t nruns = 1;
    while (nruns * 2 * run * 8 <= bytes) nruns *= 2;
    if (kind <= 3)
      packed_runs_kernel<true><<<dim3(uint32_t(cap)), dim3(256), 0, s>>>(static_cast<v2i*>(buf), nruns, run,
                                                                         kind == 3);
    else
      packed_runs_kernel<false><<<dim3(uint32_t(cap)), dim3(256), 0, s>>>(static_cast<v2i*>(buf), nruns, run,
                                                                          kind == 5);
  }
  GAT_LAUNCH_CHECK("HBM pattern kernel");
  return GAT_OK;
}

extern "C" int gat_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
  GAT_REQUIRE(bytes % 16 == 0, "stream copy: size must be a multiple of 16 bytes");
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "stream copy: null pointer");
  const uint64_t n = bytes / 16;
  uint64_t blocks = (n + 255) / 256;
  // 64 workgroups per CU: 5.22 TB/s, against 4.72-4.81 from 16 per CU.
  const uint64_t cap = uint64_t(gat::cu_count()) * 64;
  if (blocks > cap) blocks = cap;
  stream_copy_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, gat::as_stream(stream)>>>(
      static_cast<const v4f*>(src), static_cast<v4f*>(dst), n);
  GAT_LAUNCH_CHECK("stream_copy_kernel");
  return GAT_OK;
}
