// Calibration kernel: a plain streaming copy.  It gives the stream-copy HBM
// bandwidth the range scan is compared with, and the known-size reference that
// rocprofv3's FETCH_SIZE / WRITE_SIZE readings of other kernels are scaled by.
#include "gat_common.hpp"

namespace {

__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int gat_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
  GAT_REQUIRE(bytes % 16 == 0, "stream copy: size must be a multiple of 16 bytes");
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "stream copy: null pointer");
  const uint64_t n = bytes / 16;
  uint64_t blocks = (n + 255) / 256;
  const uint64_t cap = uint64_t(gat::cu_count()) * 32;
  if (blocks > cap) blocks = cap;
  stream_copy_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, gat::as_stream(stream)>>>(
      static_cast<const float4*>(src), static_cast<float4*>(dst), n);
  GAT_LAUNCH_CHECK("stream_copy_kernel");
  return GAT_OK;
}
