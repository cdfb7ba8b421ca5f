// Calibration kernel: a plain streaming copy.  It gives the stream-copy HBM
// bandwidth the range scan is compared with, and the known-size reference that
// rocprofv3's FETCH_SIZE / WRITE_SIZE readings of other kernels are scaled by.
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

}  // namespace

extern "C" int gat_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
  GAT_REQUIRE(bytes % 16 == 0, "stream copy: size must be a multiple of 16 bytes");
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "stream copy: null pointer");
  const uint64_t n = bytes / 16;
  uint64_t blocks = (n + 255) / 256;
  const uint64_t cap = uint64_t(gat::cu_count()) * 16;
  if (blocks > cap) blocks = cap;
  stream_copy_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, gat::as_stream(stream)>>>(
      static_cast<const v4f*>(src), static_cast<v4f*>(dst), n);
  GAT_LAUNCH_CHECK("stream_copy_kernel");
  return GAT_OK;
}
