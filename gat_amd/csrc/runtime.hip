// Device, memory, stream and event management for libgat_hip.so: the
// replacement for the CUDA runtime calls the reference's host classes make
// directly (Grid.cpp:470-476, STIG.cpp:271-273, FSG.cpp:165-170) and for its
// abort-on-error CUDA_CALL macro (cudaKernel.h:10).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>

#include "gat_common.hpp"

namespace gat {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

}  // namespace gat

extern "C" {

const char* gat_last_error(void) { return gat::g_last_error.c_str(); }

int gat_init(int device) {
  int n = 0;
  GAT_HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    gat::set_error("gat_init: device %d out of range (%d visible)", device, n);
    return GAT_ERR_ARG;
  }
  GAT_HIP_CHECK(hipSetDevice(device));
  GAT_HIP_CHECK(hipFree(nullptr));  // force context creation
  return gat::scratch_setup();
}

int gat_device_name(char* buf, int len) {
  GAT_REQUIRE(buf && len > 0, "gat_device_name: bad buffer");
  int dev = 0;
  GAT_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  GAT_HIP_CHECK(hipGetDeviceProperties(&p, dev));
  snprintf(buf, size_t(len), "%s (%s)", p.name, p.gcnArchName);
  return GAT_OK;
}

int gat_device_props(int64_t* out) {
  GAT_REQUIRE(out, "gat_device_props: null output");
  int dev = 0;
  GAT_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  GAT_HIP_CHECK(hipGetDeviceProperties(&p, dev));
  out[0] = p.multiProcessorCount;
  out[1] = int64_t(p.sharedMemPerBlock);
  out[2] = int64_t(p.totalGlobalMem);
  out[3] = p.clockRate;
  out[4] = p.warpSize;
  return GAT_OK;
}

int gat_mem_info(size_t* free_bytes, size_t* total_bytes) {
  GAT_REQUIRE(free_bytes && total_bytes, "gat_mem_info: null output");
  GAT_HIP_CHECK(hipMemGetInfo(free_bytes, total_bytes));
  return GAT_OK;
}

int gat_malloc(void** ptr, size_t bytes) {
  GAT_REQUIRE(ptr, "gat_malloc: null output");
  *ptr = nullptr;
  if (bytes == 0) return GAT_OK;
  GAT_HIP_CHECK(hipMalloc(ptr, bytes));
  return GAT_OK;
}

int gat_free(void* ptr) {
  if (ptr) GAT_HIP_CHECK(hipFree(ptr));
  return GAT_OK;
}

int gat_host_alloc(void** ptr, size_t bytes) {
  GAT_REQUIRE(ptr, "gat_host_alloc: null output");
  *ptr = nullptr;
  if (bytes == 0) return GAT_OK;
  GAT_HIP_CHECK(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return GAT_OK;
}

int gat_host_free(void* ptr) {
  if (ptr) GAT_HIP_CHECK(hipHostFree(ptr));
  return GAT_OK;
}

int gat_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "gat_memcpy_h2d: null pointer");
  GAT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, gat::as_stream(stream)));
  return GAT_OK;
}

int gat_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "gat_memcpy_d2h: null pointer");
  GAT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, gat::as_stream(stream)));
  return GAT_OK;
}

int gat_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst && src, "gat_memcpy_d2d: null pointer");
  GAT_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, gat::as_stream(stream)));
  return GAT_OK;
}

int gat_memset(void* dst, int value, size_t bytes, void* stream) {
  if (bytes == 0) return GAT_OK;
  GAT_REQUIRE(dst, "gat_memset: null pointer");
  GAT_HIP_CHECK(hipMemsetAsync(dst, value, bytes, gat::as_stream(stream)));
  return GAT_OK;
}

int gat_stream_create(void** stream) {
  GAT_REQUIRE(stream, "gat_stream_create: null output");
  hipStream_t s;
  GAT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return GAT_OK;
}

int gat_stream_destroy(void* stream) {
  if (stream) GAT_HIP_CHECK(hipStreamDestroy(gat::as_stream(stream)));
  return GAT_OK;
}

int gat_stream_sync(void* stream) {
  GAT_HIP_CHECK(hipStreamSynchronize(gat::as_stream(stream)));
  return GAT_OK;
}

int gat_device_sync(void) {
  GAT_HIP_CHECK(hipDeviceSynchronize());
  return GAT_OK;
}

int gat_event_create(void** event) {
  GAT_REQUIRE(event, "gat_event_create: null output");
  hipEvent_t e;
  GAT_HIP_CHECK(hipEventCreate(&e));
  *event = e;
  return GAT_OK;
}

int gat_event_destroy(void* event) {
  if (event) GAT_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return GAT_OK;
}

int gat_event_record(void* event, void* stream) {
  GAT_REQUIRE(event, "gat_event_record: null event");
  GAT_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event), gat::as_stream(stream)));
  return GAT_OK;
}

int gat_stream_wait_event(void* stream, void* event) {
  GAT_REQUIRE(event, "gat_stream_wait_event: null event");
  GAT_HIP_CHECK(hipStreamWaitEvent(gat::as_stream(stream), static_cast<hipEvent_t>(event), 0));
  return GAT_OK;
}

int gat_event_sync(void* event) {
  GAT_REQUIRE(event, "gat_event_sync: null event");
  GAT_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(event)));
  return GAT_OK;
}

int gat_event_elapsed_ms(float* ms, void* start, void* stop) {
  GAT_REQUIRE(ms && start && stop, "gat_event_elapsed_ms: null argument");
  GAT_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return GAT_OK;
}

// (double)|d| < eps holds exactly for the floats |d| below the smallest float
// that is >= eps, so that float is the single-precision threshold.
float gat_eps_threshold(double eps) {
  if (!(eps > 0)) return 0.0f;
  if (eps > double(FLT_MAX)) return INFINITY;
  float u = float(eps);  // round to nearest
  if (double(u) < eps) u = nextafterf(u, INFINITY);
  return u;
}

}  // extern "C"
