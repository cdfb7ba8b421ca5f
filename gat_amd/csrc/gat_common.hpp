// Internal helpers shared by the translation units of libgat_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "gat_hip.h"

namespace gat {

// Records the calling thread's last error message (read by gat_last_error).
void set_error(const char* fmt, ...);

inline int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%d) at %s:%d", what, hipGetErrorString(e), int(e), file, line);
  return e == hipErrorOutOfMemory ? GAT_ERR_NOMEM : GAT_ERR_HIP;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Number of compute units of the current device (cached per device).
int cu_count();

// Resolves the current device's address of the no-alloc handler's bitmap
// scratch (range.hip); gat_init calls it once per device.
int scratch_setup();

}  // namespace gat

#define GAT_HIP_CHECK(call)                                              \
  do {                                                                   \
    hipError_t gat_e_ = (call);                                          \
    if (gat_e_ != hipSuccess) return gat::hip_fail(gat_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define GAT_REQUIRE(cond, ...)          \
  do {                                  \
    if (!(cond)) {                      \
      gat::set_error(__VA_ARGS__);      \
      return GAT_ERR_ARG;               \
    }                                   \
  } while (0)

// Launch-error check after a kernel launch.
#define GAT_LAUNCH_CHECK(name)                                                    \
  do {                                                                            \
    hipError_t gat_e_ = hipGetLastError();                                        \
    if (gat_e_ != hipSuccess) return gat::hip_fail(gat_e_, name, __FILE__, __LINE__); \
  } while (0)

namespace gat {

// A result pair written with a streaming (nontemporal) store, for kernels
// whose output runs are long and written once (the per-query scan).
__device__ __forceinline__ void store_pair(int2* dst, int2 v) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  v2i w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2i*>(dst));
}

// Two consecutive result pairs (16 bytes) in one streaming store; dst is
// 16-byte aligned.
__device__ __forceinline__ void store_pair2(int2* dst, const int4& v) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  const v4i w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<v4i*>(dst));
}

// Order-independent digest of a (tid, pid) pair: a 64-bit mix (the splitmix64
// finaliser) of tid in the high and pid in the low half.  A query's digest is
// the wrapping sum over its pairs.
__device__ __forceinline__ uint64_t pair_mix(int2 p) {
  uint64_t z = (uint64_t(uint32_t(p.x)) << 32) | uint32_t(p.y);
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

}  // namespace gat
