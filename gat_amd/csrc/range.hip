// Range refine kernels for gfx950: the per-point containment test behind
// Grid::rangeQueryBatchGPU, STIG::rangeQueryGPU and FSG::rangeQueryBatchGPU
// (reference kernels cudaRangeQueryTest, kernel.cu:1004-1074, and
// cudaRangeQuerySTIG, kernel.cu:1165-1209).
//
// Layout: one wave64 refines one task (a run of consecutive store points and a
// rectangle).  Lanes read the run as float2 (lon, lat) pairs, 64 consecutive
// points per load instruction, up to sixteen loads (1024 points) in flight per
// lane, and test them against the rectangle held in scalar registers.  Hits are
// compacted with wave ballots and per-lane prefixes (mbcnt).  The ids of the
// hit points are gathered before the wave's single atomic on the query's
// counter returns, so that latency overlaps the gathers; every result pair is
// then written exactly once and no pass re-reads the points.
//
// Ids are stored compactly: (tid << pid_bits) | pid in 32 bits when the
// trajectory and point ids fit, otherwise a 32-bit tid array and a 16-bit pid
// array — 4 or 6 bytes per hit instead of the 8 of an (int, int) pair.
#include <atomic>
#include <mutex>

#include "gat_common.hpp"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;               // 4 waves per workgroup
constexpr int kSlots = 16;                // float2 loads in flight per lane
constexpr int kSpan = kWave * kSlots;     // points per wave iteration

struct Task {
  uint32_t start, count, qid;
};
static_assert(sizeof(Task) == sizeof(gat_range_task), "task layout");

struct Store {
  const float2* xy;
  const uint32_t* id;
  const uint16_t* pid;
  uint64_t npoints;
  uint32_t pid_bits;
};

// MBB::pInBox (MBB.cpp:21-26): inclusive on all four edges.
__device__ __forceinline__ bool in_box(const float4& b, float x, float y) {
  return x <= b.z && x >= b.x && y <= b.w && y >= b.y;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <bool kPacked>
__device__ __forceinline__ int2 point_ids(const Store& s, uint64_t i) {
  if constexpr (kPacked) {
    const uint32_t v = s.id[i];
    return make_int2(int(v >> s.pid_bits), int(v & ((1u << s.pid_bits) - 1u)));
  } else {
    return make_int2(int(s.id[i]), int(s.pid[i]));
  }
}

template <bool kPairs, bool kPacked>
__global__ __launch_bounds__(kBlock) void range_refine_kernel(
    Store store, const Task* __restrict__ tasks, uint64_t ntasks, const float4* __restrict__ boxes,
    uint32_t nqueries, const uint64_t* __restrict__ qbase, uint32_t* __restrict__ qcount, int2* __restrict__ pairs,
    uint64_t capacity, uint8_t* __restrict__ bitmap, uint64_t traj_stride) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = uniform((blockIdx.x * kBlock + threadIdx.x) / kWave);
  const uint64_t nwaves = uint64_t(gridDim.x) * (kBlock / kWave);
  for (uint64_t t = wave; t < ntasks; t += nwaves) {
    const Task task = tasks[t];
    const uint32_t qid = uniform(task.qid);
    const uint32_t count = uniform(task.count);
    const uint64_t start = uniform(task.start);
    // Malformed tasks are skipped rather than read or written out of bounds.
    if (qid >= nqueries || start + count > store.npoints) continue;
    const float4 box = boxes[qid];
    const float2* p = store.xy + start;
    for (uint32_t base = 0; base < count; base += kSpan) {
      const uint32_t left = count - base;
      float2 v[kSlots];
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        const uint32_t i = k * kWave + lane;
        v[k] = i < left ? p[base + i] : make_float2(__builtin_nanf(""), 0.0f);
      }
      uint64_t mask[kSlots];
      uint32_t total = 0;
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        mask[k] = __ballot(in_box(box, v[k].x, v[k].y));  // NaN fails every comparison
        total += __popcll(mask[k]);
      }
      if (total == 0) continue;
      if constexpr (kPairs) {
        int2 ids[kSlots];
#pragma unroll
        for (int k = 0; k < kSlots; ++k)
          if ((mask[k] >> lane) & 1) ids[k] = point_ids<kPacked>(store, start + base + k * kWave + lane);
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(qcount + qid, total);
        slot = uniform(slot);
        const uint64_t first = qbase[qid] + slot;
        uint32_t before = 0;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
          const uint64_t at = first + before + lanes_below(mask[k]);
          if (((mask[k] >> lane) & 1) && at < capacity) gat::store_pair(pairs + at, ids[k]);
          before += __popcll(mask[k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < kSlots; ++k)
          if ((mask[k] >> lane) & 1) {
            const uint32_t tid = uint32_t(point_ids<kPacked>(store, start + base + k * kWave + lane).x);
            if (tid < traj_stride) bitmap[uint64_t(qid) * traj_stride + tid] = 1;
          }
      }
    }
  }
}

int check_store(const gat_point_store* s) {
  GAT_REQUIRE(s, "range refine: null store");
  GAT_REQUIRE(s->npoints >= 0, "range refine: negative store size");
  GAT_REQUIRE(s->pid_bits >= 0 && s->pid_bits < 32, "range refine: bad pid_bits");
  GAT_REQUIRE(s->npoints == 0 || (s->xy_dev && s->id_dev && (s->pid_bits > 0 || s->pid_dev)),
              "range refine: null store array");
  return GAT_OK;
}

int launch_refine(bool pairs_mode, const gat_point_store* sp, const gat_range_task* tasks, int64_t ntasks,
                  const float* boxes, int64_t nqueries, const uint64_t* qbase, uint32_t* qcount, int32_t* pairs,
                  int64_t capacity, uint8_t* bitmap, int64_t traj_stride, void* stream) {
  if (int rc = check_store(sp)) return rc;
  GAT_REQUIRE(ntasks >= 0 && nqueries >= 0 && nqueries <= int64_t(UINT32_MAX), "range refine: bad sizes");
  if (ntasks == 0) return GAT_OK;
  GAT_REQUIRE(tasks && boxes, "range refine: null input pointer");
  const Store st{reinterpret_cast<const float2*>(sp->xy_dev), sp->id_dev, sp->pid_dev, uint64_t(sp->npoints),
                 uint32_t(sp->pid_bits)};
  const bool packed = sp->pid_bits > 0;
  // Enough waves to keep every SIMD busy; beyond that the grid-stride loop
  // hands each wave further tasks.
  const int64_t blocks_cap = int64_t(gat::cu_count()) * 16;
  int64_t blocks = (ntasks + (kBlock / kWave) - 1) / (kBlock / kWave);
  if (blocks > blocks_cap) blocks = blocks_cap;
  hipStream_t s = gat::as_stream(stream);
  const dim3 grid{uint32_t(blocks)}, block{uint32_t(kBlock)};
  const auto* T = reinterpret_cast<const Task*>(tasks);
  const auto* B = reinterpret_cast<const float4*>(boxes);
  if (pairs_mode) {
    GAT_REQUIRE(qbase && qcount && pairs && capacity >= 0, "range refine: bad output arguments");
    auto* P = reinterpret_cast<int2*>(pairs);
    if (packed)
      range_refine_kernel<true, true><<<grid, block, 0, s>>>(st, T, uint64_t(ntasks), B, uint32_t(nqueries), qbase,
                                                               qcount, P, uint64_t(capacity), nullptr, 0);
    else
      range_refine_kernel<true, false><<<grid, block, 0, s>>>(st, T, uint64_t(ntasks), B, uint32_t(nqueries), qbase,
                                                                qcount, P, uint64_t(capacity), nullptr, 0);
  } else {
    GAT_REQUIRE(bitmap && traj_stride > 0, "range refine: null bitmap");
    if (packed)
      range_refine_kernel<false, true><<<grid, block, 0, s>>>(st, T, uint64_t(ntasks), B, uint32_t(nqueries), nullptr,
                                                                nullptr, nullptr, 0, bitmap, uint64_t(traj_stride));
    else
      range_refine_kernel<false, false><<<grid, block, 0, s>>>(st, T, uint64_t(ntasks), B, uint32_t(nqueries), nullptr,
                                                                 nullptr, nullptr, 0, bitmap, uint64_t(traj_stride));
  }
  GAT_LAUNCH_CHECK("range_refine_kernel");
  return GAT_OK;
}

// Shared body of the two reference-signature handlers.
int bitmap_handler(const gat_point_store* store, const gat_range_task* tasks, int ntasks, const float* boxes,
                   uint8_t* result_host, int maxTrajNum, int maxQueryNum, void* stream) {
  GAT_REQUIRE(result_host, "range handler: null result buffer");
  GAT_REQUIRE(maxTrajNum > 0 && maxQueryNum >= 0 && ntasks >= 0, "range handler: bad sizes");
  const size_t bytes = size_t(maxTrajNum) * size_t(maxQueryNum);
  if (bytes == 0) return GAT_OK;
  hipStream_t s = gat::as_stream(stream);
  uint8_t* dev = nullptr;
  GAT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dev), bytes));
  int rc = GAT_OK;
  hipError_t e = hipMemsetAsync(dev, 0, bytes, s);
  if (e != hipSuccess) rc = gat::hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__);
  if (rc == GAT_OK)
    rc = launch_refine(false, store, tasks, ntasks, boxes, maxQueryNum, nullptr, nullptr, nullptr, 0, dev, maxTrajNum,
                       stream);
  if (rc == GAT_OK) {
    e = hipMemcpyAsync(result_host, dev, bytes, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = gat::hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  hipError_t es = hipStreamSynchronize(s);
  hipError_t ef = hipFree(dev);
  if (rc == GAT_OK && es != hipSuccess) rc = gat::hip_fail(es, "hipStreamSynchronize", __FILE__, __LINE__);
  if (rc == GAT_OK && ef != hipSuccess) rc = gat::hip_fail(ef, "hipFree", __FILE__, __LINE__);
  return rc;
}

__global__ __launch_bounds__(kBlock) void pairs_digest_kernel(const int2* __restrict__ pairs,
                                                             const uint64_t* __restrict__ qbase,
                                                             const uint32_t* __restrict__ qcount, uint32_t nq,
                                                             uint64_t capacity, uint64_t* __restrict__ out) {
  __shared__ uint64_t part[kBlock / kWave];
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const uint64_t b = qbase[q];
    const uint64_t n = b < capacity ? (qcount[q] < capacity - b ? qcount[q] : capacity - b) : 0;
    uint64_t acc = 0;
    for (uint64_t i = threadIdx.x; i < n; i += kBlock) acc += gat::pair_mix(pairs[b + i]);
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t total = 0;
      for (int w = 0; w < kBlock / kWave; ++w) total += part[w];
      out[q] = total;
    }
    __syncthreads();
  }
}

}  // namespace

// Device bitmap of gatRangeQueryTestHandlerNoAlloc, for callers that cannot
// allocate: one copy per GPU, held until the process ends, shared by every call
// on that GPU (calls that overlap on one GPU race on it).  In host code the
// name `scratch` only stands for the symbol; kernels and copies get the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = size_t(GAT_NOALLOC_BITMAP_MAX);
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::atomic<uint8_t*> g_scratch_ptr[kMaxDevices];

// The current device's copy of `scratch`, resolved once per device (by
// gat_init, or by the first no-alloc call on a device gat_init never saw).
int scratch_address(uint8_t** out) {
  int dev = 0;
  GAT_HIP_CHECK(hipGetDevice(&dev));
  GAT_REQUIRE(dev >= 0 && dev < kMaxDevices, "scratch: device %d beyond the %d supported", dev, kMaxDevices);
  uint8_t* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* addr = nullptr;
      GAT_HIP_CHECK(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
      GAT_REQUIRE(addr, "scratch: hipGetSymbolAddress returned no address");
      p = static_cast<uint8_t*>(addr);
      g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return GAT_OK;
}

}  // namespace

namespace gat {

int scratch_setup() {
  uint8_t* p = nullptr;
  return scratch_address(&p);
}

}  // namespace gat

extern "C" {

int gat_range_refine_bitmap(const gat_point_store* store, const gat_range_task* tasks, int64_t ntasks,
                            const float* boxes, int64_t nqueries, uint8_t* bitmap, int64_t traj_stride, void* stream) {
  return launch_refine(false, store, tasks, ntasks, boxes, nqueries, nullptr, nullptr, nullptr, 0, bitmap, traj_stride,
                       stream);
}

int gat_range_refine_pairs(const gat_point_store* store, const gat_range_task* tasks, int64_t ntasks,
                           const float* boxes, int64_t nqueries, const uint64_t* qbase, uint32_t* qcount,
                           int32_t* pairs, int64_t capacity, void* stream) {
  return launch_refine(true, store, tasks, ntasks, boxes, nqueries, qbase, qcount, pairs, capacity, nullptr, 0,
                       stream);
}

int gatRangeQueryTestHandler(const gat_point_store* store, const gat_range_task* tasks, int stateTableLength,
                             const float* boxes, uint8_t* result_host, int maxTrajNum, int maxQueryNum,
                             void* stream) {
  return bitmap_handler(store, tasks, stateTableLength, boxes, result_host, maxTrajNum, maxQueryNum, stream);
}

int gatRangeQuerySTIGHandler(const gat_point_store* store, const gat_range_task* tasks, int stateTableLength,
                             const float* boxes, uint8_t* result_host, int maxTrajNum, int maxQueryNum,
                             void* stream) {
  return bitmap_handler(store, tasks, stateTableLength, boxes, result_host, maxTrajNum, maxQueryNum, stream);
}

int gatRangeQueryTestHandlerNoAlloc(const gat_point_store* store, const gat_range_task* tasks, int stateTableLength,
                                    const float* boxes, uint8_t* result_host, int maxTrajNum, int maxQueryNum,
                                    void* stream) {
  GAT_REQUIRE(result_host, "range handler (no-alloc): null result buffer");
  GAT_REQUIRE(maxTrajNum > 0 && maxQueryNum >= 0 && stateTableLength >= 0, "range handler (no-alloc): bad sizes");
  const size_t bytes = size_t(maxTrajNum) * size_t(maxQueryNum);
  if (bytes > kScratchBytes) {
    gat::set_error("range handler (no-alloc): a %d x %d bitmap is %zu bytes, above the %zu-byte device scratch",
                   maxQueryNum, maxTrajNum, bytes, kScratchBytes);
    return GAT_ERR_CAPACITY;
  }
  if (bytes == 0) return GAT_OK;
  uint8_t* bitmap = nullptr;
  if (int rc = scratch_address(&bitmap)) return rc;
  hipStream_t s = gat::as_stream(stream);
  // The previous call's bitmap is still there: clear what this call reads.
  int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(bitmap), 0u, int64_t((bytes + 3) / 4), stream);
  if (rc == GAT_OK)
    rc = launch_refine(false, store, tasks, stateTableLength, boxes, maxQueryNum, nullptr, nullptr, nullptr, 0, bitmap,
                       maxTrajNum, stream);
  if (rc == GAT_OK) {
    hipError_t e = hipMemcpyAsync(result_host, bitmap, bytes, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = gat::hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  hipError_t es = hipStreamSynchronize(s);
  if (rc == GAT_OK && es != hipSuccess) rc = gat::hip_fail(es, "hipStreamSynchronize", __FILE__, __LINE__);
  return rc;
}

int gat_pairs_digest(const int32_t* pairs, const uint64_t* qbase, const uint32_t* qcount, int64_t nq,
                     int64_t pairs_capacity, uint64_t* digest, void* stream) {
  GAT_REQUIRE(nq >= 0 && nq <= int64_t(UINT32_MAX), "pairs digest: bad query count");
  if (nq == 0) return GAT_OK;
  GAT_REQUIRE(pairs && qbase && qcount && digest && pairs_capacity >= 0, "pairs digest: bad argument");
  const int64_t blocks = nq < int64_t(gat::cu_count()) * 8 ? nq : int64_t(gat::cu_count()) * 8;
  pairs_digest_kernel<<<dim3(uint32_t(blocks)), dim3(kBlock), 0, gat::as_stream(stream)>>>(
      reinterpret_cast<const int2*>(pairs), qbase, qcount, uint32_t(nq), uint64_t(pairs_capacity), digest);
  GAT_LAUNCH_CHECK("pairs_digest_kernel");
  return GAT_OK;
}

}  // extern "C"
