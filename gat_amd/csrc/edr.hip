// EDR kernels for gfx950: the dynamic program of Grid::SimilarityExecuter
// (Grid.cpp:1222-1318) behind top-k trajectory similarity.
//
// The reference's cell rule (Grid.cpp:1288-1300) is not the textbook minimum:
// with M = D[r-1][c-1] + s, U = D[r-1][c] + 1 and L = D[r][c-1] + 1 it returns M
// when M is strictly smallest, L when L is strictly smallest, and U otherwise —
// including M == L < U.  Every kernel here evaluates that rule exactly.
//
// This file holds the plain drop-in for EDRDistance_Batch (kernel.cu:249-501):
// one workgroup per (query, candidate) task, the table swept anti-diagonal by
// anti-diagonal.  Unlike the reference kernel it stages both trajectories in
// LDS once (instead of re-reading global memory for every cell), accepts
// lengths up to GAT_MAX_TRAJ_LEN (the reference stops at 1024, kernel.cu:263),
// and rotates three diagonal buffers so the single barrier per diagonal orders
// every read before the next overwrite (the reference copies state[1] into
// state[0] and rewrites state[1] between barriers, kernel.cu:449-496).
#include "gat_common.hpp"

namespace {

constexpr int kPlainBlock = 256;
constexpr int kMaxLen = GAT_MAX_TRAJ_LEN;

// Grid.cpp:1297-1300.
__device__ __forceinline__ int edr_rule(int m, int u, int l) {
  const bool c1 = (m < u) && (m < l);
  const bool c2 = (l < u) && (l < m);
  return c1 ? m : (c2 ? l : u);
}

// |dx| < u and |dy| < u, where u is the smallest float >= eps: the exact
// single-precision form of the double comparison of Grid.cpp:1281.
__device__ __forceinline__ bool edr_match(float2 a, float2 b, float u) {
  return fabsf(a.x - b.x) < u && fabsf(a.y - b.y) < u;
}

__global__ __launch_bounds__(kPlainBlock) void edr_plain_kernel(
    int ntasks, const gat_edr_task* __restrict__ tasks, const float2* __restrict__ qxy, const int32_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, const float2* __restrict__ cxy, const int64_t* __restrict__ coff,
    const int32_t* __restrict__ clen, float u, int32_t* __restrict__ result) {
  __shared__ float2 rows[kMaxLen];
  __shared__ float2 cols[kMaxLen];
  __shared__ int16_t diag[3][kMaxLen + 2];

  const int t = blockIdx.x;
  if (t >= ntasks) return;
  const gat_edr_task task = tasks[t];
  const int qn = qlen[task.qid];
  const int cn = clen[t];
  // Rows are the shorter trajectory; on equal lengths the query (Grid.cpp:1234-1247).
  const bool query_rows = cn >= qn;
  const int n = query_rows ? qn : cn;
  const int m = query_rows ? cn : qn;
  if (n < 0 || m > kMaxLen) {  // rejected by the host wrapper; never indexes out of bounds
    if (threadIdx.x == 0) result[t] = -1;
    return;
  }
  if (n == 0) {  // D[0][m] = m (Grid.cpp:1269-1272)
    if (threadIdx.x == 0) result[t] = m;
    return;
  }
  const float2* rp = query_rows ? qxy + qoff[task.qid] : cxy + coff[t];
  const float2* cp = query_rows ? cxy + coff[t] : qxy + qoff[task.qid];
  for (int i = threadIdx.x; i < n; i += kPlainBlock) rows[i] = rp[i];
  for (int i = threadIdx.x; i < m; i += kPlainBlock) cols[i] = cp[i];
  // Diagonal d stores D[r][d - r] at index r.
  int16_t* d2 = diag[0];  // diagonal d-2
  int16_t* d1 = diag[1];  // diagonal d-1
  int16_t* d0 = diag[2];  // diagonal d
  if (threadIdx.x == 0) {
    d2[0] = 0;  // D[0][0]
    d1[0] = 1;  // D[0][1]
    d1[1] = 1;  // D[1][0]
  }
  __syncthreads();
  for (int d = 2; d <= n + m; ++d) {
    const int rlo = d - m > 0 ? d - m : 0;
    const int rhi = n < d ? n : d;
    for (int r = rlo + int(threadIdx.x); r <= rhi; r += kPlainBlock) {
      const int c = d - r;
      int v;
      if (r == 0) {
        v = c;  // D[0][c] = c
      } else if (c == 0) {
        v = r;  // D[r][0] = r
      } else {
        const int s = edr_match(rows[r - 1], cols[c - 1], u) ? 0 : 1;
        v = edr_rule(int(d2[r - 1]) + s, int(d1[r - 1]) + 1, int(d1[r]) + 1);
      }
      d0[r] = int16_t(v);
    }
    __syncthreads();
    int16_t* tmp = d2;
    d2 = d1;
    d1 = d0;
    d0 = tmp;
  }
  if (threadIdx.x == 0) result[t] = d1[n];
}

}  // namespace

extern "C" {

int gatEDRDistanceBatchHandler(int queryTaskNum, const gat_edr_task* tasks, const float* query_xy,
                               const int32_t* query_offset, const int32_t* query_length, const float* cand_xy,
                               const int64_t* cand_offset, const int32_t* cand_length, double eps, int32_t* result,
                               void* stream) {
  GAT_REQUIRE(queryTaskNum >= 0, "EDR batch: negative task count");
  if (queryTaskNum == 0) return GAT_OK;
  GAT_REQUIRE(tasks && query_xy && query_offset && query_length && cand_xy && cand_offset && cand_length && result,
              "EDR batch: null pointer");
  GAT_REQUIRE(eps > 0, "EDR batch: eps must be positive");
  edr_plain_kernel<<<dim3(uint32_t(queryTaskNum)), dim3(kPlainBlock), 0, gat::as_stream(stream)>>>(
      queryTaskNum, tasks, reinterpret_cast<const float2*>(query_xy), query_offset, query_length,
      reinterpret_cast<const float2*>(cand_xy), cand_offset, cand_length, gat_eps_threshold(eps), result);
  GAT_LAUNCH_CHECK("edr_plain_kernel");
  return GAT_OK;
}

}  // extern "C"
