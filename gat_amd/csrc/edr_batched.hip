// Batched EDR for gfx950: many (query, candidate) pairs in one persistent
// launch.  Same recurrence as Grid::SimilarityExecuter (Grid.cpp:1222-1318);
// no reference counterpart for the batching itself.
//
// Wave-systolic layout.  A wave64 runs four independent pairs, one per 16-lane
// DPP row ("segment"), or one long pair over all 64 lanes.  Inside a segment of
// S lanes, lane l owns DP row r = S*s + l + 1 of strip s and keeps that row's
// point in registers; at step t it computes column c = t - l + 1, so a strip
// sweeps its S rows across all columns in m + S - 1 steps.  The three
// neighbours of a cell come from registers:
//   U = D[r-1][c]   the left lane's value of the previous step (DPP shift)
//   L = D[r][c-1]   the lane's own value of the previous step
//   M = D[r-1][c-1] the left lane's value two steps back, i.e. the U the lane
//                   received one step earlier
// and the cell rule of Grid.cpp:1297-1300 reduces exactly to
//   D = (M == L) ? U : min(M, L, U)
// (M strictly smallest -> M; L strictly smallest -> L; every tie for the
// minimum, and M == L < U in particular, -> U).
// The first lane of a segment takes U + 1 from the row above the strip, kept
// in LDS as 16-bit values of D + 1 and read once per step as a broadcast (the
// value is the `old` operand of the shift, so lane 0 keeps it).  Each lane
// stores its cell plus one into the same buffer at its own column; the strip's
// last lane stores every column last, which leaves its row there for the next
// strip, and each column is read (step c - 1) no later than its first store of
// the strip (lane 0, same step, after the read).  Column points are read
// straight from the NaN-padded trajectory store (64 consecutive points per
// wave access, L1 resident across the lanes that reuse them); a NaN point
// never matches, so lanes before their first column keep D[r][0] = r untouched
// and lanes past the last column compute values nobody reads.  The cell
// D[n][m] is captured by the lane that owns row n at the step it reaches
// column m.
//
// Work: jobs of up to four pairs (one for 64-lane segments) with the same
// strip count, ordered longest first, handed out by one atomic counter per
// launch, so the batch's tail is short and no host round trip separates length
// buckets.
#include <cmath>

#include "gat_common.hpp"

namespace {

constexpr int kBlk = 16;   // steps per feed block
constexpr int kPad = 16;   // NaN points before every trajectory of a padded store

struct Job {
  int32_t first, count;
};
static_assert(sizeof(Job) == sizeof(gat_edr_job), "job layout");

// Stores of different lanes to one LDS column come from different steps'
// store instructions, and the row left behind is the last lane's only if those
// instructions keep program order.  Per lane their addresses differ, so the
// scheduler may reorder them; this barrier lets every other kind of
// instruction (ALU, global loads, LDS reads) move across it, but not LDS
// writes.
__device__ __forceinline__ void keep_store_order() {
  // 0x1 ALU, 0x2 VALU, 0x4 SALU, 0x10 VMEM, 0x20 VMEM read, 0x100 DS read.
  __builtin_amdgcn_sched_barrier(0x1 | 0x2 | 0x4 | 0x10 | 0x20 | 0x100);
}

// Lane l of a segment takes src from lane l - 1; lane 0 keeps old.  Segments
// of 16 lanes are DPP rows (row_shr:1); a 64-lane segment is the wave
// (wave_shr:1).
template <int SEG>
__device__ __forceinline__ int shift_in(int old, int src) {
  static_assert(SEG == 16 || SEG == 64, "segments of 16 or 64 lanes");
  return __builtin_amdgcn_update_dpp(old, src, SEG == 16 ? 0x111 : 0x138, 0xf, 0xf, false);
}

// A launch over a window of the job array: jobs [window[0], window[0] +
// window[1]) of the njobs the array holds, the window written on the device
// (by edr_jobs_kernel) after the launch was queued.
__device__ __forceinline__ int32_t take_window(const int32_t* window, int32_t& njobs) {
  if (!window) return 0;
  const int32_t first = min(max(window[0], 0), njobs);
  njobs = min(njobs - first, max(window[1], 0));
  return first;
}


// Issue priority of a wave by the length of its job's schedule in steps:
// the longest pairs of a batch set its end, so their waves go first.
__device__ __forceinline__ void job_priority(int64_t steps) {
  if (steps >= 16384) __builtin_amdgcn_s_setprio(3);
  else if (steps >= 8192) __builtin_amdgcn_s_setprio(2);
  else if (steps >= 4096) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(0);
}

// SEG lanes per pair: 16 (four pairs per wave) or 64 (one pair per wave, a
// quarter of the strips, for the longest pairs, whose strips would otherwise
// run one after another for milliseconds).  Waves per SIMD each class can hold:
// four 16-lane pairs' rows of 2048 columns (16.8 KB of LDS per wave) cap that
// class at two; the others are held to 128 VGPRs so four fit.
template <int MAXM, int SEG>
__global__ __launch_bounds__(64, (MAXM >= 2048 && SEG == 16) ? 2 : 4) void edr_wave_kernel(const int2* __restrict__ pairs, int64_t npairs,
                                                      const Job* __restrict__ jobs, int32_t njobs,
                                                      const int32_t* __restrict__ window,
                                                      int32_t* __restrict__ counter, int32_t nqueries, int32_t ncands,
                                                      const float2* __restrict__ qxy, const int64_t* __restrict__ qoff,
                                                      const int32_t* __restrict__ qlen, const float2* __restrict__ cxy,
                                                      const int64_t* __restrict__ coff,
                                                      const int32_t* __restrict__ clen, float u,
                                                      int32_t* __restrict__ out) {
  constexpr int kSeg = SEG;        // lanes per pair
  constexpr int kSegs = 64 / SEG;  // pairs per wave
  // Row above the current strip plus one, D[r0][c] + 1, at [c + kOff][segment]
  // for columns 2 - SEG .. MAXM + SEG + kBlk - 1 (the lanes' look-behind, the
  // columns, and the last blocks' overrun).  Segments interleave, so one
  // segment's consecutive columns are 2 * kSegs bytes apart and each step's
  // feed is a load of its own.
  constexpr int kOff = SEG;
  constexpr int kRows = MAXM + 2 * SEG + kBlk;
  __shared__ int16_t above[kRows][kSegs];
  const int lane = threadIdx.x;
  const int seg = lane / kSeg;
  const int l = lane & (kSeg - 1);
  const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
  jobs += take_window(window, njobs);
  for (;;) {
    int ji = 0;
    if (lane == 0) ji = atomicAdd(counter, 1);
    ji = __builtin_amdgcn_readfirstlane(ji);
    if (ji < 0 || ji >= njobs) break;  // a counter that was not cleared hands out nothing
    const Job job = jobs[ji];
    // Jobs and pairs that point outside their arrays are skipped.
    const bool live = seg < job.count && job.first >= 0 && int64_t(job.first) + seg < npairs;
    int n = 0, m = 0;
    // Dead segments read rows nowhere (n = 0) and columns from the store's
    // first trajectory, whose leading padding covers the lanes' look-behind
    // (loads further back than the padding are replaced by NaN, below).
    const float2* R = cxy;
    const float2* C = cxy + kPad;
    bool valid = false;
    if (live) {
      const int2 pr = pairs[job.first + seg];
      const bool ids_ok = pr.x >= 0 && pr.x < nqueries && pr.y >= 0 && pr.y < ncands;
      const int nq = ids_ok ? qlen[pr.x] : 0, nc = ids_ok ? clen[pr.y] : 0;
      // Lengths outside 1..MAXM (and bad ids) are reported as -1 and never indexed.
      valid = nq >= 1 && nc >= 1 && nq <= MAXM && nc <= MAXM;
      if (valid) {
        // Rows are the shorter trajectory; on equal lengths the query
        // (Grid.cpp:1234-1247).
        if (nc >= nq) {
          R = qxy + qoff[pr.x]; n = nq; C = cxy + coff[pr.y]; m = nc;
        } else {
          R = cxy + coff[pr.y]; n = nc; C = qxy + qoff[pr.x]; m = nq;
        }
      }
    }
    // Wave-uniform extents: the largest column count and strip count of the job.
    const int m0 = __builtin_amdgcn_readlane(m, 0), m1 = __builtin_amdgcn_readlane(m, 16);
    const int m2 = __builtin_amdgcn_readlane(m, 32), m3 = __builtin_amdgcn_readlane(m, 48);
    const int mmax = max(max(m0, m1), max(m2, m3));
    const int n0 = __builtin_amdgcn_readlane(n, 0), n1 = __builtin_amdgcn_readlane(n, 16);
    const int n2 = __builtin_amdgcn_readlane(n, 32), n3 = __builtin_amdgcn_readlane(n, 48);
    const int nmax = max(max(n0, n1), max(n2, n3));
    // Smallest column and row counts among the live segments: capture starts
    // at the first strip that is some segment's last and the first column
    // that is some segment's last.
    int mmin = mmax, nmin = nmax;
    if (n0 > 0) { mmin = min(mmin, m0); nmin = min(nmin, n0); }
    if (n1 > 0) { mmin = min(mmin, m1); nmin = min(nmin, n1); }
    if (n2 > 0) { mmin = min(mmin, m2); nmin = min(nmin, n2); }
    if (n3 > 0) { mmin = min(mmin, m3); nmin = min(nmin, n3); }
    const int first_last_strip = (nmin + kSeg - 1) / kSeg - 1;
    const int strips = (nmax + kSeg - 1) / kSeg;
    const int my_last = (n + kSeg - 1) / kSeg - 1;
    const int steps = mmax + kSeg - 1;
    job_priority(int64_t(strips) * steps);
    // Row 0: D[0][c] + 1 = c + 1.  (Kept rolled: unrolled, its length grows with MAXM
    // and so does the register count.)
#pragma unroll 2
    for (int i = l; i < kRows; i += kSeg) above[i][seg] = int16_t(i - kOff + 1);
    int res = 0;
    for (int s = 0; s < strips; ++s) {
      // Orders the previous strip's (or the initial) writes of `above` before
      // this strip's reads of them by other lanes.
      __syncthreads();
      const int r = kSeg * s + l + 1;
      const float2 rp = r <= n ? R[r - 1] : nan2;
      const float rx = rp.x, ry = rp.y;
      int c1 = r + 1;  // D[r][c - 1] + 1, from D[r][0] = r
      int u1p = r;     // D[r-1][c - 1] + 1, from D[r-1][0] + 1
      const int tcap = (s == my_last) ? m + l - 1 : -1;  // step at which lane l holds D[r][m]
      const bool last_strip_somewhere = s >= first_last_strip;
      const float2* cp = C - l;  // column point of step t is cp[t]
      // Column points of steps tb .. tb + 15.  Lanes look back up to SEG - 1
      // points before the trajectory.  A lane whose block starts more than
      // kPad points before it (64-lane segments, first blocks only) has all 16
      // points before it: it takes NaN, never matching, and loads from the
      // padding instead.
      auto load_block = [&](float2(&dst)[kBlk], int tb) __attribute__((always_inline)) {
        if (SEG <= kPad + 1 || tb - (kSeg - 1) >= -kPad) {
#pragma unroll
          for (int j = 0; j < kBlk; ++j) dst[j] = cp[tb + j];
        } else {
          const bool before = tb - l < -kPad;
          const float2* bp = before ? C - kPad : cp + tb;
#pragma unroll
          for (int j = 0; j < kBlk; ++j) {
            const float2 v = bp[j];
            dst[j] = before ? nan2 : v;
          }
        }
      };
      // One feed block of kBlk steps over the column points in `src` (steps
      // t0 .. t0 + 15) while the next block's points load into `dst`.  Two
      // register sets alternate, so no block copies one into the other.
      auto sweep = [&](const float2(&src)[kBlk], float2(&dst)[kBlk], int t0) __attribute__((always_inline)) {
        // Feed: U + 1 for the segment's first lane at step t0 + j (column
        // t0 + j + 1), read for the whole block before its first store: this
        // strip stores column c first at step c - 1, and no earlier block of
        // the strip stores past column t0.
        const int16_t(*feed)[kSegs] = &above[kOff + t0 + 1];
        int fv[kBlk];
#pragma unroll
        for (int j = 0; j < kBlk; ++j) fv[j] = feed[j][seg];
        // Every lane stores its cell plus one at its column, t0 + j - l + 1.
        // The last lane stores a column last, so after the strip the buffer
        // holds the strip's last row (columns up to the widest segment's last);
        // lane 0 stores column c at the step that reads its feed, after the
        // read.
        int16_t(*own)[kSegs] = &above[kOff + t0 - l + 1];
        load_block(dst, t0 + kBlk);
        const bool capture = last_strip_somewhere && t0 + kBlk - 1 >= mmin - 1;
        if (capture) {
#pragma unroll
          for (int j = 0; j < kBlk; ++j) {
            const bool match = fabsf(rx - src[j].x) < u && fabsf(ry - src[j].y) < u;
            const int u1 = shift_in<SEG>(fv[j], c1);
            const int mm = u1p - (match ? 1 : 0);
            const int mn = min(min(u1, c1), mm);
            const int d = (mm == c1) ? u1 : mn;
            res = (t0 + j == tcap) ? d : res;
            c1 = d + 1;
            u1p = u1;
            own[j][seg] = int16_t(c1);
            keep_store_order();
          }
        } else {
#pragma unroll
          for (int j = 0; j < kBlk; ++j) {
            const bool match = fabsf(rx - src[j].x) < u && fabsf(ry - src[j].y) < u;
            const int u1 = shift_in<SEG>(fv[j], c1);
            const int mm = u1p - (match ? 1 : 0);
            const int mn = min(min(u1, c1), mm);
            const int d = (mm == c1) ? u1 : mn;
            c1 = d + 1;
            u1p = u1;
            own[j][seg] = int16_t(c1);
            keep_store_order();
          }
        }
      };
      float2 colA[kBlk], colB[kBlk];
      load_block(colA, 0);
      for (int t0 = 0; t0 < steps; t0 += 2 * kBlk) {
        sweep(colA, colB, t0);
        if (t0 + kBlk >= steps) break;
        sweep(colB, colA, t0 + kBlk);
      }
    }
    __syncthreads();  // the next job re-initialises `above`
    if (live && !valid && l == 0) out[job.first + seg] = -1;
    if (valid && l == ((n - 1) & (kSeg - 1))) out[job.first + seg] = res;
  }
}

// One pair over W waves of a workgroup (no reference counterpart): wave w
// runs the pair's 64-row strips w, w + W, ... in lockstep with the others,
// one block of 16 steps per iteration and a workgroup barrier between
// iterations, so a pair's strips overlap instead of running one after another.
// A strip reads the row above it from the previous strip's wave: that wave's
// last lane writes its row, D + 1 per column, into the wave's own LDS row
// buffer at step c + 62 for column c, and the consumer reads column c at step
// c - 1 of its own strip (the feed of the block holding that step, loaded when
// the block starts).  So a strip starts kLag >= 5 blocks after the previous
// one (column 16b + 16, the last a consumer block b reads, is written in the
// producer's block b + 4), and no sooner than its wave has finished its
// previous strip — which also keeps a wave's writes to its buffer behind the
// next wave's reads of the previous contents (those reads start kLag blocks
// earlier on every column).  Every wave runs the same number of iterations,
// so the barriers always match.  The cell rule, the row/column orientation
// and the capture of D[n][m] are edr_wave_kernel's with 64-lane segments.
constexpr int kLag = 6;

template <int MAXM, int W>
__global__ __launch_bounds__(64 * W) void edr_pipe_kernel(const int2* __restrict__ pairs, int64_t npairs,
                                                          const Job* __restrict__ jobs, int32_t njobs,
                                                          const int32_t* __restrict__ window,
                                                          int32_t* __restrict__ counter, int32_t nqueries,
                                                          int32_t ncands, const float2* __restrict__ qxy,
                                                          const int64_t* __restrict__ qoff,
                                                          const int32_t* __restrict__ qlen,
                                                          const float2* __restrict__ cxy,
                                                          const int64_t* __restrict__ coff,
                                                          const int32_t* __restrict__ clen, float u,
                                                          int32_t* __restrict__ out) {
  // Columns -62 .. 0 sit before column 0 (lane 63's look-behind); a block's
  // feed reads up to column 16 * nblocks <= m + 78 (the last block's overrun).
  constexpr int kOff = 64;
  constexpr int kLen = kOff + MAXM + 96;
  __shared__ int16_t rows[W][kLen];
  // The column trajectory in LDS, columns -63 .. 16 * nblocks - 1 (<= m + 77)
  // at [kColOff + column]: a pair over W waves keeps one wave per SIMD, too
  // few to hide a global load's latency behind a block's 16 steps, so each
  // block's points come from LDS instead.
  constexpr int kColOff = 64;
  constexpr int kColLen = kColOff + MAXM + 80;
  __shared__ float2 cols[kColLen];
  __shared__ int32_t job_slot;
  const int wv = threadIdx.x / 64;
  const int l = threadIdx.x & 63;
  const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
  jobs += take_window(window, njobs);
  for (;;) {
    if (threadIdx.x == 0) job_slot = atomicAdd(counter, 1);
    __syncthreads();
    const int ji = job_slot;
    __syncthreads();  // every wave has read the slot before the next job overwrites it
    if (ji < 0 || ji >= njobs) break;  // a counter that was not cleared hands out nothing
    const Job job = jobs[ji];
    // One pair per job; everything below is uniform across the workgroup.
    const bool live = job.count >= 1 && job.first >= 0 && int64_t(job.first) < npairs;
    int n = 0, m = 0;
    const float2* R = cxy;
    const float2* C = cxy + kPad;
    bool valid = false;
    if (live) {
      const int2 pr = pairs[job.first];
      const bool ids_ok = pr.x >= 0 && pr.x < nqueries && pr.y >= 0 && pr.y < ncands;
      const int nq = ids_ok ? qlen[pr.x] : 0, nc = ids_ok ? clen[pr.y] : 0;
      valid = nq >= 1 && nc >= 1 && nq <= MAXM && nc <= MAXM;
      if (valid) {
        if (nc >= nq) {
          R = qxy + qoff[pr.x]; n = nq; C = cxy + coff[pr.y]; m = nc;
        } else {
          R = cxy + coff[pr.y]; n = nc; C = qxy + qoff[pr.x]; m = nq;
        }
      }
    }
    if (!valid) {
      if (live && threadIdx.x == 0) out[job.first] = -1;
      continue;
    }
    const int strips = (n + 63) / 64;
    const int steps = m + 63;
    const int nblocks = (steps + kBlk - 1) / kBlk;
    // Strip s starts at block (s / W) * P + (s % W) * kLag, P = max(nblocks,
    // W * kLag): kLag after strip s - 1 and P >= nblocks after the same wave's
    // previous strip.
    const int P = max(nblocks, W * kLag);
    const int kend = ((strips - 1) / W) * P + ((strips - 1) % W) * kLag + nblocks;
    job_priority(int64_t(kend) * kBlk);
    // Stage the columns the blocks read (lane l of block tb reads columns
    // tb - l .. tb - l + 15); those more than kPad before the trajectory are
    // NaN, never matching.
    for (int i = int(threadIdx.x) - 63; i < nblocks * kBlk; i += 64 * W) cols[kColOff + i] = i < -kPad ? nan2 : C[i];
    __syncthreads();
    // Per-lane state of the wave's current strip.
    int cur = -1;       // strip
    int c1 = 0, u1p = 0, tcap = -1, res = 0;
    float rx = 0.0f, ry = 0.0f;
    // The row point of the wave's next strip, loaded a strip ahead.
    float2 next_row = 64 * wv + l + 1 <= n ? R[64 * wv + l] : nan2;
    float2 colA[kBlk], colB[kBlk];
    auto load_block = [&](float2(&dst)[kBlk], int tb) __attribute__((always_inline)) {
      const float2* cl = &cols[kColOff + tb - l];
#pragma unroll
      for (int j = 0; j < kBlk; ++j) dst[j] = cl[j];
    };
    // One iteration: the wave's block of its current strip, if any, with the
    // block's column points in src and the next block's loading into dst.
    auto step = [&](int k, float2(&src)[kBlk], float2(&dst)[kBlk]) __attribute__((always_inline)) {
      const int d = k - wv * kLag;
      if (d < 0) return;
      const int i = d / P;
      const int b = d - i * P;
      const int s = wv + i * W;
      if (b >= nblocks || s >= strips) return;
      const int t0 = b * kBlk;
      if (b == 0) {  // a new strip: row point, D[r][0] + 1, D[r-1][0] + 1
        cur = s;
        const int r = 64 * s + l + 1;
        rx = next_row.x;
        ry = next_row.y;
        const int rn = r + 64 * W;  // the wave's next strip
        next_row = rn <= n ? R[rn - 1] : nan2;
        c1 = r + 1;
        u1p = r;
        tcap = (s == strips - 1) ? m + l - 1 : -1;
        load_block(src, 0);
      }
      // Feed: D[r0][t0 + j + 1] + 1 from the previous strip's wave (row 0: c + 1).
      int fv[kBlk];
      if (s == 0) {
#pragma unroll
        for (int j = 0; j < kBlk; ++j) fv[j] = t0 + j + 2;
      } else {
        const int16_t* feed = &rows[(s - 1) % W][kOff + t0 + 1];
#pragma unroll
        for (int j = 0; j < kBlk; ++j) fv[j] = feed[j];
      }
      if (t0 + kBlk < steps) load_block(dst, t0 + kBlk);
      int16_t* own = &rows[wv][kOff + t0 - 63 + 1];
      // A wave here is often alone on its SIMD, so each step waits on its
      // dependent chain (shift, min, select, +1).  The block's matches are
      // tested before the chain and the last lane's cells are stored after
      // it, so nothing else (no per-step branch for that lane's store) sits
      // in the chain's way; the stored columns are still written within the
      // producer block the consumers' lag allows for.
      int mt[kBlk];  // 1 where the step's points match (values, not lane masks: those would crowd the SGPRs)
#pragma unroll
      for (int j = 0; j < kBlk; ++j) mt[j] = (fabsf(rx - src[j].x) < u && fabsf(ry - src[j].y) < u) ? 1 : 0;
      int out_c[kBlk];
#pragma unroll
      for (int j = 0; j < kBlk; ++j) {
        const int u1 = shift_in<64>(fv[j], c1);
        const int mm = u1p - mt[j];
        const int mn = min(min(u1, c1), mm);
        const int d = (mm == c1) ? u1 : mn;
        res = (t0 + j == tcap) ? d : res;
        c1 = d + 1;
        u1p = u1;
        out_c[j] = c1;
      }
      if (l == 63) {
#pragma unroll
        for (int j = 0; j < kBlk; ++j) own[j] = int16_t(out_c[j]);
      }
    };
    for (int k = 0; k < kend; k += 2) {
      step(k, colA, colB);
      __syncthreads();
      step(k + 1, colB, colA);
      __syncthreads();
    }
    if (cur == strips - 1 && l == ((n - 1) & 63)) out[job.first] = res;
    __syncthreads();  // the next job's strips reuse the row buffers
  }
}

// Candidate filter: a (query, candidate) pair can contain a matching point
// pair only if neither box lies a float-difference of u or more beyond the
// other on either axis.  Float subtraction is monotone, so for points a (query)
// and b (candidate) with b.x >= cand.xmin and a.x <= query.xmax,
// fl(b.x - a.x) >= fl(cand.xmin - query.xmax): when the latter is >= u no point
// pair matches on x, and likewise for the other three sides.  Pairs the filter
// drops have no match at all, so their distance is the all-mismatch value.
__device__ __forceinline__ bool may_match(const float4& q, const float4& c, float u) {
  return !((c.x - q.z) >= u || (q.x - c.z) >= u || (c.y - q.w) >= u || (q.y - c.w) >= u);
}

// A workgroup tests kFilterQ queries against its candidates: each lane keeps
// its candidate's box in registers across the queries, so the candidate boxes
// are read once per kFilterQ queries instead of once per query.  The order of
// the output pairs is not defined.
constexpr int kFilterQ = 16;

__global__ __launch_bounds__(256) void edr_filter_kernel(const float4* __restrict__ qbox, int32_t nq,
                                                         const float4* __restrict__ cbox, int32_t nc, float u,
                                                         int2* __restrict__ out, uint64_t capacity,
                                                         unsigned long long* __restrict__ count) {
  __shared__ float4 qs[kFilterQ];
  const int q0 = int(blockIdx.y) * kFilterQ;
  if (threadIdx.x < kFilterQ && q0 + int(threadIdx.x) < nq) qs[threadIdx.x] = qbox[q0 + int(threadIdx.x)];
  __syncthreads();
  const int nqb = min(kFilterQ, nq - q0);
  const int lane = threadIdx.x & 63;
  for (int c = blockIdx.x * 256 + threadIdx.x; c - lane < nc; c += gridDim.x * 256) {
    const float4 cb = c < nc ? cbox[c] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int j = 0; j < nqb; ++j) {
      const bool keep = c < nc && may_match(qs[j], cb, u);
      const uint64_t mask = __ballot(keep);
      if (mask == 0) continue;
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(count, (unsigned long long)__popcll(mask));
      base = __builtin_amdgcn_readfirstlane(uint32_t(base)) |
             (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(base >> 32))) << 32);
      const uint64_t at = base + __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
      if (keep && at < capacity) out[at] = make_int2(q0 + j, c);
    }
  }
}

// The same filter through a grid over the candidates' boxes (no reference
// counterpart): every candidate whose box spans at most kGridMaxCells cells
// is listed in each cell it touches (gat_edr_cand_grid, built by the host);
// a query tests the candidates of the cells its box, widened by u and one
// cell on every side, touches, so it sees every candidate that could pass
// may_match (keep implies c.xmin < q.xmax + u and c.xmax > q.xmin - u in
// exact arithmetic, and the widening covers the rounding of the cell of
// q.xmax + u).  A pair is reported by one cell only: the one at (max(cell of
// c.xmin, first column), max(cell of c.ymin, first row)), which both the
// candidate's cells and the query's hold.  Candidates the grid does not list
// (boxes not finite, inverted or too wide) are tested by every query, and a
// query whose box is not finite or inverted tests every candidate.  One wave
// per query; the output set equals edr_filter_kernel's.
struct CandGrid {
  int32_t g;
  float x0, y0, inv_x, inv_y;
  const uint32_t* start;  // g * g + 1 entry offsets, cells row by row
  const int32_t* id;
  const float4* box;
  const int32_t* always;
  int32_t nalways;
};

// Cell column (or row) of v: floor((v - o) * inv) clamped to 0 .. g - 1 (the
// host lists candidates with the same float operations).
__device__ __forceinline__ int grid_cell(float v, float o, float inv, int g) {
  const float t = floorf((v - o) * inv);
  return int(fminf(fmaxf(t, 0.0f), float(g - 1)));
}

constexpr int kGridFilterWaves = 4;

__global__ __launch_bounds__(64 * kGridFilterWaves) void edr_filter_grid_kernel(
    const float4* __restrict__ qbox, int32_t nq, const float4* __restrict__ cbox, int32_t nc, CandGrid gr, float u,
    int2* __restrict__ out, uint64_t capacity, unsigned long long* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int q = int(blockIdx.x) * kGridFilterWaves + int(threadIdx.x >> 6);
  if (q >= nq) return;  // the whole wave
  const float4 qb = qbox[q];
  auto emit = [&](bool keep, int c) __attribute__((always_inline)) {
    const uint64_t mask = __ballot(keep);
    if (mask == 0) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, (unsigned long long)__popcll(mask));
    base = __builtin_amdgcn_readfirstlane(uint32_t(base)) |
           (uint64_t(__builtin_amdgcn_readfirstlane(uint32_t(base >> 32))) << 32);
    const uint64_t at = base + __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
    if (keep && at < capacity) out[at] = make_int2(q, c);
  };
  const bool boxed = isfinite(qb.x) && isfinite(qb.y) && isfinite(qb.z) && isfinite(qb.w) && qb.x <= qb.z &&
                     qb.y <= qb.w;
  if (!boxed) {
    for (int c0 = 0; c0 < nc; c0 += 64) {
      const int c = c0 + lane;
      emit(c < nc && may_match(qb, cbox[c], u), c);
    }
    return;
  }
  for (int a0 = 0; a0 < gr.nalways; a0 += 64) {
    const int a = a0 + lane;
    const int c = a < gr.nalways ? gr.always[a] : 0;
    emit(a < gr.nalways && c >= 0 && c < nc && may_match(qb, cbox[c], u), c);
  }
  const int cx0 = max(grid_cell(qb.x - u, gr.x0, gr.inv_x, gr.g) - 1, 0);
  const int cx1 = min(grid_cell(qb.z + u, gr.x0, gr.inv_x, gr.g) + 1, gr.g - 1);
  const int cy0 = max(grid_cell(qb.y - u, gr.y0, gr.inv_y, gr.g) - 1, 0);
  const int cy1 = min(grid_cell(qb.w + u, gr.y0, gr.inv_y, gr.g) + 1, gr.g - 1);
  const int wx = cx1 - cx0 + 1, ncells = wx * (cy1 - cy0 + 1);
  for (int k0 = 0; k0 < ncells; k0 += 64) {
    // Lane k takes cell k0 + k; its entries are numbered after the lower lanes'.
    const int k = k0 + lane;
    int cellx = 0, celly = 0;
    uint32_t s = 0, cnt = 0;
    if (k < ncells) {
      cellx = cx0 + k % wx;
      celly = cy0 + k / wx;
      const int cell = celly * gr.g + cellx;
      s = gr.start[cell];
      cnt = gr.start[cell + 1] - s;
    }
    uint32_t incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = uint32_t(__shfl_up(int(incl), d, 64));
      if (lane >= d) incl += y;
    }
    const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
    for (uint32_t i0 = 0; i0 < total; i0 += 64) {
      const uint32_t i = i0 + uint32_t(lane);
      // The lane whose cell holds entry i: the first with incl > i.
      int j = 0;
      for (int step = 32; step >= 1; step >>= 1)
        if (uint32_t(__shfl(int(incl), j + step - 1, 64)) <= i) j += step;
      j = min(j, 63);
      const uint32_t sj = uint32_t(__shfl(int(s), j, 64));
      const uint32_t before = uint32_t(__shfl(int(incl), j, 64)) - uint32_t(__shfl(int(cnt), j, 64));
      const int jx = __shfl(cellx, j, 64), jy = __shfl(celly, j, 64);
      bool keep = false;
      int c = 0;
      if (i < total) {
        const uint32_t e = sj + (i - before);
        const float4 b = gr.box[e];
        c = gr.id[e];
        const int rx = max(grid_cell(b.x, gr.x0, gr.inv_x, gr.g), cx0);
        const int ry = max(grid_cell(b.y, gr.y0, gr.inv_y, gr.g), cy0);
        keep = rx == jx && ry == jy && c >= 0 && c < nc && may_match(qb, b, u);
      }
      emit(keep, c);
    }
  }
}

// The all-mismatch table: out[r * (maxlen + 1) + c] = D[r][c] of the recurrence
// with every substitution costing 1, for 0 <= r, c <= maxlen.  A pair of
// lengths n <= m with no matching point has distance out[n][m].  One
// workgroup sweeps the anti-diagonals.
__global__ __launch_bounds__(1024) void edr_mismatch_table_kernel(int32_t maxlen, int32_t* __restrict__ out) {
  __shared__ int32_t dg[3][GAT_MAX_TRAJ_LEN + 2];
  const int w = maxlen + 1;
  int32_t* d2 = dg[0];
  int32_t* d1 = dg[1];
  int32_t* d0 = dg[2];
  for (int d = 0; d <= 2 * maxlen; ++d) {
    const int rlo = d - maxlen > 0 ? d - maxlen : 0;
    const int rhi = d < maxlen ? d : maxlen;
    for (int r = rlo + int(threadIdx.x); r <= rhi; r += 1024) {
      const int c = d - r;
      int v;
      if (r == 0) v = c;
      else if (c == 0) v = r;
      else {
        const int mm = d2[r - 1] + 1, uu = d1[r - 1] + 1, ll = d1[r] + 1;
        const bool c1 = mm < uu && mm < ll, c2 = ll < uu && ll < mm;  // Grid.cpp:1297-1300
        v = c1 ? mm : (c2 ? ll : uu);
      }
      d0[r] = v;
      out[int64_t(r) * w + c] = v;
    }
    __syncthreads();
    int32_t* t = d2;
    d2 = d1;
    d1 = d0;
    d0 = t;
  }
}

// Exact top-k merge of one query per workgroup (no reference counterpart: the
// reference keeps a k-entry priority queue per query on the host,
// Grid.cpp:1600-1729).  A candidate's distance is its DP distance when it
// survived the filter, else the all-mismatch value of its length pair.  (The
// reference's cell rule is not a minimum, so a DP distance can exceed that
// value: survivors are never assumed to lie below it.)  The distances are
// histogrammed per trajectory length (at most GAT_MAX_TRAJ_LEN lengths,
// whatever the database size), each survivor moved from its length's value to
// its DP distance, and the cut T is the smallest distance whose running count
// reaches kk.  Every candidate below T is taken (fewer than kk: the
// non-survivors of the lengths whose value is below T, the survivors whose DP
// distance is), then the r smallest ids at distance T, found by a binary
// search for the id bound X over the ascending id lists of the lengths whose
// value is T (their survivors left out) and the survivors at T; the search
// range and each list's part of it are cut to the lists' first r + (number of
// survivors) ids.  The kk results are ranked by (distance, id).
constexpr int kTopkMax = 1024;
constexpr int kLenBins = GAT_MAX_TRAJ_LEN + 1;     // lengths 0..GAT_MAX_TRAJ_LEN
constexpr int kDistBins = 2 * GAT_MAX_TRAJ_LEN + 1;  // distances 0..n + m
constexpr int kTopkBlock = 256;

// Index of c in the ascending sid[s0, s1), or -1.
__device__ __forceinline__ int32_t survivor_index(const int32_t* __restrict__ sid, int32_t s0, int32_t s1, int32_t c) {
  int32_t lo = s0, hi = s1;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (sid[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < s1 && sid[lo] == c) ? lo : -1;
}

// Number of entries <= x in the ascending ids[0, n).
__device__ __forceinline__ int32_t count_le(const int32_t* __restrict__ ids, int32_t n, int32_t x) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (ids[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Sum over the workgroup, returned to every thread.
__device__ __forceinline__ int32_t topk_block_sum(int32_t v, int32_t* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();  // every thread has read red from the previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(kTopkBlock) void edr_topk_kernel(
    const int32_t* __restrict__ ftable, int32_t w, const int32_t* __restrict__ qlen,
    const int32_t* __restrict__ clen, int32_t nc, const int32_t* __restrict__ len_ids,
    const int32_t* __restrict__ len_off, const int32_t* __restrict__ sstart, const int32_t* __restrict__ sid,
    const int32_t* __restrict__ sdist, int32_t k, int32_t kk, int32_t* __restrict__ out_ids,
    int32_t* __restrict__ out_dist, int32_t* __restrict__ status) {
  __shared__ int32_t val[kLenBins];    // the query's all-mismatch value per candidate length
  __shared__ int32_t hist[kDistBins];  // candidates per distance
  __shared__ int32_t out_d[kTopkMax], out_i[kTopkMax];
  __shared__ int32_t red[kTopkBlock / 64];
  __shared__ int32_t n_out, sh_T, sh_below, sh_nt;
  const int q = blockIdx.x;
  const int tid = threadIdx.x;
  const int a = qlen[q];
  if (a < 1 || a > GAT_MAX_TRAJ_LEN) {
    if (tid == 0) atomicOr(status, 4);
    return;
  }
  for (int L = tid; L < kLenBins; L += kTopkBlock) val[L] = L >= 1 ? ftable[int64_t(min(a, L)) * w + max(a, L)] : 0;
  for (int b = tid; b < kDistBins; b += kTopkBlock) hist[b] = 0;
  if (tid == 0) {
    n_out = 0;
    sh_T = -1;
    sh_below = 0;
  }
  __syncthreads();
  for (int L = 1 + tid; L < kLenBins; L += kTopkBlock) {
    const int32_t cnt = len_off[L + 1] - len_off[L];
    const int32_t v = val[L];
    if (cnt > 0) {
      if (v < 0 || v >= kDistBins) atomicOr(status, 1);
      else atomicAdd(&hist[v], cnt);
    }
  }
  const int32_t s0 = sstart[q], s1 = sstart[q + 1];
  for (int32_t i = s0 + tid; i < s1; i += kTopkBlock) {
    const int32_t c = sid[i];
    const int32_t L = (c >= 0 && c < nc) ? clen[c] : 0;
    const int32_t d = sdist[i];
    if (L < 1 || L > GAT_MAX_TRAJ_LEN || d < 0 || d >= kDistBins || val[L] >= kDistBins) {
      atomicOr(status, 1);
      continue;
    }
    atomicSub(&hist[val[L]], 1);
    atomicAdd(&hist[d], 1);
  }
  __syncthreads();
  // The cut: thread t scans distances [kPer t, kPer (t + 1)).
  constexpr int kPer = (kDistBins + kTopkBlock - 1) / kTopkBlock;
  const int lane = tid & 63, wv = tid >> 6;
  int32_t local = 0;
  for (int j = 0; j < kPer; ++j) {
    const int b = tid * kPer + j;
    if (b < kDistBins) local += hist[b];
  }
  int32_t incl = local;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) red[wv] = incl;
  __syncthreads();
  int32_t before = incl - local;
  for (int j = 0; j < wv; ++j) before += red[j];
  if (before < kk && before + local >= kk) {
    int32_t acc = before;
    for (int j = 0; j < kPer; ++j) {
      const int b = tid * kPer + j;
      if (b >= kDistBins) break;
      if (acc + hist[b] >= kk) {
        sh_T = b;
        sh_below = acc;
        break;
      }
      acc += hist[b];
    }
  }
  __syncthreads();
  const int32_t T = sh_T;
  const int32_t r = kk - sh_below;
  if (T < 0) {
    if (tid == 0) atomicOr(status, 2);
    return;
  }
  // The distance histogram is done with: its space now holds the lengths
  // valued T (when they fit) and this query's survivor ids, ascending (when
  // they fit), so survivor lookups stay in LDS.
  constexpr int32_t kTlMax = 2048, kStMax = kDistBins - kTlMax;
  int32_t* tl = hist;             // lengths whose value is T
  int32_t* st = hist + kTlMax;    // staged survivor ids
  const int32_t ns = s1 - s0;
  const bool staged = ns <= kStMax;
  if (tid == 0) sh_nt = 0;
  __syncthreads();
  if (staged)
    for (int32_t i = tid; i < ns; i += kTopkBlock) st[i] = sid[s0 + i];
  for (int L = 1 + tid; L < kLenBins; L += kTopkBlock)
    if (val[L] == T && len_off[L + 1] > len_off[L]) {
      const int32_t at = atomicAdd(&sh_nt, 1);
      if (at < kTlMax) tl[at] = L;
    }
  __syncthreads();
  const int32_t nt = sh_nt;
  auto is_survivor = [&](int32_t c) -> bool {
    if (ns == 0) return false;
    if (!staged) return survivor_index(sid, s0, s1, c) >= 0;
    int32_t lo = 0, hi = ns;
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (hist[kTlMax + mid] < c) lo = mid + 1;
      else hi = mid;
    }
    return lo < ns && hist[kTlMax + lo] == c;
  };
  // Every candidate below T: non-survivors by length, survivors by DP distance.
  for (int L = 1 + tid; L < kLenBins; L += kTopkBlock) {
    if (val[L] >= T) continue;
    for (int32_t i = len_off[L], e = len_off[L + 1]; i < e; ++i) {
      const int32_t c = len_ids[i];
      if (is_survivor(c)) continue;
      const int32_t pos = atomicAdd(&n_out, 1);
      if (pos < kTopkMax) {
        out_d[pos] = val[L];
        out_i[pos] = c;
      }
    }
  }
  for (int32_t i = s0 + tid; i < s1; i += kTopkBlock) {
    const int32_t c = sid[i];
    const int32_t L = (c >= 0 && c < nc) ? clen[c] : 0;
    if (L < 1 || L > GAT_MAX_TRAJ_LEN) continue;
    if (sdist[i] < T) {
      const int32_t pos = atomicAdd(&n_out, 1);
      if (pos < kTopkMax) {
        out_d[pos] = sdist[i];
        out_i[pos] = c;
      }
    }
  }
  // The r smallest ids at T: the least id X with r candidates at T at or
  // below it, by a binary search over ids.  Each length valued T holds its ids
  // ascending and at most ns of them are survivors (counted at their DP
  // distance), so a list of at least r + ns ids has r non-survivors at or
  // below its (r + ns)-th id: the least such id over those lists bounds X
  // from above, and below that bound every list's count of ids <= X lies in
  // its first r + ns ids.  The search then runs over that short range with
  // short binary searches per list.
  const int32_t pre = r + ns;
  int32_t bound = nc > 0 ? nc - 1 : 0;
  for (int32_t t = tid; t < nt && nt <= kTlMax; t += kTopkBlock) {
    const int L = tl[t];
    if (len_off[L + 1] - len_off[L] >= pre) bound = min(bound, len_ids[len_off[L] + pre - 1]);
  }
  if (nt > kTlMax)
    for (int L = 1 + tid; L < kLenBins; L += kTopkBlock)
      if (val[L] == T && len_off[L + 1] - len_off[L] >= pre) bound = min(bound, len_ids[len_off[L] + pre - 1]);
  for (int o = 32; o > 0; o >>= 1) bound = min(bound, __shfl_xor(bound, o));
  if ((tid & 63) == 0) red[tid >> 6] = bound;
  __syncthreads();
  bound = min(min(red[0], red[1]), min(red[2], red[3]));
  __syncthreads();  // every thread has read red before ties_le reuses it
  auto list_le = [&](int L, int32_t X) -> int32_t {
    const int32_t size = len_off[L + 1] - len_off[L];
    return count_le(len_ids + len_off[L], size < pre ? size : pre, X);
  };
  auto ties_le = [&](int32_t X) -> int32_t {
    int32_t part = 0;
    if (nt <= kTlMax) {
      for (int32_t t = tid; t < nt; t += kTopkBlock) part += list_le(tl[t], X);
    } else {
      for (int L = 1 + tid; L < kLenBins; L += kTopkBlock)
        if (val[L] == T) part += list_le(L, X);
    }
    for (int32_t i = s0 + tid; i < s1; i += kTopkBlock) {
      const int32_t c = sid[i];
      if (c > X || c < 0 || c >= nc) continue;
      const int32_t L = clen[c];
      if (L < 1 || L > GAT_MAX_TRAJ_LEN) continue;
      part += (sdist[i] == T ? 1 : 0) - (val[L] == T ? 1 : 0);
    }
    return topk_block_sum(part, red);
  };
  int32_t lo = 0, hi = bound;
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (ties_le(mid) >= r) hi = mid;
    else lo = mid + 1;
  }
  const int32_t X = lo;
  auto collect = [&](int L) {
    const int32_t* ids = len_ids + len_off[L];
    const int32_t u = list_le(L, X);
    for (int32_t i = 0; i < u; ++i) {
      const int32_t c = ids[i];
      if (is_survivor(c)) continue;  // counted at its DP distance
      const int32_t pos = atomicAdd(&n_out, 1);
      if (pos < kTopkMax) {
        out_d[pos] = T;
        out_i[pos] = c;
      }
    }
  };
  if (nt <= kTlMax) {
    for (int32_t t = tid; t < nt; t += kTopkBlock) collect(tl[t]);
  } else {
    for (int L = 1 + tid; L < kLenBins; L += kTopkBlock)
      if (val[L] == T) collect(L);
  }
  for (int32_t i = s0 + tid; i < s1; i += kTopkBlock) {
    const int32_t c = sid[i];
    if (c > X || c < 0 || c >= nc) continue;
    const int32_t L = clen[c];
    if (L < 1 || L > GAT_MAX_TRAJ_LEN) continue;
    if (sdist[i] == T) {
      const int32_t pos = atomicAdd(&n_out, 1);
      if (pos < kTopkMax) {
        out_d[pos] = T;
        out_i[pos] = c;
      }
    }
  }
  __syncthreads();
  const int32_t n = n_out;
  if (n != kk) {
    if (tid == 0) atomicOr(status, 2);
    return;
  }
  for (int i = tid; i < n; i += kTopkBlock) {
    const int32_t d = out_d[i], c = out_i[i];
    int32_t rank = 0;
    for (int j = 0; j < n; ++j) rank += (out_d[j] < d || (out_d[j] == d && out_i[j] < c)) ? 1 : 0;
    out_dist[int64_t(q) * k + rank] = d;
    out_ids[int64_t(q) * k + rank] = c + 1;
  }
}

// ---------------------------------------------------------------- device jobs
// The batch's job list built on the device from the filter's survivors, so a
// top-k batch runs filter -> jobs -> DP -> survivor lists -> merge with no host
// round trip.  One workgroup: per pair the shorter and longer length, the
// length bucket (runBatched's rules: short up to `split`, long up to `mid`,
// the rest one pair per wave, and in a batch of one round of waves the pairs
// that would outlast the longest pair's several-wave schedule on one wave
// take several), then a counting sort in LDS by (bucket, strips descending,
// longer length descending in sixteen steps of the bucket's length range) and
// jobs of up to four pairs of one strip count (one pair in the long buckets).  Pairs of equal sort key
// keep no particular order: it only changes which pairs share a job.
constexpr int kJobThreads = 1024;
constexpr int64_t kJobsOneBlock = 131072;  // larger batches take the whole device
constexpr int kJobsScratchExtra = GAT_EDR_JOBS_SCRATCH_EXTRA;
constexpr int kStripsMax = GAT_MAX_TRAJ_LEN / 16 + 1;  // strips_of(n) is 0..128
constexpr int kJobGroups = 4 * kStripsMax;             // (bucket, strips)
constexpr int kLenSteps = 16;                          // longer length in 128-point steps
constexpr int kSortBins = kJobGroups * kLenSteps;
constexpr int kSortPer = (kSortBins + kJobThreads - 1) / kJobThreads;

// Exclusive scan of a[0..n) in LDS by the whole workgroup (n <= per *
// kJobThreads); returns the total.  red holds kJobThreads / 64 + 1 words.
template <int PER>
__device__ uint32_t block_exclusive_scan(uint32_t* a, int n, uint32_t* red) {
  const int t = threadIdx.x;
  uint32_t v[PER], sum = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = t * PER + k;
    v[k] = i < n ? a[i] : 0u;
    sum += v[k];
  }
  uint32_t incl = sum;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = uint32_t(__shfl_up(int(incl), d, 64));
    if ((t & 63) >= d) incl += y;
  }
  if ((t & 63) == 63) red[t / 64] = incl;
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (int w = 0; w < kJobThreads / 64; ++w) {
      const uint32_t x = red[w];
      red[w] = run;
      run += x;
    }
    red[kJobThreads / 64] = run;
  }
  __syncthreads();
  uint32_t base = red[t / 64] + incl - sum;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = t * PER + k;
    if (i < n) a[i] = base;
    base += v[k];
  }
  const uint32_t total = red[kJobThreads / 64];
  __syncthreads();
  return total;
}

// Several-wave schedule of a pair of s 64-row strips over a row of b blocks
// (runBatched's pipe_blocks: a lag of kLag blocks between strips).
__device__ __forceinline__ int64_t pipe_blocks(int64_t s, int64_t b, int waves) {
  const int64_t P = b > int64_t(waves) * kLag ? b : int64_t(waves) * kLag;
  return ((s - 1) / waves) * P + ((s - 1) % waves) * kLag + b;
}

struct JobRules {
  int32_t split, waves, every_long, cus;
};

// The bucket rules of a batch of np pairs: in a batch of fewer than 64 jobs
// per CU (one round of waves) long means longer than `split`, otherwise
// longer than max(split, 1024); with several waves per pair such a batch
// sends some long pairs to the several-wave bucket.
struct JobShape {
  int mid;
  bool multi;
  __device__ JobShape(const JobRules& r, uint32_t np) {
    const bool latency_bound = (int64_t(np) + 3) / 4 < int64_t(r.cus) * 16;
    mid = latency_bound ? r.split : max(r.split, 1024);
    multi = latency_bound && r.waves > 1;
  }
};

// Sort key of a pair of shorter length n and longer m: (bucket, strips
// descending, longer length descending in sixteen steps of the bucket's
// range); critical is the batch's longest several-wave schedule.
__device__ __forceinline__ uint32_t job_key(int n, int m, const JobRules& rules, const JobShape& sh,
                                            int64_t critical) {
  int bucket = m <= rules.split ? 0 : m <= sh.mid ? 1 : 2;
  if (bucket == 2 && sh.multi && (rules.every_long || int64_t((n + 63) / 64) * ((m + 63 + 15) / 16) > critical))
    bucket = 3;
  const int strips = (n + 15) / 16;
  const int top = bucket == 0 ? rules.split : bucket == 1 ? sh.mid : GAT_MAX_TRAJ_LEN;
  const int step = max(1, (top + kLenSteps - 1) / kLenSteps);
  const int len = m >= 1 ? min((m - 1) / step, kLenSteps - 1) : 0;
  return uint32_t((bucket * kStripsMax + (kStripsMax - 1 - strips)) * kLenSteps + (kLenSteps - 1 - len));
}

__global__ __launch_bounds__(kJobThreads) void edr_jobs_kernel(const int2* __restrict__ pairs,
                                                               const unsigned long long* __restrict__ count,
                                                               uint32_t cap, const int32_t* __restrict__ qlen,
                                                               int32_t nq, const int32_t* __restrict__ clen,
                                                               int32_t nc, JobRules rules, int2* __restrict__ sorted,
                                                               int32_t* __restrict__ origin,
                                                               int2* __restrict__ nm, Job* __restrict__ jobs,
                                                               int32_t* __restrict__ windows,
                                                               int32_t* __restrict__ status,
                                                               unsigned long long* __restrict__ cells) {
  __shared__ uint32_t bins[kSortBins];
  __shared__ uint32_t gjobs[kJobGroups + 1];
  __shared__ uint32_t red[kJobThreads / 64 + 1];
  __shared__ unsigned long long red64[kJobThreads / 64];
  const int t = threadIdx.x;
  const unsigned long long total = *count;
  if (total > cap) {  // uniform: nothing to run; the caller grows the buffers
    if (t < 8) windows[t] = 0;
    if (t == 0) atomicOr(status, 1);
    return;
  }
  const uint32_t np = uint32_t(total);
  // Lengths, the cell count and the several-wave threshold.
  const JobShape sh0(rules, np);
  const int mid = sh0.mid;
  const bool multi = sh0.multi;
  unsigned long long my_cells = 0;
  int64_t my_crit = 0;
  for (uint32_t i = t; i < np; i += kJobThreads) {
    const int2 p = pairs[i];
    const bool ok = p.x >= 0 && p.x < nq && p.y >= 0 && p.y < nc;
    const int a = ok ? qlen[p.x] : 0, b = ok ? clen[p.y] : 0;
    const int n = min(a, b), m = max(a, b);
    nm[i] = make_int2(n, m);
    my_cells += (unsigned long long)(n) * (unsigned long long)(m);
    if (multi && m > mid) {
      const int64_t c = pipe_blocks((n + 63) / 64, (m + 63 + 15) / 16, rules.waves);
      if (c > my_crit) my_crit = c;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    my_cells += __shfl_xor(my_cells, off);
    const int64_t o = __shfl_xor(my_crit, off);
    if (o > my_crit) my_crit = o;
  }
  __shared__ int64_t crit_w[kJobThreads / 64];
  if ((t & 63) == 0) {
    red64[t / 64] = my_cells;
    crit_w[t / 64] = my_crit;
  }
  for (int i = t; i < kSortBins; i += kJobThreads) bins[i] = 0;
  __syncthreads();
  int64_t critical = 0;
  for (int w = 0; w < kJobThreads / 64; ++w) critical = crit_w[w] > critical ? crit_w[w] : critical;
  if (t == 0) {
    unsigned long long c = 0;
    for (int w = 0; w < kJobThreads / 64; ++w) c += red64[w];
    *cells = c;
  }
  const JobShape shape(rules, np);
  auto key_of = [&](int2 v) { return job_key(v.x, v.y, rules, shape, critical); };
  for (uint32_t i = t; i < np; i += kJobThreads) atomicAdd(&bins[key_of(nm[i])], 1u);
  __syncthreads();
  block_exclusive_scan<kSortPer>(bins, kSortBins, red);
  for (uint32_t i = t; i < np; i += kJobThreads) {
    const uint32_t pos = atomicAdd(&bins[key_of(nm[i])], 1u);
    sorted[pos] = pairs[i];
    if (origin) origin[pos] = int32_t(i);
  }
  __syncthreads();
  // bins[b] now ends bin b: group g spans [end of bin g*16 - 1, end of bin g*16 + 15).
  auto group_start = [&](int g) { return g == 0 ? 0u : bins[g * kLenSteps - 1]; };
  auto per_job = [](int g) { return g / kStripsMax >= 2 ? 1u : 4u; };
  for (int g = t; g < kJobGroups; g += kJobThreads) {
    const uint32_t len = bins[g * kLenSteps + kLenSteps - 1] - group_start(g);
    gjobs[g] = (len + per_job(g) - 1) / per_job(g);
  }
  __syncthreads();
  const uint32_t njobs = block_exclusive_scan<1>(gjobs, kJobGroups, red);
  if (t == 0) gjobs[kJobGroups] = njobs;
  __syncthreads();
  for (uint32_t j = t; j < njobs; j += kJobThreads) {
    int lo = 0, hi = kJobGroups - 1;  // the last group whose first job is at or below j
    while (lo < hi) {
      const int mid_g = (lo + hi + 1) / 2;
      if (gjobs[mid_g] <= j) lo = mid_g; else hi = mid_g - 1;
    }
    const uint32_t g = uint32_t(lo), k = j - gjobs[g], per = per_job(int(g));
    const uint32_t first = group_start(int(g)) + k * per;
    const uint32_t end = bins[g * kLenSteps + kLenSteps - 1];
    jobs[j] = Job{int32_t(first), int32_t(min(per, end - first))};
  }
  if (t < 4) {
    windows[2 * t] = int32_t(gjobs[t * kStripsMax]);
    windows[2 * t + 1] = int32_t(gjobs[(t + 1) * kStripsMax] - gjobs[t * kStripsMax]);
  }
}

// The same job list for large batches, over the whole device: lengths and
// the batch's cell count and several-wave threshold; global bin totals;
// their scan, the groups' job counts and the windows (one workgroup); the
// pairs to their bins (each workgroup counts its contiguous share in LDS,
// reserves its range of every bin with one atomic per bin, then places its
// pairs); the jobs.  Scratch after the capacity's (n, m): bin totals, bin
// cursors, group starts and group job offsets.
constexpr int kJobsBlock = 256;

__device__ __forceinline__ bool jobs_overflow(const unsigned long long* count, uint32_t cap, uint32_t* np) {
  const unsigned long long total = *count;
  *np = total > cap ? 0u : uint32_t(total);
  return total > cap;
}

__global__ __launch_bounds__(kJobsBlock) void edr_jobs_lengths_kernel(
    const int2* __restrict__ pairs, const unsigned long long* __restrict__ count, uint32_t cap,
    const int32_t* __restrict__ qlen, int32_t nq, const int32_t* __restrict__ clen, int32_t nc, JobRules rules,
    int2* __restrict__ nm, unsigned long long* __restrict__ cells, unsigned long long* __restrict__ crit) {
  uint32_t np;
  if (jobs_overflow(count, cap, &np)) return;
  const JobShape sh(rules, np);
  unsigned long long my_cells = 0, my_crit = 0;
  for (uint32_t i = blockIdx.x * kJobsBlock + threadIdx.x; i < np; i += gridDim.x * kJobsBlock) {
    const int2 p = pairs[i];
    const bool ok = p.x >= 0 && p.x < nq && p.y >= 0 && p.y < nc;
    const int a = ok ? qlen[p.x] : 0, b = ok ? clen[p.y] : 0;
    const int n = min(a, b), m = max(a, b);
    nm[i] = make_int2(n, m);
    my_cells += (unsigned long long)(n) * (unsigned long long)(m);
    if (sh.multi && m > sh.mid) {
      const unsigned long long c =
          (unsigned long long)pipe_blocks((n + 63) / 64, (m + 63 + 15) / 16, rules.waves);
      my_crit = c > my_crit ? c : my_crit;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    my_cells += __shfl_xor(my_cells, off);
    const unsigned long long o = __shfl_xor(my_crit, off);
    my_crit = o > my_crit ? o : my_crit;
  }
  if ((threadIdx.x & 63) == 0) {
    if (my_cells) atomicAdd(cells, my_cells);
    if (my_crit) atomicMax(crit, my_crit);
  }
}

__device__ __forceinline__ void jobs_share(uint32_t np, uint32_t* lo, uint32_t* hi) {
  const uint32_t per = (np + gridDim.x - 1) / gridDim.x;
  *lo = min(np, blockIdx.x * per);
  *hi = min(np, *lo + per);
}

__global__ __launch_bounds__(kJobsBlock) void edr_jobs_hist_kernel(const unsigned long long* __restrict__ count,
                                                                   uint32_t cap, JobRules rules,
                                                                   const int2* __restrict__ nm,
                                                                   const unsigned long long* __restrict__ crit,
                                                                   uint32_t* __restrict__ gbins) {
  __shared__ uint32_t h[kSortBins];
  uint32_t np;
  if (jobs_overflow(count, cap, &np)) return;
  const JobShape sh(rules, np);
  const int64_t critical = int64_t(*crit);
  for (int b = threadIdx.x; b < kSortBins; b += kJobsBlock) h[b] = 0;
  __syncthreads();
  uint32_t lo, hi;
  jobs_share(np, &lo, &hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kJobsBlock) {
    const int2 v = nm[i];
    atomicAdd(&h[job_key(v.x, v.y, rules, sh, critical)], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kSortBins; b += kJobsBlock)
    if (h[b]) atomicAdd(gbins + b, h[b]);
}

__global__ __launch_bounds__(kJobThreads) void edr_jobs_scan_kernel(const unsigned long long* __restrict__ count,
                                                                   uint32_t cap, const uint32_t* __restrict__ gbins,
                                                                   uint32_t* __restrict__ gcur,
                                                                   uint32_t* __restrict__ gstart,
                                                                   uint32_t* __restrict__ gjobs_out,
                                                                   int32_t* __restrict__ windows,
                                                                   int32_t* __restrict__ status) {
  __shared__ uint32_t bins[kSortBins];
  __shared__ uint32_t gjobs[kJobGroups + 1];
  __shared__ uint32_t red[kJobThreads / 64 + 1];
  const int t = threadIdx.x;
  uint32_t np;
  if (jobs_overflow(count, cap, &np)) {
    if (t < 8) windows[t] = 0;
    if (t == 0) {
      atomicOr(status, 1);
      gjobs_out[kJobGroups] = 0;
    }
    return;
  }
  for (int b = t; b < kSortBins; b += kJobThreads) bins[b] = gbins[b];
  __syncthreads();
  block_exclusive_scan<kSortPer>(bins, kSortBins, red);
  for (int b = t; b < kSortBins; b += kJobThreads) gcur[b] = bins[b];
  auto per_job = [](int g) { return g / kStripsMax >= 2 ? 1u : 4u; };
  auto group_end = [&](int g) { return g + 1 < kJobGroups ? bins[(g + 1) * kLenSteps] : np; };
  for (int g = t; g < kJobGroups; g += kJobThreads) {
    const uint32_t first = bins[g * kLenSteps];
    gstart[g] = first;
    gjobs[g] = (group_end(g) - first + per_job(g) - 1) / per_job(g);
  }
  if (t == 0) gstart[kJobGroups] = np;
  __syncthreads();
  const uint32_t njobs = block_exclusive_scan<1>(gjobs, kJobGroups, red);
  for (int g = t; g < kJobGroups; g += kJobThreads) gjobs_out[g] = gjobs[g];
  if (t == 0) gjobs_out[kJobGroups] = njobs;
  if (t < 4) {
    const uint32_t a = gjobs[t * kStripsMax];
    const uint32_t b = t == 3 ? njobs : gjobs[(t + 1) * kStripsMax];
    windows[2 * t] = int32_t(a);
    windows[2 * t + 1] = int32_t(b - a);
  }
}

__global__ __launch_bounds__(kJobsBlock) void edr_jobs_scatter_kernel(
    const int2* __restrict__ pairs, const unsigned long long* __restrict__ count, uint32_t cap, JobRules rules,
    const int2* __restrict__ nm, const unsigned long long* __restrict__ crit, uint32_t* __restrict__ gcur,
    int2* __restrict__ sorted, int32_t* __restrict__ origin) {
  __shared__ uint32_t h[kSortBins];
  uint32_t np;
  if (jobs_overflow(count, cap, &np)) return;
  const JobShape sh(rules, np);
  const int64_t critical = int64_t(*crit);
  for (int b = threadIdx.x; b < kSortBins; b += kJobsBlock) h[b] = 0;
  __syncthreads();
  uint32_t lo, hi;
  jobs_share(np, &lo, &hi);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kJobsBlock) {
    const int2 v = nm[i];
    atomicAdd(&h[job_key(v.x, v.y, rules, sh, critical)], 1u);
  }
  __syncthreads();
  // This workgroup's range of every bin it uses; h becomes its cursors.
  for (int b = threadIdx.x; b < kSortBins; b += kJobsBlock)
    if (h[b]) h[b] = atomicAdd(gcur + b, h[b]);
  __syncthreads();
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kJobsBlock) {
    const int2 v = nm[i];
    const uint32_t pos = atomicAdd(&h[job_key(v.x, v.y, rules, sh, critical)], 1u);
    if (pos < np) {
      sorted[pos] = pairs[i];
      if (origin) origin[pos] = int32_t(i);
    }
  }
}

__global__ __launch_bounds__(kJobsBlock) void edr_jobs_write_kernel(const unsigned long long* __restrict__ count,
                                                                    uint32_t cap, const uint32_t* __restrict__ gstart,
                                                                    const uint32_t* __restrict__ gjobs_in,
                                                                    Job* __restrict__ jobs) {
  __shared__ uint32_t gs[kJobGroups + 1], gj[kJobGroups + 1];
  uint32_t np;
  if (jobs_overflow(count, cap, &np)) return;
  for (int g = threadIdx.x; g <= kJobGroups; g += kJobsBlock) {
    gs[g] = gstart[g];
    gj[g] = gjobs_in[g];
  }
  __syncthreads();
  const uint32_t njobs = gj[kJobGroups];
  for (uint32_t j = blockIdx.x * kJobsBlock + threadIdx.x; j < njobs; j += gridDim.x * kJobsBlock) {
    int lo = 0, hi = kJobGroups - 1;  // the last group whose first job is at or below j
    while (lo < hi) {
      const int mid_g = (lo + hi + 1) / 2;
      if (gj[mid_g] <= j) lo = mid_g; else hi = mid_g - 1;
    }
    const uint32_t per = lo / kStripsMax >= 2 ? 1u : 4u;
    const uint32_t first = gs[lo] + (j - gj[lo]) * per;
    jobs[j] = Job{int32_t(first), int32_t(min(per, gs[lo + 1] - first))};
  }
}

// Per-query survivor lists for edr_topk_kernel from the DP's sorted pairs and
// distances: counts in LDS, their scan as the start offsets, then each pair
// to its query's next slot (in no order; edr_lists_sort_kernel sorts by id).
constexpr int kListQueries = 8192;

__global__ __launch_bounds__(kJobThreads) void edr_lists_kernel(const int2* __restrict__ sorted,
                                                                const int32_t* __restrict__ dist,
                                                                const unsigned long long* __restrict__ count,
                                                                uint32_t cap, int32_t nq, int32_t* __restrict__ start,
                                                                int32_t* __restrict__ ids,
                                                                int32_t* __restrict__ dists) {
  __shared__ uint32_t cnt[kListQueries];
  __shared__ uint32_t red[kJobThreads / 64 + 1];
  const int t = threadIdx.x;
  const unsigned long long total = *count;
  const uint32_t np = total > cap ? 0u : uint32_t(total);
  for (int q = t; q < nq; q += kJobThreads) cnt[q] = 0;
  __syncthreads();
  for (uint32_t i = t; i < np; i += kJobThreads) {
    const int q = sorted[i].x;
    if (q >= 0 && q < nq) atomicAdd(&cnt[q], 1u);
  }
  __syncthreads();
  const uint32_t n = block_exclusive_scan<kListQueries / kJobThreads>(cnt, nq, red);
  for (int q = t; q < nq; q += kJobThreads) start[q] = int32_t(cnt[q]);
  if (t == 0) start[nq] = int32_t(n);
  __syncthreads();
  for (uint32_t i = t; i < np; i += kJobThreads) {
    const int2 p = sorted[i];
    if (p.x < 0 || p.x >= nq) continue;
    const uint32_t pos = atomicAdd(&cnt[p.x], 1u);
    ids[pos] = p.y;
    dists[pos] = dist[i];
  }
}

// Each query's survivors sorted by id (ids are distinct within a query): a
// bitonic sort in LDS of up to kSortMax (id, distance) pairs per workgroup;
// a query with more sets status bit 2 and keeps its list unsorted.
constexpr int kSortMax = 2048;
constexpr int kSortThreads = 256;

__global__ __launch_bounds__(kSortThreads) void edr_lists_sort_kernel(const int32_t* __restrict__ start,
                                                                      int32_t* __restrict__ ids,
                                                                      int32_t* __restrict__ dists,
                                                                      int32_t* __restrict__ status) {
  __shared__ unsigned long long v[kSortMax];
  const int q = blockIdx.x, t = threadIdx.x;
  const int s0 = start[q], m = start[q + 1] - s0;
  if (m <= 1) return;
  if (m > kSortMax) {
    if (t == 0) atomicOr(status, 2);
    return;
  }
  int P = 2;
  while (P < m) P <<= 1;
  for (int i = t; i < P; i += kSortThreads)
    v[i] = i < m ? ((unsigned long long)(uint32_t(ids[s0 + i])) << 32) | uint32_t(dists[s0 + i]) : ~0ull;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += kSortThreads) {
        const int o = i ^ j;
        if (o > i) {
          const bool up = (i & k) == 0;
          const unsigned long long x = v[i], y = v[o];
          if ((x > y) == up) {
            v[i] = y;
            v[o] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int i = t; i < m; i += kSortThreads) {
    ids[s0 + i] = int32_t(v[i] >> 32);
    dists[s0 + i] = int32_t(uint32_t(v[i]));
  }
}

// Every (query, candidate) pair of nq x nc, pair i = (i / nc, i % nc).
__global__ __launch_bounds__(256) void edr_all_pairs_kernel(int32_t nq, int32_t nc, int2* __restrict__ pairs,
                                                            unsigned long long* __restrict__ count) {
  const uint64_t n = uint64_t(nq) * uint64_t(nc);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    pairs[i] = make_int2(int32_t(i / uint64_t(nc)), int32_t(i % uint64_t(nc)));
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

// out[origin[p]] = sorted_dist[p]: distances back in the pairs' first order.
__global__ __launch_bounds__(256) void edr_unsort_kernel(const int32_t* __restrict__ origin,
                                                         const int32_t* __restrict__ sorted_dist,
                                                         const unsigned long long* __restrict__ count, uint32_t cap,
                                                         int32_t* __restrict__ out) {
  const unsigned long long total = *count;
  const uint32_t n = total > cap ? 0u : uint32_t(total);
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int32_t o = origin[p];
    if (o >= 0 && uint32_t(o) < n) out[o] = sorted_dist[p];
  }
}

}  // namespace

extern "C" int gat_edr_filter(const float* qbox, int32_t nq, const float* cbox, int32_t nc, double eps,
                              int32_t* pairs_out, int64_t capacity, uint64_t* count, void* stream) {
  GAT_REQUIRE(nq >= 0 && nc >= 0 && capacity >= 0 && nq <= 65535, "EDR filter: bad sizes");
  GAT_REQUIRE(count, "EDR filter: null count");
  hipStream_t s = gat::as_stream(stream);
  GAT_HIP_CHECK(hipMemsetAsync(count, 0, sizeof(uint64_t), s));
  if (nq == 0 || nc == 0) return GAT_OK;
  GAT_REQUIRE(qbox && cbox && (pairs_out || capacity == 0), "EDR filter: null pointer");
  GAT_REQUIRE(eps > 0, "EDR filter: eps must be positive");
  int64_t bx = (int64_t(nc) + 255) / 256;
  if (bx > 64) bx = 64;
  edr_filter_kernel<<<dim3(uint32_t(bx), uint32_t((nq + kFilterQ - 1) / kFilterQ)), dim3(256), 0, s>>>(
      reinterpret_cast<const float4*>(qbox), nq, reinterpret_cast<const float4*>(cbox), nc, gat_eps_threshold(eps),
      reinterpret_cast<int2*>(pairs_out), uint64_t(capacity), reinterpret_cast<unsigned long long*>(count));
  GAT_LAUNCH_CHECK("edr_filter_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_filter_grid(const float* qbox, int32_t nq, const float* cbox, int32_t nc,
                                   const gat_edr_cand_grid* grid, double eps, int32_t* pairs_out, int64_t capacity,
                                   uint64_t* count, void* stream) {
  GAT_REQUIRE(nq >= 0 && nc >= 0 && capacity >= 0, "EDR grid filter: bad sizes");
  GAT_REQUIRE(count && grid, "EDR grid filter: null count or grid");
  GAT_REQUIRE(grid->cells_per_side >= 1 && grid->cells_per_side <= 4096 && grid->nalways >= 0 &&
                  grid->cell_start_dev && (grid->nalways == 0 || grid->always_dev) && std::isfinite(grid->x0) &&
                  std::isfinite(grid->y0) && std::isfinite(grid->inv_x) && std::isfinite(grid->inv_y) &&
                  grid->inv_x >= 0.0f && grid->inv_y >= 0.0f,
              "EDR grid filter: bad grid");
  hipStream_t s = gat::as_stream(stream);
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(count), 0u, 2, stream)) return rc;
  if (nq == 0 || nc == 0) return GAT_OK;
  GAT_REQUIRE(qbox && cbox && (pairs_out || capacity == 0), "EDR grid filter: null pointer");
  GAT_REQUIRE(eps > 0, "EDR grid filter: eps must be positive");
  const CandGrid gr{grid->cells_per_side,
                    grid->x0,
                    grid->y0,
                    grid->inv_x,
                    grid->inv_y,
                    grid->cell_start_dev,
                    grid->entry_id_dev,
                    reinterpret_cast<const float4*>(grid->entry_box_dev),
                    grid->always_dev,
                    grid->nalways};
  edr_filter_grid_kernel<<<dim3(uint32_t((nq + kGridFilterWaves - 1) / kGridFilterWaves)),
                           dim3(64 * kGridFilterWaves), 0, s>>>(
      reinterpret_cast<const float4*>(qbox), nq, reinterpret_cast<const float4*>(cbox), nc, gr,
      gat_eps_threshold(eps), reinterpret_cast<int2*>(pairs_out), uint64_t(capacity),
      reinterpret_cast<unsigned long long*>(count));
  GAT_LAUNCH_CHECK("edr_filter_grid_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_topk(const int32_t* ftable, int32_t w, const int32_t* qlen, int32_t nq, const int32_t* clen,
                            int32_t nc, const int32_t* len_ids, const int32_t* len_off, const int32_t* surv_start,
                            const int32_t* surv_id, const int32_t* surv_dist, int32_t k, int32_t kk, int32_t* out_ids,
                            int32_t* out_dist, int32_t* status, void* stream) {
  GAT_REQUIRE(w == GAT_MAX_TRAJ_LEN + 1, "EDR top-k: the mismatch table must be %d wide", GAT_MAX_TRAJ_LEN + 1);
  GAT_REQUIRE(nq >= 0 && nc >= 0, "EDR top-k: bad sizes");
  GAT_REQUIRE(k >= 1 && k <= kTopkMax && kk >= 0 && kk <= k && kk <= nc,
              "EDR top-k: k outside 1..%d or kk outside 0..min(k, candidates)", kTopkMax);
  GAT_REQUIRE(status, "EDR top-k: null status");
  hipStream_t s = gat::as_stream(stream);
  GAT_HIP_CHECK(hipMemsetAsync(status, 0, sizeof(int32_t), s));
  if (nq == 0 || kk == 0) return GAT_OK;
  GAT_REQUIRE(ftable && qlen && clen && len_ids && len_off && surv_start && out_ids && out_dist,
              "EDR top-k: null pointer");
  edr_topk_kernel<<<dim3(uint32_t(nq)), dim3(kTopkBlock), 0, s>>>(ftable, w, qlen, clen, nc, len_ids, len_off,
                                                                  surv_start, surv_id, surv_dist, k, kk, out_ids,
                                                                  out_dist, status);
  GAT_LAUNCH_CHECK("edr_topk_kernel");
  return GAT_OK;
}

namespace {

// One wave per trajectory: its points, given as separate lon / lat arrays, go
// to the padded interleaved store at dst_off[t], and its box is reduced over
// the wave.  The store is NaN-filled first by the entry point.
__global__ __launch_bounds__(64) void edr_pack_kernel(const float* __restrict__ lon, const float* __restrict__ lat,
                                                      const int64_t* __restrict__ src_off,
                                                      const int64_t* __restrict__ dst_off, int32_t ntraj,
                                                      int64_t total, float2* __restrict__ xy,
                                                      float4* __restrict__ box) {
  for (int32_t t = int32_t(blockIdx.x); t < ntraj; t += int32_t(gridDim.x)) {
    const int64_t b = src_off[t], e = src_off[t + 1], d = dst_off[t];
    float xl = INFINITY, yl = INFINITY, xh = -INFINITY, yh = -INFINITY;
    for (int64_t i = b + int64_t(threadIdx.x); i < e; i += 64) {
      const float x = lon[i], y = lat[i];
      const int64_t at = d + (i - b);
      if (at >= 0 && at < total) xy[at] = make_float2(x, y);
      xl = fminf(xl, x);
      xh = fmaxf(xh, x);
      yl = fminf(yl, y);
      yh = fmaxf(yh, y);
    }
    for (int o = 32; o > 0; o >>= 1) {
      xl = fminf(xl, __shfl_xor(xl, o));
      xh = fmaxf(xh, __shfl_xor(xh, o));
      yl = fminf(yl, __shfl_xor(yl, o));
      yh = fmaxf(yh, __shfl_xor(yh, o));
    }
    if (threadIdx.x == 0) box[t] = make_float4(xl, yl, xh, yh);
  }
}

}  // namespace

extern "C" int gat_edr_pack_trajectories(const float* lon, const float* lat, const int64_t* src_off,
                                         const int64_t* dst_off, int32_t ntraj, int64_t total, float* xy, float* box,
                                         void* stream) {
  GAT_REQUIRE(ntraj >= 0 && total >= 0, "gat_edr_pack_trajectories: bad sizes");
  GAT_REQUIRE(total == 0 || xy, "gat_edr_pack_trajectories: null store");
  if (total > 0)
    if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(xy), 0x7FC00000u, 2 * total, stream)) return rc;
  if (ntraj == 0) return GAT_OK;
  GAT_REQUIRE(lon && lat && src_off && dst_off && box, "gat_edr_pack_trajectories: null pointer");
  const int64_t cap = int64_t(gat::cu_count()) * 16;
  const uint32_t blocks = uint32_t(ntraj < cap ? ntraj : cap);
  edr_pack_kernel<<<dim3(blocks), dim3(64), 0, gat::as_stream(stream)>>>(
      lon, lat, src_off, dst_off, ntraj, total, reinterpret_cast<float2*>(xy), reinterpret_cast<float4*>(box));
  GAT_LAUNCH_CHECK("edr_pack_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_mismatch_table(int32_t maxlen, int32_t* out, void* stream) {
  GAT_REQUIRE(maxlen >= 0 && maxlen <= GAT_MAX_TRAJ_LEN, "mismatch table: maxlen outside 0..%d", GAT_MAX_TRAJ_LEN);
  GAT_REQUIRE(out, "mismatch table: null output");
  edr_mismatch_table_kernel<<<dim3(1), dim3(1024), 0, gat::as_stream(stream)>>>(maxlen, out);
  GAT_LAUNCH_CHECK("edr_mismatch_table_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_batched(const int32_t* pairs, int64_t npairs, const gat_edr_job* jobs, int64_t njobs,
                               const int32_t* window, int32_t* counter, const float* qxy, const int64_t* qoff, const int32_t* qlen,
                               int32_t nqueries, const float* cxy, const int64_t* coff, const int32_t* clen,
                               int32_t ncands, int32_t maxlen, double eps, int32_t per_cu, int32_t lanes_per_pair,
                               int32_t* out, void* stream) {
  GAT_REQUIRE(npairs >= 0 && njobs >= 0 && njobs <= INT32_MAX && nqueries >= 0 && ncands >= 0,
              "EDR batched: bad sizes");
  GAT_REQUIRE(lanes_per_pair == 16 || lanes_per_pair == 64 || lanes_per_pair == 128 || lanes_per_pair == 256 ||
                  lanes_per_pair == 512,
              "EDR batched: lanes_per_pair must be 16, 64, 128, 256 or 512");
  if (njobs == 0) return GAT_OK;
  GAT_REQUIRE(pairs && jobs && counter && qxy && qoff && qlen && cxy && coff && clen && out,
              "EDR batched: null pointer");
  GAT_REQUIRE(maxlen >= 1 && maxlen <= GAT_MAX_TRAJ_LEN, "EDR batched: maxlen outside 1..%d", GAT_MAX_TRAJ_LEN);
  GAT_REQUIRE(eps > 0, "EDR batched: eps must be positive");
  GAT_REQUIRE(per_cu >= 0 && per_cu <= 64, "EDR batched: per_cu outside 0..64");
  hipStream_t s = gat::as_stream(stream);
  // The job counter is cleared by a kernel, so the launch can be captured in a graph.
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(counter), 0u, 1, stream)) return rc;
  const float u = gat_eps_threshold(eps);
  const auto* P = reinterpret_cast<const int2*>(pairs);
  const auto* J = reinterpret_cast<const Job*>(jobs);
  const auto* Q = reinterpret_cast<const float2*>(qxy);
  const auto* Cs = reinterpret_cast<const float2*>(cxy);
  // LDS class: the smallest row buffer that holds maxlen columns (64-lane
  // pairs: 1024 or 2048).  Persistent grid: per_cu single-wave workgroups per
  // CU (default by class), capped by the work.
  const bool wide = lanes_per_pair >= 64;
  const int waves = lanes_per_pair / 64;  // > 1: one pair over several waves (edr_pipe_kernel)
  int cls = maxlen <= 256 ? 256 : maxlen <= 512 ? 512 : maxlen <= 1024 ? 1024 : GAT_MAX_TRAJ_LEN;
  if (wide && cls < 1024) cls = 1024;
  const int64_t pc = per_cu > 0 ? per_cu : waves > 1 ? 16 / waves : wide ? 16 : (cls <= 512 ? 20 : cls == 1024 ? 16 : 9);
  int64_t blocks = int64_t(gat::cu_count()) * pc;
  if (blocks > njobs) blocks = njobs;
  const dim3 grid{uint32_t(blocks)}, block{64};
  if (waves > 1) {
    const dim3 wblock{uint32_t(64 * waves)};
#define GAT_EDR_PIPE(M, W)                                                                                       \
  edr_pipe_kernel<M, W><<<grid, wblock, 0, s>>>(P, npairs, J, int32_t(njobs), window, counter, nqueries, ncands, Q, \
                                                qoff, qlen, Cs, coff, clen, u, out)
    if (cls == 1024) {
      if (waves == 2) GAT_EDR_PIPE(1024, 2);
      else if (waves == 4) GAT_EDR_PIPE(1024, 4);
      else GAT_EDR_PIPE(1024, 8);
    } else {
      if (waves == 2) GAT_EDR_PIPE(GAT_MAX_TRAJ_LEN, 2);
      else if (waves == 4) GAT_EDR_PIPE(GAT_MAX_TRAJ_LEN, 4);
      else GAT_EDR_PIPE(GAT_MAX_TRAJ_LEN, 8);
    }
#undef GAT_EDR_PIPE
    GAT_LAUNCH_CHECK("edr_pipe_kernel");
    return GAT_OK;
  }
#define GAT_EDR_LAUNCH(M, SEG)                                                                                  \
  edr_wave_kernel<M, SEG><<<grid, block, 0, s>>>(P, npairs, J, int32_t(njobs), window, counter, nqueries, ncands, Q, \
                                                 qoff, qlen, Cs, coff, clen, u, out)
  if (wide) {
    if (cls == 1024)
      GAT_EDR_LAUNCH(1024, 64);
    else
      GAT_EDR_LAUNCH(GAT_MAX_TRAJ_LEN, 64);
  } else {
    switch (cls) {
      case 256: GAT_EDR_LAUNCH(256, 16); break;
      case 512: GAT_EDR_LAUNCH(512, 16); break;
      case 1024: GAT_EDR_LAUNCH(1024, 16); break;
      default: GAT_EDR_LAUNCH(GAT_MAX_TRAJ_LEN, 16);
    }
  }
#undef GAT_EDR_LAUNCH
  GAT_LAUNCH_CHECK("edr_wave_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_jobs(const int32_t* pairs, const uint64_t* count, int64_t capacity, const int32_t* qlen,
                            int32_t nq, const int32_t* clen, int32_t nc, int32_t split, int32_t waves,
                            int32_t every_long, int32_t* sorted, int32_t* origin, int32_t* scratch,
                            gat_edr_job* jobs, int32_t* windows, int32_t* status, uint64_t* cells, void* stream) {
  GAT_REQUIRE(capacity >= 0 && capacity <= int64_t(INT32_MAX) - 8 && nq >= 0 && nc >= 0,
              "EDR jobs: bad sizes");
  GAT_REQUIRE(split >= 1 && split <= GAT_MAX_TRAJ_LEN && waves >= 1 && waves <= 8, "EDR jobs: bad rules");
  GAT_REQUIRE(pairs && count && qlen && clen && sorted && scratch && jobs && windows && status && cells,
              "EDR jobs: null pointer");
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(status), 0u, 1, stream)) return rc;
  const JobRules rules{split, waves, every_long ? 1 : 0, int32_t(gat::cu_count())};
  hipStream_t s = gat::as_stream(stream);
  const auto* P = reinterpret_cast<const int2*>(pairs);
  const auto* C = reinterpret_cast<const unsigned long long*>(count);
  auto* NM = reinterpret_cast<int2*>(scratch);
  auto* cells_u = reinterpret_cast<unsigned long long*>(cells);
  if (capacity <= kJobsOneBlock) {
    edr_jobs_kernel<<<dim3(1), dim3(kJobThreads), 0, s>>>(P, C, uint32_t(capacity), qlen, nq, clen, nc, rules,
                                                          reinterpret_cast<int2*>(sorted), origin, NM,
                                                          reinterpret_cast<Job*>(jobs), windows, status, cells_u);
    GAT_LAUNCH_CHECK("edr_jobs_kernel");
    return GAT_OK;
  }
  // Scratch after the (n, m) pairs (8-byte aligned: 2 * capacity words): the
  // several-wave threshold, bin totals, bin cursors, group starts, group job
  // offsets.
  static_assert(2 + 2 * kSortBins + 2 * (kJobGroups + 1) <= kJobsScratchExtra, "job scratch");
  auto* crit = reinterpret_cast<unsigned long long*>(scratch + 2 * capacity);
  uint32_t* gbins = reinterpret_cast<uint32_t*>(scratch + 2 * capacity + 2);
  uint32_t* gcur = gbins + kSortBins;
  uint32_t* gstart = gcur + kSortBins;
  uint32_t* gjobs = gstart + kJobGroups + 1;
  if (int rc = gat_fill_u32(gbins, 0u, kSortBins, stream)) return rc;
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(cells), 0u, 2, stream)) return rc;
  if (int rc = gat_fill_u32(reinterpret_cast<uint32_t*>(crit), 0u, 2, stream)) return rc;
  int64_t blocks = (capacity + 4095) / 4096;
  const int64_t most = int64_t(gat::cu_count()) * 4;
  if (blocks > most) blocks = most;
  if (blocks < 1) blocks = 1;
  const dim3 grid{uint32_t(blocks)}, block{kJobsBlock};
  const uint32_t cap = uint32_t(capacity);
  edr_jobs_lengths_kernel<<<grid, block, 0, s>>>(P, C, cap, qlen, nq, clen, nc, rules, NM, cells_u, crit);
  GAT_LAUNCH_CHECK("edr_jobs_lengths_kernel");
  edr_jobs_hist_kernel<<<grid, block, 0, s>>>(C, cap, rules, NM, crit, gbins);
  GAT_LAUNCH_CHECK("edr_jobs_hist_kernel");
  edr_jobs_scan_kernel<<<dim3(1), dim3(kJobThreads), 0, s>>>(C, cap, gbins, gcur, gstart, gjobs, windows, status);
  GAT_LAUNCH_CHECK("edr_jobs_scan_kernel");
  edr_jobs_scatter_kernel<<<grid, block, 0, s>>>(P, C, cap, rules, NM, crit, gcur, reinterpret_cast<int2*>(sorted),
                                                 origin);
  GAT_LAUNCH_CHECK("edr_jobs_scatter_kernel");
  edr_jobs_write_kernel<<<grid, block, 0, s>>>(C, cap, gstart, gjobs, reinterpret_cast<Job*>(jobs));
  GAT_LAUNCH_CHECK("edr_jobs_write_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_survivor_lists(const int32_t* sorted, const int32_t* dist, const uint64_t* count,
                                      int64_t capacity, int32_t nq, int32_t* start, int32_t* ids, int32_t* dists,
                                      int32_t* status, void* stream) {
  GAT_REQUIRE(capacity >= 0 && capacity <= int64_t(INT32_MAX) && nq >= 0 && nq <= kListQueries,
              "EDR survivor lists: at most %d queries", kListQueries);
  GAT_REQUIRE(sorted && dist && count && start && ids && dists && status, "EDR survivor lists: null pointer");
  if (nq == 0) return GAT_OK;
  hipStream_t s = gat::as_stream(stream);
  edr_lists_kernel<<<dim3(1), dim3(kJobThreads), 0, s>>>(reinterpret_cast<const int2*>(sorted), dist,
                                                          reinterpret_cast<const unsigned long long*>(count),
                                                          uint32_t(capacity), nq, start, ids, dists);
  GAT_LAUNCH_CHECK("edr_lists_kernel");
  edr_lists_sort_kernel<<<dim3(uint32_t(nq)), dim3(kSortThreads), 0, s>>>(start, ids, dists, status);
  GAT_LAUNCH_CHECK("edr_lists_sort_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_all_pairs(int32_t nq, int32_t nc, int32_t* pairs, uint64_t* count, void* stream) {
  GAT_REQUIRE(nq >= 0 && nc >= 0 && int64_t(nq) * nc <= int64_t(INT32_MAX), "EDR all pairs: bad sizes");
  GAT_REQUIRE(pairs && count, "EDR all pairs: null pointer");
  const int64_t n = int64_t(nq) * nc;
  int64_t blocks = (n + 255) / 256;
  const int64_t cap = int64_t(gat::cu_count()) * 16;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  edr_all_pairs_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, gat::as_stream(stream)>>>(
      nq, nc, reinterpret_cast<int2*>(pairs), reinterpret_cast<unsigned long long*>(count));
  GAT_LAUNCH_CHECK("edr_all_pairs_kernel");
  return GAT_OK;
}

extern "C" int gat_edr_unsort(const int32_t* origin, const int32_t* sorted_dist, const uint64_t* count,
                              int64_t capacity, int32_t* out, void* stream) {
  GAT_REQUIRE(capacity >= 0 && capacity <= int64_t(INT32_MAX), "EDR unsort: bad capacity");
  GAT_REQUIRE(origin && sorted_dist && count && out, "EDR unsort: null pointer");
  int64_t blocks = (capacity + 255) / 256;
  const int64_t cap = int64_t(gat::cu_count()) * 16;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  edr_unsort_kernel<<<dim3(uint32_t(blocks)), dim3(256), 0, gat::as_stream(stream)>>>(
      origin, sorted_dist, reinterpret_cast<const unsigned long long*>(count), uint32_t(capacity), out);
  GAT_LAUNCH_CHECK("edr_unsort_kernel");
  return GAT_OK;
}

