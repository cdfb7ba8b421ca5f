// SystemTest — the reference's query driver (SystemTest.h:13-35): the range
// tests of the three indexes and the similarity test, run on the GPU path and
// timed.  The reference replicates one rectangle / one query trajectory N times
// (SystemTest.cpp:31-33, 167-189); the overloads taking arrays run distinct
// inputs through the same classes.  Times are wall-clock milliseconds
// (std::chrono::steady_clock; the reference's MyTimer reports seconds under an
// "ms" label, ConstDefine.h:144-147).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "FSG.h"
#include "Grid.h"
#include "MortonGrid.h"
#include "STIG.h"
#include "Trajectory.h"

class SystemTest {
 public:
  SystemTest() = default;
  SystemTest(Trajectory* tradb, Grid* g, STIG* stig, FSG* fsg) : tradb(tradb), g(g), stig(stig), fsg(fsg) {}
  // SystemTest.cpp:15-22.
  SystemTest(Trajectory* tradb, Grid* g, STIG* stig, FSG* fsg, MortonGrid* mgrid)
      : tradb(tradb), g(g), stig(stig), fsg(fsg), mgrid(mgrid) {}

  int rangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum);
  int rangeQueryTest(const MBB* bounds, int rangeQueryNum);
  int STIGrangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum);
  int STIGrangeQueryTest(const MBB* bounds, int rangeQueryNum);
  int FSGrangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum);
  int FSGrangeQueryTest(const MBB* bounds, int rangeQueryNum);
  // SystemTest.cpp:386-445 (the reference issues 50 rectangles per call).
  int MortonGridRangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum);
  int MortonGridRangeQueryTest(const MBB* bounds, int rangeQueryNum);
  // SystemTest.cpp:164-249: the query replicated similarityScale times, the
  // CPU top-k timed first (SimilarityQueryBatchCPUParallel, SystemTest.cpp:207),
  // then the GPU top-k (SimilarityQueryBatchOnGPU, SystemTest.cpp:226).  This
  // library keeps no CPU query path: the CPU leg runs the caller's cpuSimilarity
  // when one is set and is skipped otherwise.  The ids of both legs stay in
  // cpuIds / gpuIds (query-major, k per query).
  int similarityQueryTest(Trajectory t, int similarityScale, int similarityKValue);

  // The CPU top-k of the similarity test: (queries, query count, k, ids out).
  using CpuSimilarity = std::function<int(Trajectory*, int, int, int*)>;
  CpuSimilarity cpuSimilarity;
  std::vector<int> cpuIds, gpuIds;
  double cpuMs = 0;

  struct Result {
    double wallMs = 0;      // whole call: host candidate selection, upload, kernel, count readback
    double selectMs = 0;    // host candidate selection
    float kernelMs = 0;     // refine kernel (device events)
    int64_t candidates = 0;
    int64_t pairs = 0;
  };
  Result last;
  // The last range test's result pairs (per-query counts, digests and pairs
  // stay readable on the device).
  RangeBatch lastBatch;
  bool verbose = true;

  Trajectory* tradb = nullptr;
  Grid* g = nullptr;
  STIG* stig = nullptr;
  FSG* fsg = nullptr;
  MortonGrid* mgrid = nullptr;

 private:
  int report(const char* what, int n, const RangeBatch& b, double ms);
};
