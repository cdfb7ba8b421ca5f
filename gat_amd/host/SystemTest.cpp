#include "SystemTest.h"

#include <chrono>
#include <cstdio>

namespace {
double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

int SystemTest::report(const char* what, int n, const RangeBatch& b, double ms) {
  last.wallMs = ms;
  last.selectMs = b.selectMs;
  last.kernelMs = b.kernelMs;
  last.candidates = b.candidates;
  last.pairs = b.pairTotal;
  if (verbose)
    std::printf("single GPU %s range query #query=%d: %.3f ms (select %.3f ms, refine kernel %.3f ms, %lld pairs)\n",
                what, n, ms, b.selectMs, double(b.kernelMs), (long long)b.pairTotal);
  return GAT_OK;
}

int SystemTest::rangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum) {
  std::vector<MBB> bounds(size_t(std::max(rangeQueryNum, 0)), rangeQueryMBB);
  return rangeQueryTest(bounds.data(), rangeQueryNum);
}

int SystemTest::rangeQueryTest(const MBB* bounds, int rangeQueryNum) {
  if (!g) return GAT_ERR_STATE;
  lastBatch = RangeBatch{};
  RangeBatch& b = lastBatch;
  const double t0 = now_ms();
  if (int rc = g->rangeQueryBatchGPUPairs(bounds, rangeQueryNum, b)) return rc;
  return report("GT-quadtree", rangeQueryNum, b, now_ms() - t0);
}

int SystemTest::STIGrangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum) {
  std::vector<MBB> bounds(size_t(std::max(rangeQueryNum, 0)), rangeQueryMBB);
  return STIGrangeQueryTest(bounds.data(), rangeQueryNum);
}

int SystemTest::STIGrangeQueryTest(const MBB* bounds, int rangeQueryNum) {
  if (!stig) return GAT_ERR_STATE;
  lastBatch = RangeBatch{};
  RangeBatch& b = lastBatch;
  const double t0 = now_ms();
  if (int rc = stig->rangeQueryGPUPairs(bounds, rangeQueryNum, b)) return rc;
  return report("STIG", rangeQueryNum, b, now_ms() - t0);
}

int SystemTest::FSGrangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum) {
  std::vector<MBB> bounds(size_t(std::max(rangeQueryNum, 0)), rangeQueryMBB);
  return FSGrangeQueryTest(bounds.data(), rangeQueryNum);
}

int SystemTest::FSGrangeQueryTest(const MBB* bounds, int rangeQueryNum) {
  if (!fsg) return GAT_ERR_STATE;
  lastBatch = RangeBatch{};
  RangeBatch& b = lastBatch;
  const double t0 = now_ms();
  if (int rc = fsg->rangeQueryBatchGPUPairs(bounds, rangeQueryNum, b)) return rc;
  return report("FSG", rangeQueryNum, b, now_ms() - t0);
}

int SystemTest::MortonGridRangeQueryTest(MBB rangeQueryMBB, int rangeQueryNum) {
  std::vector<MBB> bounds(size_t(std::max(rangeQueryNum, 0)), rangeQueryMBB);
  return MortonGridRangeQueryTest(bounds.data(), rangeQueryNum);
}

int SystemTest::MortonGridRangeQueryTest(const MBB* bounds, int rangeQueryNum) {
  if (!mgrid) return GAT_ERR_STATE;
  lastBatch = RangeBatch{};
  RangeBatch& b = lastBatch;
  const double t0 = now_ms();
  if (int rc = mgrid->rangeQueryBatchGPUPairs(bounds, rangeQueryNum, b)) return rc;
  return report("MortonGrid", rangeQueryNum, b, now_ms() - t0);
}

int SystemTest::similarityQueryTest(Trajectory t, int similarityScale, int similarityKValue) {
  if (!g) return GAT_ERR_STATE;
  if (similarityScale < 0 || similarityKValue < 0) return GAT_ERR_ARG;
  std::vector<Trajectory> q(size_t(similarityScale), t);
  const size_t n = size_t(similarityScale) * size_t(similarityKValue);
  last = Result{};
  cpuMs = 0;
  cpuIds.assign(cpuSimilarity ? n : 0, 0);
  if (cpuSimilarity) {
    const double t0 = now_ms();
    if (int rc = cpuSimilarity(q.data(), similarityScale, similarityKValue, cpuIds.data())) return rc;
    cpuMs = now_ms() - t0;
    if (verbose)
      std::printf("multi-core CPU similarity @ k=%d and #query=%d: %.3f ms\n", similarityKValue, similarityScale,
                  cpuMs);
  }
  gpuIds.assign(n, 0);
  const double t0 = now_ms();
  if (int rc = g->SimilarityQueryBatchOnGPU(q.data(), similarityScale, gpuIds.data(), similarityKValue)) return rc;
  last.wallMs = now_ms() - t0;
  if (verbose)
    std::printf("one GPU similarity @ k=%d and #query=%d: %.3f ms\n", similarityKValue, similarityScale, last.wallMs);
  return GAT_OK;
}
