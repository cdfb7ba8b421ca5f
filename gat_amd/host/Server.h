// Serving loop over a resident GT-quadtree index (no reference counterpart:
// the reference answers one pre-built batch per call, SystemTest.cpp:24-249).
//
// The index's point store, tree and trajectory store stay in HBM.  Requests —
// range rectangles and EDR top-k queries — arrive over time; a range request
// costs the device memory of its result and list buffers, bounded exactly from
// the leaves it selects (found on arrival by the host tree walk whose rule the
// device selection applies), so an admitted range batch never outgrows the
// workspace reserved for its slot; a top-k request costs its query points, its
// k results and its filter survivors (12 bytes each, estimated by the survivors
// per query of the batches already run).  Continuous batching: each range slot
// (its own stream and workspace over the shared store) and the top-k slot
// take, whenever they are free, every waiting request of their kind that fits
// their memory budget, run
// the batch with the batched form, and hand each request its result; work of
// different slots overlaps on the device.  Requests that arrive meanwhile wait
// for the next free slot.  The plain mode answers one request at a time with
// the plain drop-in path, for comparison.
#pragma once

#include <cstdint>
#include <vector>

#include "Grid.h"
#include "gat_host.h"

struct ServeConfig {
  int rangeSlots = 2;             // range batches in flight (each its own stream and workspace)
  int maxRangeBatch = 4096;       // rectangles per range batch
  int maxTopkBatch = 256;         // queries per top-k batch
  double rangeBudgetBytes = 4e9;  // candidate/result bytes one range batch may reserve
  double topkBudgetBytes = 1e9;   // query/survivor bytes one top-k batch may reserve
  int k = 20;
  double eps = 5e-4;
  bool batched = true;  // false: the plain per-request path, one request at a time in arrival order
};

struct ServeStats {
  int64_t rangeBatches = 0, topkBatches = 0;
  int64_t rangeDone = 0, topkDone = 0;
  double wallMs = 0;
  double peakRangeBytes = 0, peakTopkBytes = 0;  // largest admitted batch charge
  double rangeBudgetBytes = 0, topkBudgetBytes = 0;  // budgets after the free-memory check
  int64_t peakTopkSurvivors = 0;  // most filter survivors of one top-k batch
  int64_t deviceAllocs = 0;       // device allocations while serving (buffers that grew)
};

class Server {
 public:
  Server(Grid& grid, const ServeConfig& cfg) : grid_(grid), cfg_(cfg) {}
  // Replays `reqs` (arrival times in ms from the start of the run, ascending)
  // and fills one outcome per request; top-k requests also fill k ids and
  // distances at topkIds / topkDists + i * k (ascending (distance, id)).
  // Top-k query points are (qlon, qlat)[qoff, qoff + qlen), inside the
  // arrays' nqpoints points.
  int run(const gath_request* reqs, int n, const float* qlon, const float* qlat, int64_t nqpoints, gath_outcome* out,
          int32_t* topkIds, int32_t* topkDists, ServeStats* stats);

  // Device bytes the result and list buffers of one range request need: an
  // upper bound (kRangeBytesPerPair per bounded result pair).
  double rangeBytes(const float* box) const;
  // Result pairs (bound) of a rectangle selecting `leaves` leaves of `points`.
  static int64_t rangePairBound(int64_t points, int64_t leaves);
  // Device bytes one top-k request is charged, with `survivors` filter
  // survivors expected.
  double topkBytes(int qlen, double survivors) const;
  // 8 B per result pair plus list entries (64 B per selected leaf) at one per
  // 16 pairs of the bound.
  static constexpr double kRangeBytesPerPair = 12.0;

 private:
  Grid& grid_;
  ServeConfig cfg_;
};
