// Autotuner of the batched forms (no reference counterpart): times candidate
// launch shapes on a sample of the caller's own workload, on the device, and
// leaves the fastest one set on the index.
//
//  * Batched range: rectangles per leaf-major work item (the share of a leaf's
//    rectangle list one wave takes), over {4, 8, 16, 32, 64}, then the
//    leaf-major form against the direct form at the sample's batch size (the
//    threshold between them moves to keep the faster one for that size);
//    timed as the device pipeline (selection + join).
//  * Batched EDR: the length split between the short and long buckets (which
//    LDS row-buffer class each pair runs in) and the workgroups per CU of each
//    bucket's launch, by coordinate descent (split first, then each bucket's
//    occupancy), timed as the DP kernels' device time.
#pragma once

#include <vector>

#include "FSG.h"
#include "Grid.h"
#include "STIG.h"

struct TuneRangeResult {
  int best = 0;                   // chosen rectangles per work item
  std::vector<int> candidates;
  std::vector<double> ms;         // median device pipeline time per candidate (leaf-major form)
  double directMs = 0;            // the same batch in the direct form
  int directMaxQueries = 0;       // resulting threshold between the two forms
};

struct TuneEdrResult {
  int split = 0, perCuShort = 0, perCuLong = 0;
  int wavesPerPair = 1;           // long pairs of one-round batches
  double smallMs = 0;             // that batch's kernel time at the chosen value
  double ms = 0;                  // best kernel time (median of reps)
  int evaluations = 0;
};

// Index is Grid, STIG or FSG (anything with rangeQueryBatchedGPU and engine()).
template <class Index>
int tuneRangeBatched(Index& index, const MBB* boxes, int nq, int reps, TuneRangeResult* out);

int tuneEdrBatched(Grid& grid, const TrajectoryView& queries, int reps, TuneEdrResult* out);
