// Deterministic synthetic workloads: Gaussian random-walk trajectories,
// range rectangles and EDR query trajectories of the benchmark shapes.
//
// Generator: xoshiro256** seeded through splitmix64, with an in-house
// Box-Muller transform, so the same seed gives the same data on every machine
// with the same libm.  Every trajectory and every query draws from its own
// stream (seeded from the workload seed and its index), so generation runs on
// many threads and stays deterministic.
#pragma once

#include <cstdint>

#include "Trajectory.h"

class Xoshiro256ss {
 public:
  explicit Xoshiro256ss(uint64_t seed);
  uint64_t next();
  double uniform();  // [0, 1), 53 random bits
  double normal();   // N(0, 1), Box-Muller with the second variate cached

 private:
  uint64_t s_[4];
  double spare_ = 0;
  bool has_spare_ = false;
};

uint64_t synth_mix(uint64_t a, uint64_t b);

// Lengths uniform in [lo, hi].
void synth_lengths_uniform(uint64_t seed, int64_t n, int lo, int hi, int32_t* out);
// Lengths round(exp(ln(median) + sigma * z)) clipped to [lo, hi].
void synth_lengths_lognormal(uint64_t seed, int64_t n, double median, double sigma, int lo, int hi, int32_t* out);
// Random walks: start uniform in [0,1)^2, steps N(0, step_sigma) per axis added
// in float, reflected at the edges of [0,1]^2.  Trajectory t (1-based) fills
// [offsets[t-1], offsets[t]).
void synth_walks(uint64_t seed, int64_t ntraj, const int64_t* offsets, double step_sigma, float* lon, float* lat,
                 int threads);
// Axis-aligned squares (xmin, ymin, xmax, ymax) with side uniform in
// [side_lo, side_hi] (log-uniform when log_uniform), centres uniform in [0,1)^2
// except a hot_frac share drawn from N(hotspot, hot_sigma) around one of
// `nhot` uniform hotspots.
void synth_rects(uint64_t seed, int64_t n, double side_lo, double side_hi, bool log_uniform, double hot_frac,
                 int nhot, double hot_sigma, float* out4);
// EDR query trajectories with lengths qlen[i], written from q_offsets[i].  Query
// i is a noisy copy of a random database trajectory when copy_mode says so
// (0: never, 1: always, 2: odd i), otherwise an independent walk.  A copy takes
// the database trajectory's first points with N(0, noise_sigma) noise added and,
// when the source is shorter, continues it as a walk.
void synth_queries(uint64_t seed, int nq, const int32_t* qlen, const int64_t* q_offsets, const TrajectoryView& db,
                   int copy_mode, double noise_sigma, double step_sigma, float* qlon, float* qlat);
