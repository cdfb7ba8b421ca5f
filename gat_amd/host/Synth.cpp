#include "Synth.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {
uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
constexpr double kTwoPi = 6.283185307179586476925286766559;

inline float reflect01(float v) {
  if (v < 0.0f) v = -v;
  if (v > 1.0f) v = 2.0f - v;
  return v;
}
}  // namespace

Xoshiro256ss::Xoshiro256ss(uint64_t seed) {
  uint64_t x = seed;
  for (auto& w : s_) w = splitmix64(x);
}

uint64_t Xoshiro256ss::next() {
  const uint64_t result = rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl(s_[3], 45);
  return result;
}

double Xoshiro256ss::uniform() { return double(next() >> 11) * 0x1.0p-53; }

double Xoshiro256ss::normal() {
  if (has_spare_) {
    has_spare_ = false;
    return spare_;
  }
  const double u1 = 1.0 - uniform();  // (0, 1]
  const double u2 = uniform();
  const double r = std::sqrt(-2.0 * std::log(u1));
  spare_ = r * std::sin(kTwoPi * u2);
  has_spare_ = true;
  return r * std::cos(kTwoPi * u2);
}

uint64_t synth_mix(uint64_t a, uint64_t b) {
  uint64_t x = a ^ (b * 0xd1b54a32d192ed03ULL);
  splitmix64(x);
  return splitmix64(x);
}

void synth_lengths_uniform(uint64_t seed, int64_t n, int lo, int hi, int32_t* out) {
  Xoshiro256ss rng(synth_mix(seed, 0x4c454e55ULL));
  const double span = double(hi - lo + 1);
  for (int64_t i = 0; i < n; ++i) {
    int v = lo + int(rng.uniform() * span);
    out[i] = std::min(v, hi);
  }
}

void synth_lengths_lognormal(uint64_t seed, int64_t n, double median, double sigma, int lo, int hi, int32_t* out) {
  Xoshiro256ss rng(synth_mix(seed, 0x4c454e4cULL));
  const double mu = std::log(median);
  for (int64_t i = 0; i < n; ++i) {
    double v = std::round(std::exp(mu + sigma * rng.normal()));
    v = std::min(std::max(v, double(lo)), double(hi));
    out[i] = int32_t(v);
  }
}

namespace {
void walk_into(Xoshiro256ss& rng, float x, float y, int64_t n, double step_sigma, float* lon, float* lat) {
  for (int64_t i = 0; i < n; ++i) {
    lon[i] = x;
    lat[i] = y;
    x = reflect01(x + float(step_sigma * rng.normal()));
    y = reflect01(y + float(step_sigma * rng.normal()));
  }
}
}  // namespace

void synth_walks(uint64_t seed, int64_t ntraj, const int64_t* offsets, double step_sigma, float* lon, float* lat,
                 int threads) {
  threads = std::max(1, threads);
  auto body = [&](int part) {
    for (int64_t t = part; t < ntraj; t += threads) {
      Xoshiro256ss rng(synth_mix(seed, uint64_t(t + 1)));
      const float x0 = float(rng.uniform());
      const float y0 = float(rng.uniform());
      walk_into(rng, x0, y0, offsets[t + 1] - offsets[t], step_sigma, lon + offsets[t], lat + offsets[t]);
    }
  };
  if (threads == 1) {
    body(0);
    return;
  }
  std::vector<std::thread> pool;
  for (int p = 0; p < threads; ++p) pool.emplace_back(body, p);
  for (auto& th : pool) th.join();
}

void synth_rects(uint64_t seed, int64_t n, double side_lo, double side_hi, bool log_uniform, double hot_frac,
                 int nhot, double hot_sigma, float* out4) {
  Xoshiro256ss rng(synth_mix(seed, 0x52454354ULL));
  std::vector<double> hx(std::max(nhot, 0)), hy(std::max(nhot, 0));
  for (int h = 0; h < nhot; ++h) {
    hx[h] = rng.uniform();
    hy[h] = rng.uniform();
  }
  const double llo = std::log10(side_lo), lhi = std::log10(side_hi);
  for (int64_t i = 0; i < n; ++i) {
    double side = log_uniform ? std::pow(10.0, llo + (lhi - llo) * rng.uniform())
                              : side_lo + (side_hi - side_lo) * rng.uniform();
    double cx, cy;
    if (nhot > 0 && rng.uniform() < hot_frac) {
      int h = std::min(int(rng.uniform() * nhot), nhot - 1);
      cx = hx[h] + hot_sigma * rng.normal();
      cy = hy[h] + hot_sigma * rng.normal();
    } else {
      cx = rng.uniform();
      cy = rng.uniform();
    }
    out4[4 * i + 0] = float(cx - side / 2);
    out4[4 * i + 1] = float(cy - side / 2);
    out4[4 * i + 2] = float(cx + side / 2);
    out4[4 * i + 3] = float(cy + side / 2);
  }
}

void synth_queries(uint64_t seed, int nq, const int32_t* qlen, const int64_t* q_offsets, const TrajectoryView& db,
                   int copy_mode, double noise_sigma, double step_sigma, float* qlon, float* qlat) {
  for (int i = 0; i < nq; ++i) {
    Xoshiro256ss rng(synth_mix(seed ^ 0x5155455259ULL, uint64_t(i)));
    const bool copy = db.trajNum > 0 && (copy_mode == 1 || (copy_mode == 2 && (i & 1)));
    float* ox = qlon + q_offsets[i];
    float* oy = qlat + q_offsets[i];
    const int64_t len = qlen[i];
    if (!copy) {
      const float x0 = float(rng.uniform());
      const float y0 = float(rng.uniform());
      walk_into(rng, x0, y0, len, step_sigma, ox, oy);
      continue;
    }
    int src = 1 + std::min(int(rng.uniform() * db.trajNum), db.trajNum - 1);
    const int64_t b = db.offsets[src - 1];
    const int64_t slen = db.offsets[src] - b;
    const int64_t ncopy = std::min(len, slen);
    for (int64_t k = 0; k < ncopy; ++k) {
      ox[k] = reflect01(db.lon[b + k] + float(noise_sigma * rng.normal()));
      oy[k] = reflect01(db.lat[b + k] + float(noise_sigma * rng.normal()));
    }
    if (ncopy < len) {
      const float x = db.lon[b + ncopy - 1], y = db.lat[b + ncopy - 1];
      walk_into(rng, reflect01(x + float(step_sigma * rng.normal())), reflect01(y + float(step_sigma * rng.normal())),
                len - ncopy, step_sigma, ox + ncopy, oy + ncopy);
    }
  }
}
