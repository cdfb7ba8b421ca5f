// Trajectory containers.  SamplePoint and Trajectory keep the reference's
// per-trajectory form (SamplePoint.h:5-20, Trajectory.h:8-27; trajectory ids are
// 1-based and tradb[0] is unused, main.cpp:116, Grid.cpp:215-221).
// TrajectoryDB is the bulk form the indexes are built from: every trajectory's
// points back to back, trajectory t (1-based) occupying
// [offsets[t-1], offsets[t]).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

class SamplePoint {
 public:
  SamplePoint() = default;
  SamplePoint(float lon1, float lat1, int time1, int tid1) : lon(lon1), lat(lat1), time(time1), tid(tid1) {}
  float lon = 0;
  float lat = 0;
  int time = 0;
  int tid = 0;
};

class Trajectory {
 public:
  Trajectory() = default;
  explicit Trajectory(int tid1) : tid(tid1) {}
  int tid = 0;
  std::vector<SamplePoint> points;
  int length = 0;
  void addSamplePoint(float lon, float lat, int time = 0) {
    points.emplace_back(lon, lat, time, tid);
    length = int(points.size());
  }
};

struct TrajectoryDB {
  std::vector<int64_t> offsets{0};  // trajNum + 1 entries
  std::vector<float> lon, lat;
  int trajNum() const { return int(offsets.size()) - 1; }
  int64_t pointNum() const { return offsets.back(); }
  int length(int tid) const { return int(offsets[tid] - offsets[tid - 1]); }

  // From the reference's 1-based array (db[1..traNum]).
  static TrajectoryDB fromArray(const Trajectory* db, int traNum) {
    TrajectoryDB out;
    out.offsets.reserve(size_t(traNum) + 1);
    for (int t = 1; t <= traNum; ++t) {
      for (int i = 0; i < db[t].length; ++i) {
        out.lon.push_back(db[t].points[i].lon);
        out.lat.push_back(db[t].points[i].lat);
      }
      out.offsets.push_back(int64_t(out.lon.size()));
    }
    return out;
  }
};

// A non-owning view of a TrajectoryDB-shaped set of arrays.
struct TrajectoryView {
  const int64_t* offsets = nullptr;  // trajNum + 1
  int trajNum = 0;
  const float* lon = nullptr;
  const float* lat = nullptr;
  int64_t pointNum() const { return offsets ? offsets[trajNum] : 0; }
  static TrajectoryView of(const TrajectoryDB& db) {
    return TrajectoryView{db.offsets.data(), db.trajNum(), db.lon.data(), db.lat.data()};
  }
};
