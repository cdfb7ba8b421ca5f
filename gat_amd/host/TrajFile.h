// On-disk trajectory formats of the reference (PreProcess.cpp):
//   * the preprocessed text file (writeTraDataToFile / readTraFromFormatedFile,
//     PreProcess.cpp:203-315): line 1 the number of trajectories, line 2
//     "xmin,ymin,xmax,ymax", line 3 every trajectory's length followed by ';',
//     then one line per trajectory of "lat,lon,tid,time;" per point — latitude
//     first, every float printed with "%f";
//   * raw vehicle CSV ingest (the PreProcess(fileName, outFileName)
//     constructor, PreProcess.cpp:48-201, with Trajectory::addSamplePoints,
//     Trajectory.cpp:22-48): field 2 the vehicle id, 3 longitude, 4 latitude,
//     8 "yyyy-mm-dd hh:mm:ss"; a vehicle's points form one trajectory until it
//     reaches 1024 points or a gap of more than an hour, and a point implying
//     50 m/s or more is dropped.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

// Trajectories 1..n in file order, flattened.
struct TrajectoryTable {
  std::vector<int64_t> offsets{0};  // n + 1
  std::vector<float> lon, lat;
  std::vector<int32_t> tid, time;   // per point, as the file holds them
  float xmin = 180.0f, ymin = 90.0f, xmax = 0.0f, ymax = 0.0f;  // PreProcess.cpp:50-53
  int trajNum() const { return int(offsets.size()) - 1; }
};

// Returns GAT_OK or an error status (message through gath_last_error).
int readTrajectoryFile(const std::string& path, TrajectoryTable* out);
int writeTrajectoryFile(const std::string& path, const TrajectoryTable& table);
int ingestVehicleCsv(const std::string& path, const std::string& baseDate, TrajectoryTable* out);

// Helpers of the ingest, exposed for tests: whole days between two
// "yyyy-mm-dd" dates (DateTime.cpp:44-95, -1 for a malformed date) and the
// distance estimate of CalcInAxis.h:9-21.
int daysBetweenDates(const std::string& a, const std::string& b);
float vehicleDistance(float latA, float lonA, float latB, float lonB);
