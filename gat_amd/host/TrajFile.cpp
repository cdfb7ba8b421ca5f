#include "TrajFile.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "DeviceArena.h"
#include "gat_host.h"

namespace {

// Whole file as a string; false when it cannot be opened.
bool slurp(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

// Lines as std::getline yields them (a last line without '\n' included).
std::vector<std::string> lines_of(const std::string& text) {
  std::vector<std::string> out;
  size_t at = 0;
  while (at < text.size()) {
    size_t nl = text.find('\n', at);
    if (nl == std::string::npos) nl = text.size();
    out.emplace_back(text, at, nl - at);
    at = nl + 1;
  }
  return out;
}

// Every field between separators, empty ones included (so "a,b," has three).
std::vector<std::string> fields_of(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t at = 0;
  for (;;) {
    const size_t k = s.find(sep, at);
    if (k == std::string::npos) {
      out.emplace_back(s, at);
      return out;
    }
    out.emplace_back(s, at, k - at);
    at = k + 1;
  }
}

// Non-empty fields, as strtok splits (PreProcess.cpp:240-300).
std::vector<std::string> tokens_of(const std::string& s, char sep) {
  std::vector<std::string> out;
  for (std::string& f : fields_of(s, sep))
    if (!f.empty()) out.push_back(std::move(f));
  return out;
}

// atof / atoi: the longest numeric prefix after leading blanks, 0 when none.
float to_float(const std::string& s) { return float(std::strtod(s.c_str(), nullptr)); }
int to_int(const std::string& s) { return int(std::strtol(s.c_str(), nullptr, 10)); }

// DateTime.cpp:8-11: note that 2000 and 2400 are not leap years here.
bool leap(int y) { return (y % 4 == 0 || y % 400 == 0) && (y % 100 != 0); }

int month_days(int y, int m) {
  static const int kDays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  return (m == 2 && leap(y)) ? 29 : kDays[m - 1];
}

// DateTime.cpp:15-25: "yyyy-mm-dd" by position.  A string too short for its
// fields (the reference's substr would throw) is malformed.
bool parse_date(const std::string& d, int* y, int* m, int* day) {
  if (d.size() < 8) return false;
  *y = to_int(d.substr(0, 4));
  *m = to_int(d.substr(5, 2));
  *day = to_int(d.substr(8, 2));
  return *y >= 0 && *m >= 1 && *m <= 12 && *day >= 1 && *day <= month_days(*y, *m);
}

// DateTime.cpp:28-42: 1-based day of the year.
int day_of_year(int y, int m, int day) {
  for (int k = 1; k < m; ++k) day += month_days(y, k);
  return day;
}

}  // namespace

int daysBetweenDates(const std::string& a, const std::string& b) {
  int y1, m1, d1, y2, m2, d2;
  if (!parse_date(a, &y1, &m1, &d1) || !parse_date(b, &y2, &m2, &d2)) return -1;
  if (y1 == y2 && m1 == m2) return std::abs(d1 - d2);
  if (y1 == y2) return std::abs(day_of_year(y1, m1, d1) - day_of_year(y2, m2, d2));
  if (y1 > y2) {
    std::swap(y1, y2);
    std::swap(m1, m2);
    std::swap(d1, d2);
  }
  int days = (leap(y1) ? 366 : 365) - day_of_year(y1, m1, d1) + day_of_year(y2, m2, d2);
  for (int y = y1 + 1; y < y2; ++y) days += leap(y) ? 366 : 365;
  return days;
}

// CalcInAxis.h:9-21, in float as the reference evaluates it (its sin / cos /
// acos take the float overloads; the colatitudes are used as radians).
float vehicleDistance(float latA, float lonA, float latB, float lonB) {
  const float pi = 3.1415926f;
  const float radius = 6371004.0f;
  const float ca = 90.0f - latA, cb = 90.0f - latB;
  const float c = std::sin(ca) * std::sin(cb) * std::cos(lonA - lonB) + std::cos(ca) * std::cos(cb);
  return radius * std::acos(c) * pi / 180.0f;
}

int readTrajectoryFile(const std::string& path, TrajectoryTable* out) {
  std::string text;
  if (!slurp(path, &text)) {
    gath_set_error("trajectory file %s: cannot be read", path.c_str());
    return GAT_ERR_ARG;
  }
  const std::vector<std::string> lines = lines_of(text);
  if (lines.size() < 3) {
    gath_set_error("trajectory file %s: fewer than 3 header lines", path.c_str());
    return GAT_ERR_ARG;
  }
  TrajectoryTable t;
  // PreProcess.cpp:239-242: the trajectory count.
  const int n = to_int(lines[0]);
  if (n < 0) {
    gath_set_error("trajectory file %s: negative trajectory count", path.c_str());
    return GAT_ERR_ARG;
  }
  // PreProcess.cpp:243-256: the bounds.
  const std::vector<std::string> b = tokens_of(lines[1], ',');
  if (b.size() < 4) {
    gath_set_error("trajectory file %s: line 2 needs xmin,ymin,xmax,ymax", path.c_str());
    return GAT_ERR_ARG;
  }
  t.xmin = to_float(b[0]);
  t.ymin = to_float(b[1]);
  t.xmax = to_float(b[2]);
  t.ymax = to_float(b[3]);
  // PreProcess.cpp:257-271: the lengths.
  const std::vector<std::string> lens = tokens_of(lines[2], ';');
  if (int(lens.size()) < n) {
    gath_set_error("trajectory file %s: %d lengths for %d trajectories", path.c_str(), int(lens.size()), n);
    return GAT_ERR_ARG;
  }
  t.offsets.assign(size_t(n) + 1, 0);
  for (int i = 0; i < n; ++i) {
    const int len = to_int(lens[size_t(i)]);
    if (len < 0) {
      gath_set_error("trajectory file %s: trajectory %d has a negative length", path.c_str(), i + 1);
      return GAT_ERR_ARG;
    }
    t.offsets[size_t(i) + 1] = t.offsets[size_t(i)] + len;
  }
  const size_t np = size_t(t.offsets.back());
  t.lon.resize(np);
  t.lat.resize(np);
  t.tid.resize(np);
  t.time.resize(np);
  // PreProcess.cpp:272-309: trajectory i on line i + 3, "lat,lon,tid,time;" per point.
  for (int i = 0; i < n; ++i) {
    const int64_t b0 = t.offsets[size_t(i)], len = t.offsets[size_t(i) + 1] - b0;
    if (len == 0) continue;
    if (size_t(i) + 3 >= lines.size()) {
      gath_set_error("trajectory file %s: trajectory %d is missing", path.c_str(), i + 1);
      return GAT_ERR_ARG;
    }
    const std::vector<std::string> pts = tokens_of(lines[size_t(i) + 3], ';');
    if (int64_t(pts.size()) < len) {
      gath_set_error("trajectory file %s: trajectory %d has %lld of %lld points", path.c_str(), i + 1,
                     (long long)pts.size(), (long long)len);
      return GAT_ERR_ARG;
    }
    for (int64_t k = 0; k < len; ++k) {
      const std::vector<std::string> f = fields_of(pts[size_t(k)], ',');
      if (f.size() < 4) {
        gath_set_error("trajectory file %s: trajectory %d point %lld needs lat,lon,tid,time", path.c_str(), i + 1,
                       (long long)k);
        return GAT_ERR_ARG;
      }
      const size_t d = size_t(b0 + k);
      t.lat[d] = to_float(f[0]);
      t.lon[d] = to_float(f[1]);
      t.tid[d] = to_int(f[2]);
      t.time[d] = to_int(f[3]);
    }
  }
  *out = std::move(t);
  return GAT_OK;
}

int writeTrajectoryFile(const std::string& path, const TrajectoryTable& t) {
  const int n = t.trajNum();
  if (n < 0 || t.offsets.empty() || t.offsets[0] != 0 || size_t(t.offsets.back()) != t.lon.size() ||
      t.lat.size() != t.lon.size() || t.tid.size() != t.lon.size() || t.time.size() != t.lon.size()) {
    gath_set_error("trajectory file %s: inconsistent table", path.c_str());
    return GAT_ERR_ARG;
  }
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    gath_set_error("trajectory file %s: cannot be written", path.c_str());
    return GAT_ERR_ARG;
  }
  // PreProcess.cpp:203-223: floats are printed as doubles with "%f".
  bool ok = std::fprintf(f, "%d\n%f,%f,%f,%f\n", n, double(t.xmin), double(t.ymin), double(t.xmax),
                         double(t.ymax)) > 0;
  for (int i = 0; ok && i < n; ++i)
    ok = std::fprintf(f, "%lld;", (long long)(t.offsets[size_t(i) + 1] - t.offsets[size_t(i)])) > 0;
  ok = ok && std::fputc('\n', f) != EOF;
  for (int i = 0; ok && i < n; ++i) {
    for (int64_t d = t.offsets[size_t(i)]; ok && d < t.offsets[size_t(i) + 1]; ++d)
      ok = std::fprintf(f, "%f,%f,%d,%d;", double(t.lat[size_t(d)]), double(t.lon[size_t(d)]), t.tid[size_t(d)],
                        t.time[size_t(d)]) > 0;
    ok = ok && std::fputc('\n', f) != EOF;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    gath_set_error("trajectory file %s: write failed", path.c_str());
    return GAT_ERR_ARG;
  }
  return GAT_OK;
}

int ingestVehicleCsv(const std::string& path, const std::string& baseDate, TrajectoryTable* out) {
  std::string text;
  if (!slurp(path, &text)) {
    gath_set_error("vehicle file %s: cannot be read", path.c_str());
    return GAT_ERR_ARG;
  }
  struct Point {
    float lon, lat;
    int32_t time;
  };
  std::vector<std::vector<Point>> traj(1);                   // index = trajectory id; 0 unused
  std::unordered_map<std::string, int> current;             // vehicle -> its newest trajectory
  TrajectoryTable t;
  // Trajectory::addSamplePoints (Trajectory.cpp:22-48): 0 appended; 1 the
  // trajectory is full; 2 more than an hour since its last point; 3 too fast.
  auto add = [&](int id, const Point& p) -> int {
    std::vector<Point>& v = traj[size_t(id)];
    if (v.size() >= 1024) return 1;  // MAXLENGTH, ConstDefine.h:24
    if (!v.empty()) {
      const Point& q = v.back();
      if (p.time - q.time > 3600) return 2;  // MAXGAP, ConstDefine.h:26
      if (vehicleDistance(p.lat, p.lon, q.lat, q.lon) / float(p.time - q.time) >= 50.0f) return 3;
    }
    v.push_back(p);
    return 0;
  };
  for (const std::string& line : lines_of(text)) {
    // PreProcess.cpp:179-200: lines of fewer than 12 fields are skipped.
    const std::vector<std::string> f = fields_of(line, ',');
    if (f.size() < 12) continue;
    const std::string& vid = f[2];
    if (vid == "0000") continue;  // PreProcess.cpp:70-71
    const std::vector<std::string> dt = fields_of(f[8], ' ');
    if (dt.size() < 2) continue;  // no clock time: the reference indexes past its split
    const std::vector<std::string> hms = fields_of(dt[1], ':');
    if (hms.size() < 3) continue;
    Point p;
    p.lon = to_float(f[3]);
    p.lat = to_float(f[4]);
    p.time = daysBetweenDates(baseDate, dt[0]) * 86400 + to_int(hms[0]) * 3600 + to_int(hms[1]) * 60 + to_int(hms[2]);
    // PreProcess.cpp:163-175: the reference's study area.
    if (p.lat < 25 || p.lat > 50 || p.lon < 90 || p.lon > 130) continue;
    // PreProcess.cpp:150-161: the bounds grow with every accepted line,
    // including points addSamplePoints later drops.
    t.ymin = std::min(t.ymin, p.lat);
    t.ymax = std::max(t.ymax, p.lat);
    t.xmin = std::min(t.xmin, p.lon);
    t.xmax = std::max(t.xmax, p.lon);
    auto it = current.find(vid);
    if (it == current.end()) {
      traj.emplace_back();
      const int id = int(traj.size()) - 1;
      current.emplace(vid, id);
      add(id, p);
      continue;
    }
    const int r = add(it->second, p);
    if (r == 1 || r == 2) {  // PreProcess.cpp:116-132: a new trajectory for the vehicle
      traj.emplace_back();
      it->second = int(traj.size()) - 1;
      add(it->second, p);
    }
  }
  const int n = int(traj.size()) - 1;
  t.offsets.assign(size_t(n) + 1, 0);
  for (int i = 1; i <= n; ++i) t.offsets[size_t(i)] = t.offsets[size_t(i) - 1] + int64_t(traj[size_t(i)].size());
  const size_t np = size_t(t.offsets.back());
  t.lon.resize(np);
  t.lat.resize(np);
  t.tid.resize(np);
  t.time.resize(np);
  for (int i = 1; i <= n; ++i) {
    size_t d = size_t(t.offsets[size_t(i) - 1]);
    for (const Point& p : traj[size_t(i)]) {
      t.lon[d] = p.lon;
      t.lat[d] = p.lat;
      t.tid[d] = i;
      t.time[d] = p.time;
      ++d;
    }
  }
  *out = std::move(t);
  return GAT_OK;
}
