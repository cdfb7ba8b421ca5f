"""Oracle for the reference's on-disk trajectory formats — test infrastructure
only (tests/ import it; the product never does).

Restates, in Python, independently of gat_amd/host/TrajFile.cpp:
  * DaysBetween2Date and its helpers (DateTime.cpp:8-95);
  * calculateDistance (CalcInAxis.h:9-21), evaluated in float32 with the C
    library's sinf / cosf / acosf (the reference's calls take the float
    overloads);
  * the raw vehicle-record ingest (PreProcess.cpp:48-201 with
    Trajectory::addSamplePoints, Trajectory.cpp:22-48);
  * writeTraDataToFile (PreProcess.cpp:203-223) and readTraFromFormatedFile
    (PreProcess.cpp:225-315).
Pinned by hand-derived known answers in tests/test_formats.py.
"""
import ctypes
import ctypes.util

import numpy as np

f32 = np.float32

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _name in ("sinf", "cosf", "acosf"):
    getattr(_libm, _name).restype = ctypes.c_float
    getattr(_libm, _name).argtypes = [ctypes.c_float]


def _sinf(x):
    return f32(_libm.sinf(float(x)))


def _cosf(x):
    return f32(_libm.cosf(float(x)))


def _acosf(x):
    return f32(_libm.acosf(float(x)))


# ------------------------------------------------------------ DateTime.cpp

def is_leap(year: int) -> bool:
    """DateTime.cpp:8-11 (2000 is not a leap year under this rule)."""
    return (year % 4 == 0 or year % 400 == 0) and (year % 100 != 0)


_DAYS = [31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]


def _atoi(s: str) -> int:
    """atoi: optional blanks and sign, then digits; 0 when there are none."""
    s = s.lstrip(" \t\n\r\f\v")
    sign, i = 1, 0
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def string_to_date(date: str):
    """DateTime.cpp:15-25: (ok, year, month, day); strings too short for the
    reference's substr calls are malformed."""
    if len(date) < 8:
        return False, 0, 0, 0
    y, m, d = _atoi(date[0:4]), _atoi(date[5:7]), _atoi(date[8:10])
    days = list(_DAYS)
    if is_leap(y):
        days[1] = 29
    ok = y >= 0 and 0 < m <= 12 and 0 < d <= days[m - 1]
    return ok, y, m, d


def day_in_year(y: int, m: int, d: int) -> int:
    """DateTime.cpp:28-42."""
    days = list(_DAYS)
    if is_leap(y):
        days[1] = 29
    return d + sum(days[: m - 1])


def days_between(date1: str, date2: str) -> int:
    """DateTime.cpp:44-95."""
    ok1, y1, m1, d1 = string_to_date(date1)
    ok2, y2, m2, d2 = string_to_date(date2)
    if not (ok1 and ok2):
        return -1
    if y1 == y2 and m1 == m2:
        return abs(d1 - d2)
    if y1 == y2:
        return abs(day_in_year(y1, m1, d1) - day_in_year(y2, m2, d2))
    if y1 > y2:
        y1, m1, d1, y2, m2, d2 = y2, m2, d2, y1, m1, d1
    rest = (366 if is_leap(y1) else 365) - day_in_year(y1, m1, d1)
    middle = sum(366 if is_leap(y) else 365 for y in range(y1 + 1, y2))
    return rest + day_in_year(y2, m2, d2) + middle


# ------------------------------------------------------------ CalcInAxis.h

def distance(lat_a, lon_a, lat_b, lon_b) -> np.float32:
    """CalcInAxis.h:9-21 in float32, operation by operation."""
    pi, r = f32(3.1415926), f32(6371004)
    mlat_a, mlat_b = f32(90) - f32(lat_a), f32(90) - f32(lat_b)
    c = _sinf(mlat_a) * _sinf(mlat_b) * _cosf(f32(lon_a) - f32(lon_b)) + _cosf(mlat_a) * _cosf(mlat_b)
    return r * _acosf(c) * pi / f32(180)


# ------------------------------------------------------------ PreProcess.cpp

def _atof32(s: str) -> np.float32:
    """atof (strtod: the longest numeric prefix) then the float conversion."""
    s = s.lstrip(" \t\n\r\f\v")
    best = 0.0
    for k in range(len(s), 0, -1):
        try:
            best = float(s[:k])
            break
        except ValueError:
            continue
    return f32(best)


MAXLENGTH, MAXGAP = 1024, 3600  # ConstDefine.h:24, 26


def ingest(text: str, base_date: str):
    """PreProcess.cpp:48-141: returns (trajectories, bounds), trajectories a
    list (ids 1..n in creation order) of lists of (lon, lat, time)."""
    xmin, xmax, ymin, ymax = f32(180), f32(0), f32(90), f32(0)
    trajs = [None]  # index 0 unused
    newest = {}
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()

    def add(tid, lon, lat, t):
        """Trajectory.cpp:22-48."""
        pts = trajs[tid]
        if len(pts) >= MAXLENGTH:
            return 1
        if pts:
            plon, plat, pt = pts[-1]
            if t - pt > MAXGAP:
                return 2
            with np.errstate(divide="ignore", invalid="ignore"):
                speed = distance(lat, lon, plat, plon) / f32(t - pt)
            if speed >= f32(50):
                return 3
        pts.append((lon, lat, t))
        return 0

    for line in lines:
        parts = line.split(",")
        if len(parts) < 12:  # PreProcess.cpp:181-185
            continue
        vid = parts[2]
        dt = parts[8].split(" ")
        if len(dt) < 2:
            continue
        hms = dt[1].split(":")
        if len(hms) < 3:
            continue
        lon, lat = _atof32(parts[3]), _atof32(parts[4])
        t = days_between(base_date, dt[0]) * 86400 + _atoi(hms[0]) * 3600 + _atoi(hms[1]) * 60 + _atoi(hms[2])
        if lat < 25 or lat > 50 or lon < 90 or lon > 130:  # validPoint, PreProcess.cpp:163-175
            continue
        if vid == "0000":
            continue
        if lat < ymin:
            ymin = lat
        if lat > ymax:
            ymax = lat
        if lon < xmin:
            xmin = lon
        if lon > xmax:
            xmax = lon
        if vid not in newest:
            trajs.append([])
            newest[vid] = len(trajs) - 1
            add(newest[vid], lon, lat, t)
            continue
        r = add(newest[vid], lon, lat, t)
        if r in (1, 2):
            trajs.append([])
            newest[vid] = len(trajs) - 1
            assert add(newest[vid], lon, lat, t) == 0
    return trajs[1:], (xmin, ymin, xmax, ymax)


def write_text(trajs, bounds, tids=None) -> str:
    """PreProcess.cpp:203-223: trajs as lists of (lon, lat, time)."""
    out = [f"{len(trajs)}\n", "%f,%f,%f,%f\n" % tuple(float(b) for b in bounds)]
    out.append("".join(f"{len(t)};" for t in trajs) + "\n")
    for i, t in enumerate(trajs):
        tid = (i + 1) if tids is None else tids[i]
        out.append("".join("%f,%f,%d,%d;" % (float(lat), float(lon), tid, tm) for lon, lat, tm in t) + "\n")
    return "".join(out)


def read_text(text: str):
    """PreProcess.cpp:225-315: (trajectories as lists of (lon, lat, tid,
    time), bounds)."""
    lines = text.split("\n")
    n = _atoi(lines[0])
    b = [x for x in lines[1].split(",") if x]
    bounds = tuple(_atof32(x) for x in b[:4])
    lens = [_atoi(x) for x in lines[2].split(";") if x][:n]
    trajs = []
    for i in range(n):
        pts = []
        if lens[i]:
            toks = [x for x in lines[i + 3].split(";") if x]
            for k in range(lens[i]):
                lat, lon, tid, tm = toks[k].split(",")[:4]
                pts.append((_atof32(lon), _atof32(lat), _atoi(tid), _atoi(tm)))
        trajs.append(pts)
    return trajs, bounds
