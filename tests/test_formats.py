"""The reference's on-disk formats (PreProcess.cpp): the preprocessed
trajectory file and the raw vehicle-record ingest, product (TrajFile.cpp,
through the C ABI) against the oracle (oracle/formats.py), and both against
hand-derived answers.  CPU only."""
import numpy as np
import pytest

from gat_amd import GatError, api, workloads as W
from gat_amd import _native as N
from oracle import formats as F

f32 = np.float32


def _days(a, b):
    return int(N.host().gath_traj_days_between(a.encode(), b.encode()))


@pytest.mark.parametrize("a,b,want", [
    ("2014-07-01", "2014-07-03", 2),        # same month
    ("2014-07-03", "2014-07-01", 2),        # order does not matter
    ("2014-07-01", "2014-08-01", 31),       # July has 31 days
    ("2014-07-01", "2014-06-30", 1),        # before the base date: still positive
    ("2014-01-01", "2014-12-31", 364),
    ("2013-07-01", "2014-07-01", 365),
    ("2012-01-01", "2013-01-01", 366),      # 2012 is leap
    ("1999-12-31", "2000-03-01", 60),       # the reference's rule makes 2000 non-leap (61 in fact)
    ("2015-12-31", "2017-01-01", 367),      # 1 + 366 (2016) + 1 — with 2016 leap, 1 + 366 = 367
    ("2014-02-29", "2014-03-01", -1),       # invalid date
    ("2014-7-1", "2014-07-01", -1),         # malformed positions
    ("short", "2014-07-01", -1),
])
def test_days_between_known_answers(a, b, want):
    assert F.days_between(a, b) == want
    assert _days(a, b) == want


def test_leap_rule_known_answers():
    assert F.is_leap(2012) and F.is_leap(2016) and not F.is_leap(2014)
    assert not F.is_leap(1900) and not F.is_leap(2000)  # the reference's rule drops every century
    assert F.day_in_year(2014, 3, 1) == 60 and F.day_in_year(2012, 3, 1) == 61


def test_distance_closed_form_and_product():
    # Equal longitudes: C = cos(a)cos(b) + sin(a)sin(b) = cos(a - b), so the
    # estimate is R |a - b| pi / 180 with a, b the colatitudes taken as radians.
    r, pi = 6371004.0, 3.1415926
    for la, lb in [(30.0, 30.5), (31.2, 31.3), (40.0, 41.0)]:
        want = r * abs(la - lb) * pi / 180
        got = float(F.distance(la, 121.0, lb, 121.0))
        assert abs(got - want) <= 2e-3 * want, (la, lb, got, want)
    rng = np.random.default_rng(5)
    for _ in range(2000):
        la, lo, lb, lo2 = (f32(v) for v in (rng.uniform(25, 50), rng.uniform(90, 130),
                                            rng.uniform(25, 50), rng.uniform(90, 130)))
        if rng.random() < 0.3:  # nearby points, the regime the speed test sees
            lb, lo2 = f32(la + f32(rng.normal(0, 1e-3))), f32(lo + f32(rng.normal(0, 1e-3)))
        want = F.distance(la, lo, lb, lo2)
        got = f32(N.host().gath_traj_vehicle_distance(la, lo, lb, lo2))
        assert np.array_equal(np.float32(got).view(np.uint32), np.float32(want).view(np.uint32)) or \
            (np.isnan(got) and np.isnan(want))


def _line(vid, lon, lat, when, fields=12):
    row = ["x", "y", vid, f"{lon}", f"{lat}", "a", "b", "c", when] + ["z"] * (fields - 9)
    return ",".join(row)


def test_ingest_known_answers(tmp_path):
    rows = [
        _line("V1", 121.0, 31.0, "2014-07-01 00:00:00"),   # traj 1
        _line("V2", 121.5, 31.5, "2014-07-01 00:00:05"),   # traj 2
        _line("V1", 121.0001, 31.0, "2014-07-01 00:00:10"),  # traj 1 (slow)
        _line("V1", 122.0, 31.0, "2014-07-01 00:00:20"),   # 1 degree in 10 s: dropped, bounds still grow
        _line("V1", 121.0002, 31.0, "2014-07-01 01:00:10"),  # 3600 s after the last kept point: same trajectory
        _line("V1", 121.0003, 31.0, "2014-07-01 02:00:11"),  # 3601 s: new trajectory 3
        _line("V3", 121.0, 10.0, "2014-07-01 00:00:00"),   # latitude outside [25, 50]: skipped
        _line("0000", 121.0, 31.0, "2014-07-01 00:00:00"),  # id 0000: skipped
        _line("V4", 121.0, 31.0, "2014-07-01 00:00:00", fields=11),  # too few fields: skipped
        _line("V2", 121.5, 31.5, "2014-07-01 00:00:05"),   # same place, same time: 0/0 is NaN, kept
        _line("V2", 121.5, 31.5, "2014-07-02 00:00:05"),   # a day later: new trajectory 4
    ]
    path = tmp_path / "veh.csv"
    path.write_text("\n".join(rows) + "\n")
    want_lens = [3, 2, 1, 1]
    trajs, bounds = F.ingest(path.read_text(), "2014-07-01")
    assert [len(t) for t in trajs] == want_lens
    assert bounds == (f32(121.0), f32(31.0), f32(122.0), f32(31.5))
    assert [p[2] for p in trajs[0]] == [0, 10, 3610]
    assert trajs[3][0][2] == 86405
    db, tid, tm, bnd = api.ingest_vehicle_csv(path, "2014-07-01")
    assert list(np.diff(db.offsets)) == want_lens
    assert tuple(bnd) == bounds
    assert list(tm) == [p[2] for t in trajs for p in t]
    assert list(tid) == [1, 1, 1, 2, 2, 3, 4]


def test_ingest_splits_full_trajectories(tmp_path):
    rows = [_line("V", 121.0 + 1e-5 * (i % 2), 31.0, f"2014-07-01 00:{i // 60:02d}:{i % 60:02d}") for i in range(1030)]
    path = tmp_path / "veh.csv"
    path.write_text("\n".join(rows))  # no final newline
    trajs, _ = F.ingest(path.read_text(), "2014-07-01")
    assert [len(t) for t in trajs] == [1024, 6]
    db, _, _, _ = api.ingest_vehicle_csv(path, "2014-07-01")
    assert list(np.diff(db.offsets)) == [1024, 6]


def _random_csv(rng, nveh, nrows):
    vids = [f"V{i:03d}" for i in range(nveh)]
    state = {v: [rng.uniform(100, 120), rng.uniform(28, 45), int(rng.integers(0, 5000))] for v in vids}
    rows = []
    for _ in range(nrows):
        v = vids[int(rng.integers(nveh))]
        s = state[v]
        step = rng.choice([1, 5, 30, 600, 3600, 3601, 7200])
        s[2] += int(step)
        jump = rng.choice([1e-5, 1e-4, 1e-3, 0.5])
        s[0] += rng.normal(0, jump)
        s[1] += rng.normal(0, jump)
        t = s[2]
        day, rem = divmod(t, 86400)
        when = f"2014-07-{1 + day % 28:02d} {rem // 3600:02d}:{rem % 3600 // 60:02d}:{rem % 60:02d}"
        lon = f"{s[0]:.6f}" if rng.random() > 0.01 else "140.0"
        rows.append(_line(v, lon, f"{s[1]:.6f}", when, fields=12 + int(rng.integers(0, 3))))
        if rng.random() < 0.01:
            rows.append("garbage,line")
    return "\n".join(rows) + "\n"


def test_ingest_matches_oracle_on_random_records(tmp_path):
    rng = np.random.default_rng(17)
    text = _random_csv(rng, 40, 6000)
    path = tmp_path / "veh.csv"
    path.write_text(text)
    trajs, bounds = F.ingest(text, "2014-07-01")
    db, tid, tm, bnd = api.ingest_vehicle_csv(path, "2014-07-01")
    assert list(np.diff(db.offsets)) == [len(t) for t in trajs]
    want = np.array([(lon, lat, t) for tr in trajs for lon, lat, t in tr], dtype=object)
    assert np.array_equal(db.lon, np.array([p[0] for p in want], np.float32))
    assert np.array_equal(db.lat, np.array([p[1] for p in want], np.float32))
    assert list(tm) == [int(p[2]) for p in want]
    assert tuple(bnd) == bounds
    assert len(trajs) > 40  # gaps and full trajectories split vehicles


KNOWN_FILE = ("2\n"
              "121.000000,31.000000,122.000000,32.000000\n"
              "2;1;\n"
              "31.500000,121.250000,1,10;31.125000,121.750000,1,20;\n"
              "31.300000,121.400002,2,5;\n")


def test_reader_known_file(tmp_path):
    path = tmp_path / "pre.txt"
    path.write_text(KNOWN_FILE)
    db, tid, tm, bnd = api.read_trajectory_file(path)
    assert list(db.offsets) == [0, 2, 3]
    assert list(db.lat) == [f32(31.5), f32(31.125), f32(31.3)]  # latitude comes first in the file
    assert list(db.lon) == [f32(121.25), f32(121.75), f32("121.400002")]
    assert list(tid) == [1, 1, 2] and list(tm) == [10, 20, 5]
    assert tuple(bnd) == (f32(121), f32(31), f32(122), f32(32))
    trajs, obnd = F.read_text(KNOWN_FILE)
    assert obnd == tuple(bnd)
    assert [(p[0], p[1]) for t in trajs for p in t] == list(zip(db.lon, db.lat))


def test_writer_matches_oracle_text_and_reads_back(tmp_path):
    db = W.CFG1.db(200)
    lon = (db.lon * f32(40) + f32(90)).astype(np.float32)   # into the reference's study area
    lat = (db.lat * f32(25) + f32(25)).astype(np.float32)
    db = api.TrajSet(db.offsets, lon, lat)
    tm = np.arange(db.npoints, dtype=np.int32) * 7
    bounds = np.array([lon.min(), lat.min(), lon.max(), lat.max()], np.float32)
    path = tmp_path / "pre.txt"
    api.write_trajectory_file(path, db, time=tm, bounds=bounds)
    trajs = [[(db.lon[i], db.lat[i], int(tm[i])) for i in range(db.offsets[t], db.offsets[t + 1])]
             for t in range(db.ntraj)]
    assert path.read_text() == F.write_text(trajs, bounds)
    back, tid, tm2, bnd = api.read_trajectory_file(path)
    otrajs, obnd = F.read_text(path.read_text())
    assert (back.offsets == db.offsets).all() and (tm2 == tm).all()
    # "%f" keeps six decimals: the round trip is the reference's own lossy one.
    assert np.array_equal(back.lon, np.array([p[0] for t in otrajs for p in t], np.float32))
    assert np.array_equal(back.lat, np.array([p[1] for t in otrajs for p in t], np.float32))
    assert np.abs(back.lon - db.lon).max() <= 5e-6 and np.abs(back.lat - db.lat).max() <= 5e-6
    assert tuple(bnd) == obnd


def test_reader_rejects_truncated_files(tmp_path):
    path = tmp_path / "bad.txt"
    path.write_text("2\n0,0,1,1\n2;1;\n31.5,121.25,1,10;\n")
    with pytest.raises(GatError, match="trajectory 1 has 1 of 2 points"):
        api.read_trajectory_file(path)
    with pytest.raises(GatError, match="cannot be read"):
        api.read_trajectory_file(tmp_path / "missing.txt")


def test_ingested_file_feeds_an_index(tmp_path):
    """The reference's flow (main.cpp:117-131): ingest, write, read, build the
    Grid over the file's bounds."""
    rng = np.random.default_rng(3)
    csv = tmp_path / "veh.csv"
    csv.write_text(_random_csv(rng, 20, 2000))
    db, tid, tm, bnd = api.ingest_vehicle_csv(csv, "2014-07-01")
    pre = tmp_path / "pre.txt"
    api.write_trajectory_file(pre, db, tid, tm, bnd)
    db2, _, _, bnd2 = api.read_trajectory_file(pre)
    g = api.GridIndex(bnd2, 0.01, 256).build(db2, upload=False)
    assert g.info()[0] == db2.npoints
