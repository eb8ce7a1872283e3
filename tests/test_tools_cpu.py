"""Command-line tools: flag handling that needs no GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "grus_framework_amd", "bin")


@pytest.fixture(scope="module")
def cli():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "grus_framework_amd", "cli")])
    return BIN


def test_cli_usage_and_app_selection(cli):
    r = subprocess.run([os.path.join(cli, "grus_bfs"), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--input" in r.stderr and "--pull" in r.stderr
    r = subprocess.run([os.path.join(cli, "grus_bfs"), "--app", "nope"], capture_output=True,
                       text=True)
    assert r.returncode == 2


def test_cli_refuses_multi_gpu(cli):
    r = subprocess.run([os.path.join(cli, "grus_cc"), "--rmat", "8", "--ngpu", "4"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def _regions(path, windows):
    import json
    out = []
    for name, b, e, units in windows:
        out.append({"name": name, "units": units,
                    "begin": {"monotonic": b, "boottime": b + 10 ** 12},
                    "end": {"monotonic": e, "boottime": e + 10 ** 12}})
    path.write_text(json.dumps(out))


def test_rocprof_summary_keeps_only_timed_windows(tmp_path):
    # dispatches inside [1000, 2000] belong to the window; the ones before
    # (a build) and after (another phase) must not be counted
    from grus_framework_amd import rocprof_summary as rs
    _regions(tmp_path / "r.json", [("bfs", 1000, 2000, 2)])
    rows = ["Kernel_Name,Start_Timestamp,End_Timestamp",
            "k_build(int),500,900",
            "k_pull(int),1100,1400",
            "k_pull(int),1500,1800",
            "k_push(int),1900,1950",
            "k_other(int),2100,2500"]
    (tmp_path / "t.csv").write_text("\n".join(rows) + "\n")
    regions = rs.load_regions(str(tmp_path / "r.json"))
    trows = rs.load_rows(str(tmp_path / "t.csv"))
    clock = rs.pick_clock(regions, [int(r["Start_Timestamp"]) for r in trows])
    assert clock == "monotonic"
    text = rs.summarise_trace(regions, trows, clock)
    assert "k_build" not in text and "k_other" not in text
    # 2 units, 3 dispatches, (300 + 300 + 50) ns = 0.00065 ms in total
    assert "(2 units, 3 dispatches, 0.0003 ms kernel time per unit)" in text
    assert "| `k_pull` | 2 | 1.00 | 0.001 | 0.30 | 0.0003 | 92.3% |" in text


def test_rocprof_summary_counter_calibration(tmp_path):
    import json
    from grus_framework_amd import rocprof_summary as rs
    # the copy window reads 512 KiB per MiB copied -> factor 2; the timed
    # window then holds 100 + 300 KiB over 4 units -> 100 KiB x 1024 x 2
    _regions(tmp_path / "r.json", [("copy_1GiB", 0, 100, 1), ("pr_iter", 200, 400, 4)])
    rows = ["Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp",
            "copy,FETCH_SIZE,524288,10,20",
            "k_pull_window(x),FETCH_SIZE,100,210,220",
            "k_pull_window(x),FETCH_SIZE,300,300,320",
            "k_late(x),FETCH_SIZE,999999,500,520"]
    (tmp_path / "c.csv").write_text("\n".join(rows) + "\n")
    regions = rs.load_regions(str(tmp_path / "r.json"))
    crow = rs.load_rows(str(tmp_path / "c.csv"))
    out = tmp_path / "traffic.json"
    out.write_text(json.dumps({"pr_iter": {"write_bytes": 7}}))
    text = rs.summarise_counters(regions, crow, "monotonic", str(out))
    assert "bytes = KiB x 1024 x 2.0000" in text
    t = json.loads(out.read_text())
    assert t["pr_iter"]["fetch_bytes"] == 100 * 1024 * 2
    assert t["pr_iter"]["write_bytes"] == 7  # merged, not replaced
    assert "k_late" not in text


def test_traffic_kept_with_window_settings(tmp_path, monkeypatch):
    # a window's recorded settings travel into traffic.json, and bench.py
    # reports the traffic only for a run with the same settings
    import json
    import sys
    from grus_framework_amd import rocprof_summary as rs
    _regions(tmp_path / "r.json", [("copy_1GiB", 0, 100, 1), ("bfs_w", 200, 400, 2)])
    r = json.loads((tmp_path / "r.json").read_text())
    r[1]["config"] = {"bfs_tuning": [14.0, 64.0, 32]}
    (tmp_path / "r.json").write_text(json.dumps(r))
    rows = ["Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp",
            "copy,WRITE_SIZE,1048576,10,20",
            "k_pull(x),WRITE_SIZE,10,210,220"]
    (tmp_path / "c.csv").write_text("\n".join(rows) + "\n")
    out = tmp_path / "traffic.json"
    out.write_text(json.dumps({"bfs_w": {"fetch_bytes": 100}}))
    rs.summarise_counters(rs.load_regions(str(tmp_path / "r.json")),
                          rs.load_rows(str(tmp_path / "c.csv")), "monotonic", str(out))
    t = json.loads(out.read_text())
    assert t["bfs_w"]["config"] == {"bfs_tuning": [14.0, 64.0, 32]}
    assert t["bfs_w"]["write_bytes"] == 5 * 1024
    sys.path.insert(0, ROOT)
    import bench
    monkeypatch.setattr(bench, "TRAFFIC", str(out))
    assert bench.traffic("bfs_w", {"bfs_tuning": [14.0, 64.0, 32]}) == 100 + 5 * 1024
    assert bench.traffic("bfs_w", {"bfs_tuning": [14.0, 24.0, 32]}) is None
    assert bench.traffic("bfs_w") == 100 + 5 * 1024


def test_bench_watchdog_reports_without_a_stuck_leg():
    # a leg that never returns is abandoned: the line is printed with that
    # leg marked, then the process exits on its own
    code = ("import sys, time; sys.path.insert(0, %r); import bench; "
            "w = bench.Watchdog(0.5); "
            "w.line = {'value': 1.0, 'extra': {'pagerank': {'x': 1, 'roofline': {'f': 2}}}}; "
            "w.leg('stuck'); time.sleep(30); print('not reached')" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(out) == 1 and "not reached" not in r.stdout
    line = json.loads(out[0])
    assert line["extra"]["stuck"]["error"].startswith("no result within")
    assert line["roofline_pagerank"] == {"f": 2} and line["extra"]["pagerank"] == {"x": 1}
    assert "leg stuck" in r.stderr


SWEEP = os.path.join(ROOT, "grus_framework_amd", "cli", "run_sweep.sh")


def test_sweep_script_commands(tmp_path):
    # the reference's dataset list and sources by default (run.sh:9-16)
    r = subprocess.run(["bash", SWEEP, "--dry-run", "--data", "/d"], capture_output=True,
                       text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")[:-1]
    assert len(lines) == 7
    assert lines[4].endswith("grus_bfs --input /d/sk-2005.w.gr --src 10")
    assert lines[5].endswith("--input /d/friendster.w.gr --src 101")
    # own pairs, another app, extra flags; pr and cc take no source
    r = subprocess.run(["bash", SWEEP, "--dry-run", "--app", "cc", "--suffix", ".gr", "g1:3",
                        "--", "--method", "lp"], capture_output=True, text=True, timeout=30)
    assert r.stdout.strip().endswith("grus_cc --input %s/data/g1.gr --method lp"
                                     % os.path.expanduser("~"))
    # a missing input is skipped, not an error; an unknown app is
    r = subprocess.run(["bash", SWEEP, "--data", str(tmp_path), "x:0"], capture_output=True,
                       text=True, timeout=30)
    assert r.returncode == 0 and "not found, skipped" in r.stderr
    r = subprocess.run(["bash", SWEEP, "--app", "tc"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 64


def test_bench_serving_leg_failure_is_reported():
    # the serving leg runs in a child process; when it cannot run (here: no
    # GPU) the leg reports an error instead of taking the line down
    import argparse
    sys.path.insert(0, ROOT)
    import bench
    for leg in ("serving", "cu", "capped"):
        res = bench.run_leg_child(leg, argparse.Namespace(leg_limit=120.0, scale=10))
        assert set(res) == {"error"} and ("%s leg exited with status" % leg) in res["error"]
