"""bench.py --summarize on synthetic rocprofv3 CSVs: the timed window is the
dispatches between the two fxMark dispatches, whatever runs before and after
it, and the traffic per launch is the mean scaled counter over that window."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_csv(path, rows, fields):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)


def test_summarize_cuts_the_window_between_the_marks(tmp_path):
    # 3 force dispatches before the window (300 us), 2 inside (100 us), 1 after (900 us)
    seq = [("computeForce", 300)] * 3 + [("fxMark", 1), ("computeForce", 100), ("computePressure", 50),
                                         ("computeForce", 100), ("computePressure", 50), ("fxMark", 1),
                                         ("computeForce", 900)]
    trace, fetch, write = [], [], []
    t = 0
    for i, (k, us) in enumerate(seq):
        trace.append({"Dispatch_Id": i + 1, "Kernel_Name": k + "(int)", "Start_Timestamp": t,
                      "End_Timestamp": t + us * 1000})
        t += us * 1000 + 10
        inside = 3 < i < 8
        fetch.append({"Dispatch_Id": i + 1, "Kernel_Name": k, "Counter_Name": "FETCH_SIZE",
                      "Counter_Value": 1000.0 if inside else 99999.0})
        write.append({"Dispatch_Id": i + 1, "Kernel_Name": k, "Counter_Name": "WRITE_SIZE",
                      "Counter_Value": 10.0 if inside else 99999.0})
    f1 = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]
    f2 = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    write_csv(tmp_path / "kt.csv", trace, f1)
    write_csv(tmp_path / "f.csv", fetch, f2)
    write_csv(tmp_path / "w.csv", write, f2)
    line = {"value": 1.0, "ms_per_step": 0.15, "config": {
        "workload": "dambreak_4m: x", "stage_threads": [64, 128, 1024, 256, 256, 128], "stage_ppt": [1, 1, 2, 1, 1, 1],
        "timed_window": {"steps": 2, "warmup": 5, "autotuned": True, "graph": True}}}
    (tmp_path / "line.json").write_text("noise\n" + json.dumps(line) + "\n")
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize", "--line",
                          str(tmp_path / "line.json"), "--trace", str(tmp_path / "kt.csv"), "--fetch",
                          str(tmp_path / "f.csv"), "--write", str(tmp_path / "w.csv"), "--json",
                          str(tmp_path / "t.json")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "| computeForce | 2 | 100.0 |" in out.stdout
    # other counters: mean over the window's dispatches
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize", "--line",
                          str(tmp_path / "line.json"), "--counters", str(tmp_path / "f.csv")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "| computeForce | 2 | 1000 |" in out.stdout
    t = json.loads((tmp_path / "t.json").read_text())
    assert t["window"] == {"steps": 2, "warmup": 5, "autotuned": True, "graph": True}
    assert t["window_avg_us"]["computeForce"] == 100.0
    # FETCH x2 + WRITE x1, KiB -> bytes
    assert t["bytes_per_launch"]["computeForce"] == (1000.0 * 2 + 10.0) * 1024
    assert t["launch"]["computeForce"] == [256, 1] and t["launch"]["countingSortFullRanked"] == [1024, 2]


def test_traffic_attaches_only_to_the_same_window_and_launch(tmp_path, monkeypatch):
    sys.path.insert(0, ROOT)
    import bench

    os.makedirs(tmp_path / "profiles")
    win = {"steps": 20, "warmup": 5, "autotuned": True, "graph": True}
    (tmp_path / "profiles" / "traffic_dambreak_4M_literal.json").write_text(json.dumps(
        {"window": win, "launch": {"computeForce": [256, 1]}, "bytes_per_launch": {"computeForce": 5.0e8}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.traffic_for("dambreak_4m", "computeForce", win, (256, 1))[0] == 500000000
    assert bench.traffic_for("dambreak_4m", "computeForce", win, (128, 1)) == (None, None)
    assert bench.traffic_for("dambreak_4m", "computeForce", dict(win, steps=200), (256, 1)) == (None, None)
    assert bench.traffic_for("dambreak_4m", "computePressure", win, (256, 1)) == (None, None)
    # block sizes pinned for the counter runs instead of tuned: same window
    assert bench.traffic_for("dambreak_4m", "computeForce", dict(win, autotuned=False), (256, 1))[0] == 500000000
    assert bench.traffic_for("dambreak_4m", "computeForce", dict(win, graph=False), (256, 1)) == (None, None)


def test_summarize_derives_the_counter_limits(tmp_path):
    """td_busy = TD_TD_BUSY_sum / 256 CUs over GRBM_GUI_ACTIVE / 8 XCDs; VALU
    issue = SQ_INSTS_VALU / 1024 SIMDs over the dispatch's duration; both
    averaged over the window's dispatches only."""
    seq = [("computeForce", 9)] + [("fxMark", 1), ("computeForce", 1), ("computeForce", 1), ("fxMark", 1)]
    c1, c2 = [], []
    for i, (k, scale) in enumerate(seq):
        t0 = 1000 * i
        base = {"Dispatch_Id": i + 1, "Kernel_Name": k, "Start_Timestamp": t0, "End_Timestamp": t0 + 2000 * scale}
        # inside the window: 2 us per dispatch, 1.024e6 VALU -> 1000 per SIMD -> 5e8 per SIMD per second
        c1.append(dict(base, Counter_Name="SQ_INSTS_VALU", Counter_Value=1.024e6 * scale))
        # TD busy 0.9 of 1000 cycles per XCD on every CU
        c2.append(dict(base, Counter_Name="TD_TD_BUSY_sum", Counter_Value=900.0 * 256 * scale))
        c2.append(dict(base, Counter_Name="GRBM_GUI_ACTIVE", Counter_Value=1000.0 * 8 * scale))
    f = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Counter_Name", "Counter_Value"]
    write_csv(tmp_path / "c1.csv", c1, f)
    write_csv(tmp_path / "c2.csv", c2, f)
    write_csv(tmp_path / "fw.csv", [dict(r, Counter_Name=n, Counter_Value=1.0) for r in c1 for n in ("FETCH_SIZE",)], f)
    write_csv(tmp_path / "ww.csv", [dict(r, Counter_Name="WRITE_SIZE", Counter_Value=1.0) for r in c1], f)
    line = {"value": 1.0, "ms_per_step": 0.15, "config": {
        "workload": "dambreak_4m: x", "stage_threads": [64, 128, 1024, 256, 256, 128], "stage_ppt": [1] * 6,
        "timed_window": {"steps": 2, "warmup": 5, "autotuned": False, "graph": True}}}
    (tmp_path / "line.json").write_text(json.dumps(line) + "\n")
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize", "--line",
                          str(tmp_path / "line.json"), "--fetch", str(tmp_path / "fw.csv"), "--write",
                          str(tmp_path / "ww.csv"), "--counters", str(tmp_path / "c1.csv"), str(tmp_path / "c2.csv"),
                          "--json", str(tmp_path / "t.json")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lim = json.loads((tmp_path / "t.json").read_text())["limits"]["computeForce"]
    assert lim == {"td_busy": 0.9, "valu_per_simd_s": 5e8}
