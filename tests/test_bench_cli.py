"""bench.py's command line on the CPU: one GPU only, and the traffic line it
reads from the committed rocprof summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_committed_traffic_summary_is_machine_readable():
    path = os.path.join(ROOT, "profiles", "k_tile_stage_spheres3d_traffic.md")
    line = [ln for ln in open(path) if ln.startswith('{"traffic"')]
    assert len(line) == 1
    t = json.loads(line[0])["traffic"]
    assert t["workload"] == "spheres3d" and t["nodes"] == 34229841
    # calibrated by the copy kernel: two fetched bytes per counted byte, one written
    assert abs(t["fetch_factor"] - 2.0) < 0.01 and abs(t["write_factor"] - 1.0) < 0.01
    assert t["bytes_per_launch"] > 1.232e9  # at least the algorithmic bytes
