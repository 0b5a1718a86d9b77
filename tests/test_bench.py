"""CPU checks of bench.py's contract: one GPU only (N > 1 exits with a
message before touching a GPU), the workload tables it measures, and the
traffic figure it reads from the committed rocprof summary."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_workloads_and_traffic_summary():
    sys.path.insert(0, ROOT)
    import bench
    from flfhip.workloads import RAGGED, WORKLOADS
    assert WORKLOADS["rl1455"][:3] == (15, 97, 20)          # Nnum, lenslets, z = -20..20 (41 planes)
    assert RAGGED["ragged64"][0] == 64
    t = bench.measured_traffic("rl1455")
    assert t is not None and 1e9 < t < 2e10                 # bytes per k_mac launch, from profiles/
    assert bench.measured_traffic("no_such_workload") is None
