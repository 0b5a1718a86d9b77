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


def test_batch_workload_is_the_32_frame_configuration():
    # 32 frames of 1455 x 1455 (Nnum 15, 41 planes, PSF up to 345 px), 100-800
    # blobs per frame, Poisson peak 2000, seed 6, 8 plain iterations as target
    sys.path.insert(0, ROOT)
    import bench
    from flfhip.workloads import BATCHES, WORKLOADS
    base, T, bmin, bmax, peak, iters, seed = BATCHES["rl1455x32"]
    assert (base, T, bmin, bmax, peak, iters, seed) == ("rl1455", 32, 100, 800, 2000.0, 8, 6)
    assert WORKLOADS[base][1] * WORKLOADS[base][0] == 1455 and WORKLOADS[base][4] == 345
    import argparse
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'default="rl1455x32"' in src                     # the default bench line is this batch


def test_iterations_to_target():
    sys.path.insert(0, ROOT)
    import bench
    fits = [[0.5, 0.2], [0.3, 0.15], [0.1, 0.14]]
    assert bench._iters_to_target(fits, [0.3, 0.2]) == [1, 0]
    assert bench._iters_to_target(fits, [0.05, 0.14]) == [-1, 2]


def test_nils_scenario_inputs():
    """The reference benchmark's scenario (benchmarking.ipynb:14-28): the
    example image fixture reads as a 1273 x 1463 float32 image, and the
    optics give 25 planes z = -60..60 um."""
    sys.path.insert(0, os.path.join(ROOT, "fast-light-field_amd"))
    import numpy as np
    from flfhip import lfimage, psfgen
    from flfhip.workloads import NILS
    im = lfimage.LoadLightFieldTiff(os.path.join(ROOT, "tests", "golden", "nils_lfimage_1273x1463.tif"))
    assert im.shape == NILS["shape"] and im.dtype == np.float32
    assert float(im.min()) > 0 and float(im.max()) == 65535.0
    zs = psfgen.plane_depths(*NILS["optics"][8:])
    assert len(zs) == 25 and abs(zs[0] + 60e-6) < 1e-12 and abs(zs[-1] - 60e-6) < 1e-12
    assert NILS["optics"][3] == 19 and NILS["iters"] == 4
