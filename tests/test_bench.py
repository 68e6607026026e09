"""bench.py's contract on the CPU: one GPU only, and the measured HBM traffic
it reports comes from the committed profile table."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_table_matches_the_counter_sums():
    # bytes = (FETCH_SIZE x 2 + WRITE_SIZE) KiB x 1024, the calibration of
    # profiles/pt_iteration_kernels_hbm_traffic.md (k2_iter 8191^2: 2877982 and
    # 3164464 KiB; k3_iter2 511^3: 11341818 and 10856465 KiB; the table rounds
    # the median counter values to whole KiB)
    sys.path.insert(0, ROOT)
    import bench
    assert abs(bench.load_traffic("k2_iter_8191") - (2877982 * 2 + 3164464) * 1024) < 3 * 1024
    assert abs(bench.load_traffic("k3_iter2_511") - (11341818 * 2 + 10856465) * 1024) < 3 * 1024
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
        assert set(json.load(f)) >= {"k2_iter_8191", "k3_iter2_511"}
    assert bench.load_traffic("no_such_kernel") is None
