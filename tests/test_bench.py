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


def test_inversion_gradient_leg_drives_chunked_forward_solves(monkeypatch):
    # bench_c5_grad's control flow with a stand-in for the device inversion:
    # forward solves are driven in check-aligned chunks until converged and
    # every phase lands in the record
    sys.path.insert(0, ROOT)
    import numpy as np
    import bench
    import ptstokes.adjoint as adjoint

    calls = []

    class FakeCtx:
        def __init__(self):
            self.calls = calls

        def pt_solve(self, params, eps, nout, iter_max, norm_kind):
            self.calls.append((eps, nout, iter_max, norm_kind))
            done = len(self.calls) % 3 == 0
            return dict(iters=500 if done else iter_max + 1, err=1e-11 if done else 1e-3, converged=done)

        def sync(self):
            pass

    class FakeInversion:
        def __init__(self, n, npts, eps, adj_eps):
            assert (npts, eps, adj_eps) == (256, 1e-10, 1e-10)
            self.ctx, self.params, self.eps, self.nout, self.iter_max = FakeCtx(), None, eps, 200, 10 ** 7
            self.model = type("M", (), {"p": dict(norm_kind=1)})()
            self.last, self.obs = {}, None

        def grf_logmu(self, seed):
            return np.zeros((n, n))

        def set_logmu(self, m):
            pass

        def top_velocity(self):
            return np.ones(256)

        def dJ_dVx(self):
            return None, 0.25

        def gradient(self):
            self.last["adjoint"] = dict(iters=42, err=1e-11)
            return np.ones((n, n))

        def close(self):
            pass

    n = 15
    monkeypatch.setattr(adjoint, "Inversion", FakeInversion)
    out = bench.bench_c5_grad(n, chunk=1000)
    assert out["observe_iters"] == out["forward_iters"] == 2 * 1000 + 500
    assert out["adjoint_iters"] == 42 and out["J"] == 0.25 and out["grad_finite"]
    assert out["forward_err"] == 1e-11
    # each chunk runs exactly `chunk` iterations at the inversion's tolerance and check interval
    assert len(calls) == 6 and all(c == (1e-10, 200, 999, 1) for c in calls)
