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


def test_traffic_records_come_from_the_request_counters(monkeypatch):
    # profiles/hbm_traffic.json: per kernel and grid, the L2-to-memory read
    # requests by size (TCC_EA0_RDREQ_128B/64B/32B) and write requests
    # (TCC_EA0_WRREQ, of which TCC_EA0_WRREQ_64B are 64-byte and the rest
    # 32-byte), medians over the timed dispatches; bytes = read + write
    sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
        table = json.load(f)["kernels"]
    assert {"k3_iter2_511", "k2_iter_8191"} <= set(table)
    for key, rec in table.items():
        assert rec["read_bytes"] == 128 * rec["rdreq_128b"] + 64 * rec["rdreq_64b"] + 32 * rec["rdreq_32b"], key
        assert rec["write_bytes"] == 64 * rec["wrreq_64b"] + 32 * (rec["wrreq"] - rec["wrreq_64b"]), key
        assert rec["bytes"] == rec["read_bytes"] + rec["write_bytes"], key
        assert len(rec["sources_sha256"]) == 64 and rec["profile"].startswith("profiles/"), key
    # a record taken with other kernel sources is not reported as this build's
    monkeypatch.setattr(bench, "sources_sha256", lambda kernel: "0" * 64)
    rec, note = bench.load_traffic("k3_iter2", 511)
    assert rec is None and "stale" in note
    assert bench.load_traffic("k3_iter2", 12345)[0] is None


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
        def __init__(self, n, npts, eps, adj_eps, norm_kind):
            assert (npts, eps, adj_eps, norm_kind) == (256, 1e-10, 1e-10, 0)
            self.ctx, self.params, self.eps, self.nout, self.iter_max = FakeCtx(), None, eps, 200, 10 ** 7
            self.model = type("M", (), {"p": dict(norm_kind=norm_kind)})()
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
    assert len(calls) == 6 and all(c == (1e-10, 200, 999, 0) for c in calls)
