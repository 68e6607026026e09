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

        def adjoint_set_source(self, g, h):
            adj_calls.append("source")

        def adjoint_solve(self, params, eps, nout, iter_max):
            adj_calls.append((eps, nout, iter_max))
            done = len(adj_calls) == 3
            return dict(iters=42 if done else iter_max, err=1e-11 if done else 1e-4, converged=done)

        def grad_logmu(self, params):
            adj_calls.append("grad")
            return np.ones((n, n))

    adj_calls = []

    class FakeInversion:
        def __init__(self, n, npts, eps, adj_eps, norm_kind):
            assert (npts, eps, adj_eps, norm_kind) == (256, 1e-10, 1e-10, 0)
            self.ctx, self.params, self.eps, self.nout, self.iter_max = FakeCtx(), None, eps, 200, 10 ** 7
            self.adj_eps = adj_eps
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

        def close(self):
            pass

    n = 15
    monkeypatch.setattr(adjoint, "Inversion", FakeInversion)
    out = bench.bench_c5_grad(n, chunk=1000)
    assert out["observe_iters"] == out["forward_iters"] == 2 * 1000 + 500
    assert out["adjoint_iters"] == 1000 + 42 and out["J"] == 0.25 and out["grad_finite"]
    # the adjoint: source set once, check-aligned chunks from the current λ, then the gradient
    assert adj_calls == ["source", (1e-10, 200, 1000), (1e-10, 200, 1000), "grad"]
    assert out["forward_err"] == 1e-11
    # each chunk runs exactly `chunk` iterations at the inversion's tolerance and check interval
    assert len(calls) == 6 and all(c == (1e-10, 200, 999, 0) for c in calls)


def test_tts_leg_drives_chunked_fp64_and_one_correction_per_call(monkeypatch):
    # bench_tts's control flow with a stand-in model: the FP64 loop in
    # check-aligned chunks until converged, the defect correction one
    # correction per call until converged, per-field differences recorded
    sys.path.insert(0, ROOT)
    import numpy as np
    import bench
    import ptstokes

    log = []

    class FakeCtx:
        def __init__(self):
            self.dc = 0

        def sync(self):
            pass

        def pt_solve(self, params, eps, nout, iter_max, norm_kind, use_graph=True):
            log.append(("pt", eps, nout, iter_max, norm_kind))
            done = sum(1 for e in log if e[0] == "pt") == 3
            return dict(iters=iter_max + 1, err=5e-9 if done else 1e-3, converged=done)

        def dc_solve(self, params, eps, norm_kind, max_outer, rtol_inner, nout_inner):
            log.append(("dc", eps, norm_kind, max_outer, rtol_inner, nout_inner))
            self.dc += 1
            done = self.dc == 2
            return dict(outer=1, converged=done, inner_iters=100 * self.dc, err=2e-9 if done else 1e-5,
                        err_outer=[1.0, 1e-5])

    class FakeModel:
        made = 0

        def __init__(self, preset, n, nout=500):
            assert preset == "Stokes3D_grf"
            FakeModel.made += 1
            self.second = FakeModel.made == 2
            self.ctx = FakeCtx()
            self.p = dict(eps=1e-8, norm_kind=1, iter_max=10 ** 6)

        def setup(self):
            return self

        def pt_params(self):
            return None

        def get(self, name):
            a = np.arange(125, dtype=float).reshape(5, 5, 5) + 1.0
            return a * (1.0 + 1e-12) if self.second else a

        def close(self):
            pass

    monkeypatch.setattr(ptstokes, "StokesModel", FakeModel)
    out = bench.bench_tts(7, nout=500, chunk=2000)
    assert out["fp64_iters"] == 3 * 2000 and out["fp64_converged"] and out["fp64_err"] == 5e-9
    assert [e for e in log if e[0] == "pt"] == [("pt", 1e-8, 500, 1999, 1)] * 3
    assert out["dc_outer"] == 2 and out["dc_fp32_sweeps"] == 300 and out["dc_sweeps_per_correction"] == [100, 200]
    assert [e for e in log if e[0] == "dc"] == [("dc", 1e-8, 1, 1, 1e-4, 500)] * 2
    assert out["dc_converged"] and out["dc_err"] == 2e-9
    assert 0 < out["dc_vs_fp64_max_rel_err"]["Vx"] < 2e-12 and out["dc_vs_fp64_max_rel_err"]["Pt"] < 1e-11
