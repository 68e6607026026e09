#!/usr/bin/env python3
"""Benchmark of the PT iteration on one MI355X.

A "step" is one PT iteration (the single-pass kernel: Pt/τ update, V update
and free-slip copies for the whole grid).  The headline workload is the 3D
Maxwell visco-elastic block at nx = ny = nz = 511 in FP64 (G = 1, dt = 1,
x-z pure shear, 1e-3 sphere), reported as the reference's effective
throughput T_eff = A_eff / t_it with A_eff = 29·n³·8 B
(scripts/stokes_3D/Stokes3D_ve_perf.jl:229: 10 fields read and written + 9
read-only).  2D Maxwell VE at nx = ny = 8191 (dt = η/G/10, A_eff = 18·n²·8 B,
Stokes2D_ve_perf.jl:173) is reported alongside.

Timing: HIP events recorded on the library's own stream around exactly
`--steps` iterations, after `--warmup` untimed ones, with the stream drained
on both sides.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pseudotransientstokes.jl_amd"))

HBM_PEAK_GBS = 8000.0
A100_2D_VE_8191_GBS = 932.8  # output/out_Stokes2D_ve_perf_ampere.txt:45
A100_3D_VE_511_GBS = 732.2   # output/out_Stokes3D_ve_perf_ampere.txt:15
METRIC = "T_eff GB/s (% of HBM peak) + time-to-‖res‖∞<1e-8, FP32-refined vs FP64"


class HipEvents:
    """hipEvent timing on an explicit stream (the one libpts launches on)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        rc = self.hip.hipEventCreate(ctypes.byref(e))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")
        return e

    def record(self, e, stream):
        if self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)


def time_iterations(model, steps, warmup, ev):
    """Wall and event time of `steps` PT iterations after `warmup` ones."""
    model.iterate(warmup)
    model.ctx.sync()
    s = model.ctx.stream()
    a, b = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(a, s)
    model.iterate(steps)
    ev.record(b, s)
    model.ctx.sync()
    wall = time.perf_counter() - t0
    return ev.elapsed_ms(a, b), wall


def bench_2d(n, steps, warmup, ev, contexts=1):
    """2D VE iteration time on `contexts` contexts created one after another
    (each freed before the next).  As in 3D, the kernel's HBM rate depends on
    where the arrays land physically; the library selects among a few
    placements on a context's first iteration call (timed on its own, outside
    the timed region), and what it finds depends on the device's history, so
    the median context is reported with all of them."""
    from ptstokes import StokesModel
    a_eff = 18.0 * n * n * 8
    runs = []
    for _ in range(max(1, contexts)):
        # 2D Maxwell VE: G = η = 1, dt = η/G/10, pure shear εbg = 1, 1e-3 inclusion of radius 1
        m = StokesModel("Stokes2D_ve_perf", n=n, dt=0.1).setup()
        m.ctx.assign_tau()
        t0 = time.perf_counter()
        m.ctx.pt_iterate(m.pt_params(), 0)  # first call: placement of the field blocks, no iteration
        m.ctx.sync()
        place_s = round(time.perf_counter() - t0, 3)
        ms, wall = time_iterations(m, steps, warmup, ev)
        runs.append(dict(ms_per_it=ms / steps, wall_ms_per_it=wall * 1e3 / steps, device_bytes=m.ctx.device_bytes(),
                         placement_ms=m.ctx.placement_times(), placement_s=place_s))
        m.close()
    order = sorted(range(len(runs)), key=lambda i: runs[i]["ms_per_it"])
    out = dict(runs[order[len(order) // 2]])
    out.update(n=n, steps=steps, A_eff_bytes=a_eff, T_eff_GBs=a_eff / (out["ms_per_it"] / 1e3) / 1e9,
               ms_per_it_by_context=[round(r["ms_per_it"], 4) for r in runs],
               placement_ms_by_context=[r["placement_ms"] for r in runs],
               placement_s_by_context=[r["placement_s"] for r in runs])
    return out


def bench_3d(n, steps, warmup, ev, contexts=1):
    """3D VE iteration time on `contexts` freshly allocated contexts.  The
    kernel's HBM rate depends on where the 28 arrays land physically (which
    channels concurrently streamed arrays share), which differs from one
    allocation to the next, so the median over several is reported with all
    of them.  The library selects among a few placements on a context's first
    iteration call (pts_placement_times): that call is timed on its own and
    its kernel times are listed per context."""
    from ptstokes import StokesModel
    per_ctx, walls, place_ms, place_s = [], [], [], []
    a_eff = 29.0 * n ** 3 * 8
    dev_bytes = 0
    for _ in range(max(1, contexts)):
        m = StokesModel("Stokes3D_ve_perf", n=n).setup()
        m.ctx.assign_tau()
        t0 = time.perf_counter()
        m.ctx.pt_iterate(m.pt_params(), 0)  # first call: placement of the field blocks, no iteration
        m.ctx.sync()
        place_s.append(round(time.perf_counter() - t0, 3))
        place_ms.append(m.ctx.placement_times())
        ms, wall = time_iterations(m, steps, warmup, ev)
        per_ctx.append(ms / steps)
        walls.append(wall * 1e3 / steps)
        dev_bytes = m.ctx.device_bytes()
        m.close()
    order = sorted(range(len(per_ctx)), key=lambda i: per_ctx[i])
    mid = order[len(order) // 2]
    t_it = per_ctx[mid] / 1e3
    out = dict(n=n, steps=steps, ms_per_it=per_ctx[mid], wall_ms_per_it=walls[mid], A_eff_bytes=a_eff,
               T_eff_GBs=a_eff / t_it / 1e9, device_bytes=dev_bytes,
               ms_per_it_by_context=[round(t, 4) for t in per_ctx], placement_ms_by_context=place_ms,
               placement_s_by_context=place_s)
    return out


def bench_tts(n, nout=500, rtol_inner=1e-4, use_graph=True, chunk=10000, max_outer=20):
    """Time to ‖res‖∞ < 1e-8 on the 3D viscous random-viscosity problem
    (log10 μ a seeded Gaussian random field clipped to ±2, 1e-3 sphere with
    ρg_z = -1): FP64 PT against FP32 PT sweeps inside FP64 defect
    correction, both from the same initial state, with the per-field error of
    the mixed-precision solution against the FP64 one.  The FP64 loop runs in
    chunks of `chunk` iterations (a multiple of nout, so its checks are those
    of one loop) and the defect correction one correction per call (the
    outer loop carries no state between corrections), each reporting its
    progress."""
    import numpy as np
    from ptstokes import StokesModel
    names = ["Pt", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz"]
    m = StokesModel("Stokes3D_grf", n=n, nout=nout).setup()
    p = m.p
    m.ctx.sync()
    t0 = time.perf_counter()
    it64 = 0
    while True:
        r64 = m.ctx.pt_solve(m.pt_params(), p["eps"], nout, chunk - 1, p["norm_kind"], use_graph=use_graph)
        it64 += r64["iters"]
        print(f"# TTS {n}^3 FP64 PT: {it64} iterations, err {r64['err']:.3e}, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if r64["converged"] or it64 >= p["iter_max"]:
            break
    m.ctx.sync()
    t64 = time.perf_counter() - t0
    ref = {f: m.get(f) for f in names}
    m.close()
    m = StokesModel("Stokes3D_grf", n=n).setup()
    m.ctx.sync()
    t0 = time.perf_counter()
    outer, sweeps, per_corr, rdc = 0, 0, [], None
    while True:
        rdc = m.ctx.dc_solve(m.pt_params(), p["eps"], p["norm_kind"], 1, rtol_inner, nout)
        if rdc["outer"] == 0:
            break
        outer += 1
        sweeps += rdc["inner_iters"]
        per_corr.append(rdc["inner_iters"])
        print(f"# TTS {n}^3 FP32+DC: correction {outer}, {rdc['inner_iters']} sweeps, err {rdc['err']:.3e}, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if rdc["converged"] or outer >= max_outer:
            break
    m.ctx.sync()
    tdc = time.perf_counter() - t0
    rel = {}
    for f, v in ref.items():
        a = m.get(f)
        if f == "Pt":  # P is fixed up to a constant: compare interior cells, mean removed
            a, v = a[1:-1, 1:-1, 1:-1], v[1:-1, 1:-1, 1:-1]
            a, v = a - a.mean(), v - v.mean()
        rel[f] = float(np.max(np.abs(a - v)) / np.max(np.abs(v)))
    m.close()
    return dict(workload=f"3D viscous Stokes, log10 mu GRF (seed 1234, corr. 0.5, sigma 1, clip 2) + 1e-3 sphere "
                         f"with rhog_z=-1, mu smoothed by 20 passes (10 leave the PT iteration unstable on this "
                         f"contrast), nx=ny=nz={n}, stop ||res||_inf < 1e-8 (scaled max-norm)",
                fp64_s=round(t64, 3), fp64_iters=it64, fp64_err=r64["err"], fp64_converged=r64["converged"],
                dc_s=round(tdc, 3), dc_outer=outer, dc_fp32_sweeps=sweeps, dc_sweeps_per_correction=per_corr,
                dc_err=rdc["err"], dc_converged=rdc["converged"], time_ratio_dc_over_fp64=round(tdc / t64, 3),
                dc_vs_fp64_max_rel_err=rel)


def bench_solve(preset, n, nt, a_eff, label, chunk=10000, **kw):
    """Time to solution of nt physical steps of a preset (assign τ_o, then the
    PT loop to ‖res‖∞ < 1e-8 checked every 200 iterations).  The loop is
    driven in chunks of `chunk` iterations (a multiple of the check interval,
    so the checks and the result are those of one loop) to report progress."""
    from ptstokes import StokesModel
    m = StokesModel(preset, n=n, nt=nt, iter_max=10 ** 7, nout=200, norm_kind=1, eps=1e-8, **kw).setup()
    p = m.p
    m.ctx.sync()
    steps = []
    t_all = time.perf_counter()
    for it in range(nt):
        m.ctx.assign_tau()
        t0 = time.perf_counter()
        iters = 0
        while True:
            r = m.ctx.pt_solve(m.pt_params(), p["eps"], p["nout"], chunk - 1, p["norm_kind"], use_graph=True)
            iters += r["iters"]
            print(f"# {label} step {it + 1}: {iters} iterations, err {r['err']:.3e}", file=sys.stderr, flush=True)
            if r["converged"]:
                break
            if iters >= 40 * chunk:
                raise RuntimeError(f"{label} step {it + 1} did not converge in {iters} iterations")
        steps.append(dict(iters=iters, s=round(time.perf_counter() - t0, 2), err=r["err"]))
    wall = time.perf_counter() - t_all
    total = sum(st["iters"] for st in steps)
    place = m.ctx.placement_times()
    m.close()
    return dict(steps=steps, iterations=total, wall_s=round(wall, 2), ms_per_it=round(wall * 1e3 / total, 4),
                T_eff_GBs=round(a_eff * total / wall / 1e9, 1), placement_ms=place)


def bench_c3_solve(n=511, nt=5):
    """The 3D Maxwell VE block (Stokes3D_ve_perf.jl kernels, G = 1, dt = 1,
    x-z pure shear, 1e-3 sphere), FP64."""
    out = bench_solve("Stokes3D_ve_perf", n, nt, 29.0 * n ** 3 * 8, f"3D VE {n}^3")
    out["workload"] = (f"3D Maxwell VE Stokes, nx=ny=nz={n}, FP64, {nt} physical steps, each to "
                       "||res||_inf < 1e-8 (check every 200 iterations)")
    return out


def bench_c2_solve(n=8191, nt=1):
    """The 2D Maxwell VE configuration (Stokes2D_ve_perf.jl kernels, G = 1,
    dt = η/G/10 = 0.1, pure shear, 1e-3 inclusion), FP64."""
    out = bench_solve("Stokes2D_ve_perf", n, nt, 18.0 * n * n * 8, f"2D VE {n}^2", chunk=20000, dt=0.1)
    out["workload"] = (f"2D Maxwell VE Stokes, nx=ny={n}, dt=0.1, FP64, {nt} physical steps, each to "
                       "||res||_inf < 1e-8 (check every 200 iterations)")
    return out


def bench_c5_grad(n, true_seed=4323, seed=1234, chunk=20000):
    """Time to ∂J/∂ln μ of the 2D viscous inversion: ln μ a seeded Gaussian
    random field, J = ½Σ|Vx − Vx_obs|² at 256 seeded points on the top
    boundary, observations from a hidden ln μ (one forward solve).  Then a
    forward PT solve to the reference's err (scaled RMS) 1e-10, the adjoint
    PT solve to a relative change of 1e-10, and the gradient kernel; each
    phase timed on its own.  The forward solves run
    in chunks of `chunk` iterations (a multiple of the check interval, so the
    checks and the result are those of one loop) to report progress; a solve
    whose err grows 10^4-fold above its best is stopped as diverged.  The
    hidden field's seed is 4323: at 2047^2 the PT iteration diverges for the
    fields of seeds 4321 and 4322 (err falls to ~1e-2, then grows; at CFL
    0.9/sqrt(2) down to 0.6/sqrt(2) and with 20 or 40 smoothing passes alike),
    while it converges to the roundoff floor for seeds 1234 and 4323."""
    import numpy as np
    from ptstokes.adjoint import Inversion
    # the reference's err (scaled RMS): the scaled max-norm has a roundoff floor
    # above 1e-10 at 2047^2 (≈3e-11 at 511^2, growing with n)
    inv = Inversion(n=n, npts=256, eps=1e-10, adj_eps=1e-10, norm_kind=0)
    truth, m0 = inv.grf_logmu(true_seed), inv.grf_logmu(seed)

    def forward(label):
        iters, best = 0, float("inf")
        while True:
            r = inv.ctx.pt_solve(inv.params, inv.eps, inv.nout, chunk - 1, inv.model.p["norm_kind"])
            iters += r["iters"]
            print(f"# C5 {n}^2 {label}: {iters} iterations, err {r['err']:.3e}", file=sys.stderr, flush=True)
            if r["converged"]:
                return dict(iters=iters, err=r["err"])
            best = min(best, r["err"])
            if not r["err"] < 1e4 * best:
                raise RuntimeError(f"C5 {label} diverged (err {r['err']:.3e} after {iters} iterations)")
            if iters >= inv.iter_max:
                raise RuntimeError(f"C5 {label} did not converge in {iters} iterations")

    inv.set_logmu(truth)
    t0 = time.perf_counter()
    obs = forward("observation forward")
    inv.obs = inv.top_velocity().copy()
    t_obs = time.perf_counter() - t0
    inv.set_logmu(m0)
    inv.ctx.sync()
    t0 = time.perf_counter()
    fwd = forward("forward")
    t_fwd = time.perf_counter() - t0
    g, J = inv.dJ_dVx()
    t0 = time.perf_counter()
    # the adjoint loop continues from the current λ, so chunks of a multiple of
    # nout iterations check where one loop would
    inv.ctx.adjoint_set_source(g, None)
    adj = dict(iters=0)
    while True:
        r = inv.ctx.adjoint_solve(inv.params, inv.adj_eps, inv.nout, chunk)
        adj = dict(iters=adj["iters"] + r["iters"], err=r["err"])
        print(f"# C5 {n}^2 adjoint: {adj['iters']} iterations, change {r['err']:.3e}", file=sys.stderr, flush=True)
        if r["converged"]:
            break
        if adj["iters"] >= inv.iter_max:
            raise RuntimeError(f"C5 adjoint did not converge in {adj['iters']} iterations")
    grad = inv.ctx.grad_logmu(inv.params)
    t_adj = time.perf_counter() - t0
    print(f"# C5 {n}^2: adjoint + gradient {t_adj:.2f} s, {adj['iters']} iterations", file=sys.stderr, flush=True)
    out = dict(workload=f"2D viscous inversion, nx=ny={n}, ln mu GRF (seed {seed}; observations from seed "
                        f"{true_seed}), 256 top-boundary Vx points, forward PT to err (scaled RMS) < 1e-10, adjoint PT "
                        f"to a relative change < 1e-10",
               observe_s=round(t_obs, 3), observe_iters=obs["iters"], forward_s=round(t_fwd, 3),
               forward_iters=fwd["iters"], forward_err=fwd["err"], adjoint_gradient_s=round(t_adj, 3),
               adjoint_iters=adj["iters"], adjoint_err=adj["err"], J=J, grad_max_abs=float(np.max(np.abs(grad))),
               grad_finite=bool(np.all(np.isfinite(grad))), time_to_gradient_s=round(t_fwd + t_adj, 3))
    inv.close()
    return out


def bench_c5_floor(n, seed=1234, chunk=20000, max_s=300.0, stall_chunks=6, **overrides):
    """Where the FP64 forward solve of the 2D inversion stops converging: the
    reference's err (scaled RMS) and the scaled max-norm of the same residuals
    after every `chunk` iterations, until the max-norm has not fallen 5 %
    below its best for `stall_chunks` chunks (its roundoff floor) or after
    `max_s` seconds.  Each point re-evaluates the residuals (compute_Res!) of
    the current state."""
    from ptstokes.adjoint import Inversion
    inv = Inversion(n=n, npts=256, eps=1e-30, norm_kind=0, **overrides)
    inv.set_logmu(inv.grf_logmu(seed))
    p = inv.params
    dx, dy = inv.model.dx, inv.model.dy
    lx = inv.model.p["lx"]
    curve, best, since = [], float("inf"), 0
    iters = 0
    t0 = time.perf_counter()
    while True:
        r = inv.ctx.pt_solve(p, 1e-30, inv.nout, chunk - 1, 0)
        iters += r["iters"]
        inv.ctx.compute_Res_perf(1.0 / dx, 1.0 / dy)
        nr = inv.ctx.norms_stored()
        e_rms, _ = inv.ctx.err_from_norms(nr, lx, 0)
        e_inf, _ = inv.ctx.err_from_norms(nr, lx, 1)
        el = time.perf_counter() - t0
        curve.append([iters, float(e_rms), float(e_inf)])
        print(f"# C5 floor {n}^2: {iters} iterations, err {e_rms:.3e} (RMS) {e_inf:.3e} (max), {el:.1f} s",
              file=sys.stderr, flush=True)
        if e_inf < 0.95 * best:
            best, since = e_inf, 0
        else:
            since += 1
        if since >= stall_chunks or el > max_s or not e_inf < 1e6:
            break
    inv.close()
    floor = since >= stall_chunks
    return dict(workload=f"2D viscous inversion forward solve, nx=ny={n}, ln mu GRF (seed {seed}) {overrides}",
                iterations=iters, seconds=round(time.perf_counter() - t0, 2), floor_reached=floor,
                min_err_rms=min(c[1] for c in curve), min_err_max=min(c[2] for c in curve),
                curve=curve if len(curve) <= 64 else curve[:: max(1, len(curve) // 64)] + [curve[-1]])


def bench_c4_kernels(n, ev):
    """Iteration kernels of the 3D viscous random-viscosity problem (log10 μ
    a seeded Gaussian random field, 1e-3 sphere with ρg_z = -1) at full size:
    the FP64 PT iteration (HIP events over 10 iterations after 2), and one
    FP32 correction sweep of the defect correction (the difference of a
    120-sweep and a 20-sweep correction, after a warm-up correction that
    allocates the FP32 arrays).  The FP32 context keeps its FP64 arrays
    ungrouped (no ping-pong partners), so that both precisions' arrays fit."""
    from ptstokes import StokesModel
    a_eff = 24.0 * n ** 3 * 8
    m = StokesModel("Stokes3D_grf", n=n).setup()
    p = m.pt_params()
    m.ctx.pt_iterate(p, 2)
    m.ctx.sync()
    s = m.ctx.stream()
    a, b = ev.event(), ev.event()
    ev.record(a, s)
    m.ctx.pt_iterate(p, 10)
    ev.record(b, s)
    m.ctx.sync()
    ms64 = ev.elapsed_ms(a, b) / 10
    bytes64 = m.ctx.device_bytes()
    m.close()
    print(f"# 3D GRF {n}^3 FP64 iteration {ms64:.3f} ms", file=sys.stderr, flush=True)
    old = os.environ.get("PTS_3D_GROUPS")
    os.environ["PTS_3D_GROUPS"] = "0"
    try:
        m = StokesModel("Stokes3D_grf", n=n).setup()
    finally:
        if old is None:
            os.environ.pop("PTS_3D_GROUPS", None)
        else:
            os.environ["PTS_3D_GROUPS"] = old
    ts = []
    for k in (20, 20, 120):
        t0 = time.perf_counter()
        m.ctx.dc_solve(p, 1e-30, 1, max_outer=1, rtol_inner=0.0, nout_inner=k, iter_max_inner=k)
        m.ctx.sync()
        ts.append(time.perf_counter() - t0)
    ms32 = (ts[2] - ts[1]) / 100 * 1e3
    bytes32 = m.ctx.device_bytes()
    m.close()
    print(f"# 3D GRF {n}^3 FP32 correction sweep {ms32:.3f} ms", file=sys.stderr, flush=True)
    return dict(workload=f"3D viscous Stokes, log10 mu GRF + 1e-3 sphere with rhog_z=-1, 20 smoothing passes, "
                         f"nx=ny=nz={n}",
                fp64_ms_per_it=round(ms64, 4), fp64_A_eff_bytes=a_eff, fp64_T_eff_GBs=round(a_eff / ms64 / 1e6, 1),
                fp32_ms_per_sweep=round(ms32, 4), fp32_A_eff_bytes=a_eff / 2,
                fp32_T_eff_GBs=round(a_eff / 2 / ms32 / 1e6, 1), sweep_over_iteration=round(ms32 / ms64, 3),
                device_bytes_fp64=bytes64, device_bytes_fp32_context=bytes32)


def bench_c1(n=127, threads=16, repeats=3):
    """2D viscous Stokes (Stokes2D.jl, divide form) at 127²: 1e-3 inclusion of
    radius 1 with ρg_y = -1 inside, free-slip walls (εbg = 0), Re and r at the
    1D dispersion optimum, converged to ‖res‖∞ < 1e-8 with a check every 200
    iterations: time to solution on the GPU and with the oracle on the host
    cores (same iterations, same result bits), each the best of `repeats`
    solves from a fresh set-up (the first GPU solve of a process also loads
    the kernel's code).  CFL is the script's 0.9/√2: at 0.95/√2 this problem
    diverges (NaN after ~2200 iterations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    from ptstokes import StokesModel
    from ptstokes.models import R_OPT, RE_OPT
    kw = dict(n=n, CFL=0.9 / math.sqrt(2), Re=RE_OPT, r=R_OPT, rhog_in=-1.0, ebg=0.0, norm_kind=1, nout=200,
              eps=1e-8)
    t_gpu = []
    for _ in range(repeats):
        m = StokesModel("Stokes2D", **kw).setup()
        m.ctx.sync()
        t0 = time.perf_counter()
        r = m.step()
        m.ctx.sync()
        t_gpu.append(time.perf_counter() - t0)
        pt_gpu = m.get("Pt")
        m.close()
    okw = dict(n=n, perf=0, ve=0, CFL=kw["CFL"], Re=RE_OPT, r=R_OPT, rhog_in=-1.0, ebg=0.0, norm_inf=1, nout=200,
               eps=1e-8, threads=threads)
    t_cpu = []
    for _ in range(repeats):
        o = Oracle("Stokes2D", **okw).setup()
        t0 = time.perf_counter()
        ro = o.solve()
        t_cpu.append(time.perf_counter() - t0)
    import numpy as np
    same = bool(np.array_equal(o.get("Pt"), pt_gpu))
    return dict(workload="2D viscous Stokes 127^2, divide form, inclusion with rho_g=-1, free slip, "
                         "||res||_inf < 1e-8 checked every 200 iterations",
                gpu_s=round(min(t_gpu), 4), gpu_s_by_solve=[round(t, 4) for t in t_gpu], gpu_iters=r["iters"],
                cpu_s=round(min(t_cpu), 4), cpu_s_by_solve=[round(t, 4) for t in t_cpu], cpu_iters=ro["ittot"],
                cpu_threads=threads, speedup=round(min(t_cpu) / min(t_gpu), 1), identical_result=same)


def bench_c5(n=2047, steps=200):
    """Per-iteration cost of the forward and adjoint PT sweeps of the 2D
    viscous inversion problem (log μ a seeded GRF) at n²."""
    import numpy as np
    from ptstokes.adjoint import Inversion
    ev = HipEvents()
    inv = Inversion(n=n, npts=256, eps=1e-8)  # set up with the seed-1234 field
    m, c = inv.model, inv.ctx
    fwd_ms, _ = time_iterations(m, steps, 10, ev)
    g = np.zeros((n + 1, n))
    g[inv.points, n - 1] = 1.0
    c.adjoint_set_source(g, None)
    c.adjoint_iterate(inv.params, 10)
    s = c.stream()
    a, b = ev.event(), ev.event()
    ev.record(a, s)
    c.adjoint_iterate(inv.params, steps)
    ev.record(b, s)
    adj_ms = ev.elapsed_ms(a, b)
    inv.close()
    a_eff = 15.0 * n * n * 8
    f_it, a_it = fwd_ms / steps, adj_ms / steps
    return dict(workload=f"2D viscous Stokes {n}^2, log mu GRF (seed 1234), forward and adjoint PT iterations",
                forward_ms_per_it=round(f_it, 4), forward_T_eff_GBs=round(a_eff / f_it / 1e6, 1),
                adjoint_ms_per_it=round(a_it, 4), adjoint_T_eff_GBs=round(a_eff / a_it / 1e6, 1),
                A_eff_bytes=a_eff, note="T_eff with A_eff = 15 arrays (6 state read+written, 3 read-only); the "
                                       "single-pass adjoint sweep reads the 6 adjoint fields, mu, Gd, dtau_Rho and "
                                       "the dJ/dVx source, and writes 6")


def host_cpu():
    """Threads the CPU baseline uses (the CPUs this process may run on, or
    OMP_NUM_THREADS where that is set lower: the GPU box sets it to its CPU
    share), the physical cores among them and the CPU model."""
    cpus = sorted(os.sched_getaffinity(0))
    threads = len(cpus)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model, cores = "unknown", set()
    try:
        blocks = open("/proc/cpuinfo").read().strip().split("\n\n")
        for b in blocks:
            kv = dict(line.split(":", 1) for line in b.splitlines() if ":" in line)
            kv = {k.strip(): v.strip() for k, v in kv.items()}
            if int(kv.get("processor", -1)) in cpus:
                model = kv.get("model name", model)
                cores.add((kv.get("physical id", "0"), kv.get("core id", kv.get("processor"))))
    except (OSError, ValueError):
        pass
    return dict(threads=threads, affinity_cpus=len(cpus), physical_cores=len(cores) or None, model=model,
                omp_num_threads=omp or None)


def _timed_best(o, warm, timed, repeats):
    """The reference's protocol (Stokes2D_ve_perf.jl:148,172-176; runme_perf.sh:48):
    the clock starts after the first `warm` iterations; the best of `repeats`
    runs of `timed` iterations."""
    o.iterate(warm)
    best = math.inf
    for _ in range(repeats):
        t0 = time.perf_counter()
        o.iterate(timed)
        best = min(best, (time.perf_counter() - t0) / timed)
    return best


def cpu_baseline_2d(n, threads, warm=10, timed=91, repeats=5):
    """The oracle (the reference's CPU kernels restated in C, OpenMP over the
    stencil loops) on 2D VE n², timed as the reference's perf script times
    itself: iterations 11..101 of one physical step, best of 5 repeats."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    o = Oracle("Stokes2D_ve_perf", n=n, dt=0.1, threads=threads).setup()
    o.assign_tau()
    t = _timed_best(o, warm, timed, repeats)
    a_eff = 18.0 * n * n * 8
    return dict(value=a_eff / t / 1e9, unit="GB/s", cores=threads, kind="port", ms_per_it=t * 1e3,
                sample=f"2D VE {n}x{n} FP64 (perf-script kernels, {threads} OpenMP threads): iterations "
                       f"{warm + 1}..{warm + timed} of one step, best of {repeats} repeats; ms/it {t * 1e3:.1f}")


def cpu_baseline_3d(n, threads, warm=10, timed=10, repeats=3):
    """The oracle on the 3D VE block: the clock starts at iteration 11 as in
    the reference's perf script; a bounded sample (10 iterations, best of 3)
    instead of its 190 (one iteration takes a quarter second on the host)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    t0 = time.perf_counter()
    o = Oracle("Stokes3D_ve_perf", n=n, threads=threads, nsmooth=1).setup()
    t_setup = time.perf_counter() - t0
    o.assign_tau()
    t = _timed_best(o, warm, timed, repeats)
    a_eff = 29.0 * n ** 3 * 8
    return dict(value=a_eff / t / 1e9, unit="GB/s", cores=threads, kind="port", ms_per_it=t * 1e3,
                sample=f"3D VE {n}^3 FP64 (perf-script kernels, {threads} OpenMP threads): iterations "
                       f"{warm + 1}..{warm + timed} of one step, best of {repeats} repeats, after the set-up "
                       f"({t_setup:.0f} s, 1 smoothing pass); ms/it {t * 1e3:.1f}")


# The kernel sources each HBM-traffic record was measured with: a record only
# describes a build of the same sources.
TRAFFIC_SOURCES = {"k3_iter2": ["pts_iter3d.hip", "pts_internal.hpp"], "k2_iter": ["pts_step2d.hip", "pts_internal.hpp"]}


def sources_sha256(kernel):
    import hashlib
    h = hashlib.sha256()
    for f in TRAFFIC_SOURCES[kernel]:
        with open(os.path.join(ROOT, "pseudotransientstokes.jl_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def load_traffic(kernel, n):
    """Measured HBM bytes per launch (profiles/hbm_traffic.json) and a note;
    None when the record is missing or was taken with other kernel sources."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(p) as f:
            rec = json.load(f).get("kernels", {}).get(f"{kernel}_{n}")
    except (OSError, ValueError):
        rec = None
    if not rec:
        return None, f"no measured traffic for {kernel} at n={n}"
    if rec.get("sources_sha256") != sources_sha256(kernel):
        return None, f"traffic of {kernel} at n={n} was measured with other kernel sources (stale)"
    return rec, rec.get("profile", "profiles/hbm_traffic.json")


def record_traffic(rd_dir, wr_dir, sizes):
    """Write profiles/hbm_traffic.json from the counter CSVs of two rocprofv3
    --pmc runs of bench.py (TCC_EA0_RDREQ_sum, _128B, _64B, _32B; then
    TCC_EA0_WRREQ_sum, _64B): per kernel, the medians over the dispatches of
    its most dispatched instantiation (the one the timed region runs), with the
    hash of the kernel's sources.  sizes: {"k3_iter2": n3d, "k2_iter": n2d}."""
    import csv
    import glob
    import statistics

    def medians(d):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise FileNotFoundError(f"no counter_collection.csv under {d}")
        per = {}
        for f in files:
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    per.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(
                        float(r["Counter_Value"]))
        out = {}
        for kern in sizes:
            names = [k for k in per if f"{kern}<" in k]
            if not names:
                continue
            name = max(names, key=lambda k: max(len(v) for v in per[k].values()))
            out[kern] = {c: statistics.median(v) for c, v in per[name].items()}
        return out

    rd, wr = medians(rd_dir), medians(wr_dir)
    p = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(p) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    table["note"] = ("HBM bytes per launch of the single-pass kernels, medians over the dispatches of one bench run: "
                     "read = 128*TCC_EA0_RDREQ_128B + 64*TCC_EA0_RDREQ_64B + 32*TCC_EA0_RDREQ_32B, write = "
                     "64*TCC_EA0_WRREQ_64B + 32*(TCC_EA0_WRREQ - TCC_EA0_WRREQ_64B); sources_sha256: the kernel's "
                     "source files the record was measured with (bench.py reports the record only for the same "
                     "sources); written by bench.py --record-traffic")
    kernels = table.setdefault("kernels", {})
    for kern, n in sizes.items():
        if kern not in rd or kern not in wr:
            continue
        r, w = rd[kern], wr[kern]
        rec = dict(rdreq=int(r["TCC_EA0_RDREQ_sum"]), rdreq_128b=int(r["TCC_EA0_RDREQ_128B_sum"]),
                   rdreq_64b=int(r["TCC_EA0_RDREQ_64B_sum"]), rdreq_32b=int(r["TCC_EA0_RDREQ_32B_sum"]),
                   wrreq=int(w["TCC_EA0_WRREQ_sum"]), wrreq_64b=int(w["TCC_EA0_WRREQ_64B_sum"]))
        rec["read_bytes"] = 128 * rec["rdreq_128b"] + 64 * rec["rdreq_64b"] + 32 * rec["rdreq_32b"]
        rec["write_bytes"] = 64 * rec["wrreq_64b"] + 32 * (rec["wrreq"] - rec["wrreq_64b"])
        rec["bytes"] = rec["read_bytes"] + rec["write_bytes"]
        rec["sources_sha256"] = sources_sha256(kern)
        rec["profile"] = "profiles/pt_iteration_kernels_hbm_traffic.md"
        kernels[f"{kern}_{n}"] = rec
        print(f"# {kern} n={n}: read {rec['read_bytes'] / 1e9:.3f} GB, write {rec['write_bytes'] / 1e9:.3f} GB",
              file=sys.stderr)
    with open(p, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n2d", type=int, default=8191)
    ap.add_argument("--n3d", type=int, default=511)
    ap.add_argument("--contexts2d", type=int, default=3, help="2D: contexts timed one after another (median reported)")
    ap.add_argument("--contexts3d", type=int, default=3, help="3D: fresh allocations timed (median reported)")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D VE line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n-tts", type=int, default=127, help="time-to-tolerance grid (0: skip)")
    ap.add_argument("--no-extra", action="store_true", help="skip the 127^2 and 2047^2 extra lines")
    ap.add_argument("--c3-solve-steps", type=int, default=0,
                    help="also run that many converged physical steps of the 3D VE block (minutes)")
    ap.add_argument("--c4-kernels", type=int, default=0,
                    help="also time the FP64 iteration and FP32 sweep of the 3D random-viscosity problem at this n "
                         "(1023: ~225 GB of device memory)")
    ap.add_argument("--c5-floor", type=int, nargs="*", default=[],
                    help="also follow the 2D inversion's FP64 forward err (RMS and max-norm) to its floor at these n")
    ap.add_argument("--c5-floor-s", type=float, default=300.0, help="time limit of each --c5-floor run (s)")
    ap.add_argument("--c5-grad", type=int, default=0,
                    help="also time the 2D inversion's forward solve, adjoint solve and gradient at this n")
    ap.add_argument("--c2-solve-steps", type=int, default=0,
                    help="also run that many converged physical steps of 2D VE 8191^2 (minutes each)")
    ap.add_argument("--tts-no-graph", action="store_true",
                    help="FP64 time-to-tolerance without HIP graphs (for tracing tools)")
    ap.add_argument("--record-traffic", nargs=2, metavar=("RD_DIR", "WR_DIR"),
                    help="write profiles/hbm_traffic.json from two rocprofv3 --pmc runs of bench.py and exit (no GPU)")
    args = ap.parse_args()
    if args.record_traffic:
        record_traffic(args.record_traffic[0], args.record_traffic[1], {"k3_iter2": args.n3d, "k2_iter": args.n2d})
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU path is a design, not a build, in this repository)", file=sys.stderr)
        return 2

    ev = HipEvents()
    host = host_cpu()
    # headline: the 3D VE block
    r3 = bench_3d(args.n3d, args.steps, args.warmup, ev, args.contexts3d)
    print(f"# 3D VE {args.n3d}^3 placements tried (ms): {r3['placement_ms_by_context']}", file=sys.stderr)
    print(f"# 3D VE {args.n3d}^3: {r3['ms_per_it']:.4f} ms/it, T_eff {r3['T_eff_GBs']:.1f} GB/s", file=sys.stderr,
          flush=True)
    extra = {"host_cpu": host}
    traffic3, traffic3_note = load_traffic("k3_iter2", args.n3d)
    extra["3d_ve"] = dict(ms_per_it_by_context=r3["ms_per_it_by_context"],
                          placement_ms_by_context=r3["placement_ms_by_context"],
                          placement_s_by_context=r3["placement_s_by_context"], device_bytes=r3["device_bytes"],
                          wall_ms_per_it=r3["wall_ms_per_it"], traffic_record=traffic3_note)
    if traffic3:
        extra["3d_ve"].update(hbm_read_bytes=traffic3.get("read_bytes"), hbm_write_bytes=traffic3.get("write_bytes"),
                              achieved_hbm_GBs=round(traffic3["bytes"] / (r3["ms_per_it"] / 1e3) / 1e9, 1))
    if not args.no_2d:
        r2 = bench_2d(args.n2d, args.steps, args.warmup, ev, args.contexts2d)
        print(f"# 2D VE {args.n2d}^2 placements tried (ms): {r2['placement_ms_by_context']} "
              f"({r2['placement_s_by_context']} s); contexts {r2['ms_per_it_by_context']} ms/it", file=sys.stderr)
        print(f"# 2D VE {args.n2d}^2: {r2['ms_per_it']:.4f} ms/it, T_eff {r2['T_eff_GBs']:.1f} GB/s",
              file=sys.stderr, flush=True)
        r2["frac_of_peak"] = r2["T_eff_GBs"] / HBM_PEAK_GBS
        r2["vs_a100"] = round(r2["T_eff_GBs"] / A100_2D_VE_8191_GBS, 3)
        r2["workload"] = (f"2D Maxwell visco-elastic Stokes PT iteration, nx=ny={args.n2d}, FP64, G=eta=1, dt=0.1, "
                          "pure shear, 1e-3 inclusion r=1 (single-pass kernel)")
        t2, t2_note = load_traffic("k2_iter", args.n2d)
        r2["traffic_record"] = t2_note
        if t2:
            r2["hbm_bytes"] = t2["bytes"]
            r2["achieved_hbm_GBs"] = round(t2["bytes"] / (r2["ms_per_it"] / 1e3) / 1e9, 1)
        extra["2d_ve"] = r2
    if args.n_tts > 0:
        tts = bench_tts(args.n_tts, use_graph=not args.tts_no_graph)
        extra["tts_fp32_refined_vs_fp64"] = tts
        print(f"# TTS {args.n_tts}^3: FP64 {tts['fp64_s']} s, FP32+DC {tts['dc_s']} s "
              f"(ratio {tts['time_ratio_dc_over_fp64']})", file=sys.stderr, flush=True)
    if not args.no_extra:
        c5 = bench_c5()
        extra["c5_forward_adjoint_2047"] = c5
        print(f"# 2047^2 forward {c5['forward_ms_per_it']} ms/it, adjoint {c5['adjoint_ms_per_it']} ms/it",
              file=sys.stderr, flush=True)
        if not args.no_cpu:
            c1 = bench_c1(threads=host["threads"])
            extra["c1_viscous_127_time_to_solution"] = c1
            print(f"# 127^2 viscous: GPU {c1['gpu_s']} s, CPU {c1['cpu_s']} s ({c1['gpu_iters']} iterations)",
                  file=sys.stderr, flush=True)
    if args.c4_kernels > 0:
        extra["c4_kernels"] = bench_c4_kernels(args.c4_kernels, ev)
    if args.c5_floor:
        extra["c5_forward_err_floor"] = {str(n): bench_c5_floor(n, max_s=args.c5_floor_s) for n in args.c5_floor}
    if args.c5_grad > 0:
        extra["c5_inversion_gradient"] = bench_c5_grad(args.c5_grad)
    if args.c2_solve_steps > 0:
        c2s = bench_c2_solve(args.n2d, args.c2_solve_steps)
        extra["c2_2d_ve_time_to_solution"] = c2s
        print(f"# 2D VE {args.n2d}^2 time to solution: {c2s['iterations']} iterations in {c2s['wall_s']} s",
              file=sys.stderr, flush=True)
    if args.c3_solve_steps > 0:
        c3s = bench_c3_solve(args.n3d, args.c3_solve_steps)
        extra["c3_3d_ve_time_to_solution"] = c3s
        print(f"# 3D VE {args.n3d}^3 time to solution: {c3s['iterations']} iterations in {c3s['wall_s']} s",
              file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_3d(args.n3d, host["threads"])
        extra["3d_ve"]["gpu_over_cpu"] = round(r3["T_eff_GBs"] / cpu["value"], 1)
        print(f"# CPU 3D: {cpu['value']:.2f} GB/s ({cpu['sample']})", file=sys.stderr, flush=True)
        if not args.no_2d:
            c2 = cpu_baseline_2d(args.n2d, host["threads"])
            extra["2d_ve"]["cpu_baseline"] = c2
            extra["2d_ve"]["gpu_over_cpu"] = round(extra["2d_ve"]["T_eff_GBs"] / c2["value"], 1)
            print(f"# CPU 2D: {c2['value']:.2f} GB/s ({c2['sample']})", file=sys.stderr, flush=True)

    t_eff = r3["T_eff_GBs"]
    line = {
        "metric": METRIC,
        "value": round(t_eff, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r3["ms_per_it"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(t_eff / A100_3D_VE_511_GBS, 3),
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Maxwell visco-elastic Stokes PT iteration, nx=ny=nz={args.n3d}, FP64, G=1, dt=1, "
                        "x-z pure shear, 1e-3 sphere r=1 (single-pass kernel k3_iter2)",
            "A_eff_bytes_per_step": r3["A_eff_bytes"],
            "T_eff_frac_of_hbm_peak": round(t_eff / HBM_PEAK_GBS, 4),
            "baseline": "A100 T_eff 732.2 GB/s, output/out_Stokes3D_ve_perf_ampere.txt:15",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(t_eff, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(t_eff / HBM_PEAK_GBS, 4),
            "traffic": traffic3["bytes"] if traffic3 else None,
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
