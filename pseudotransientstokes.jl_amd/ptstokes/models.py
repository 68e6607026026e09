"""Drivers that replay the reference scripts on the GPU.

``StokesModel(preset, **overrides)`` reproduces the setup pipeline, time loop
and PT loop of one reference script (scripts/Stokes2D.jl, Stokes2D_ve.jl,
Stokes2D_ve_perf.jl, stokes_3D/Stokes3D_ve_perf.jl, Stokes3D_ve_multixpu.jl
and _multixpu_perf.jl, on one GPU) through libpts.  The presets copy each
script's hard-coded physics and numerics; every constant can be overridden,
which is how the larger measured configurations (body force, dt, nt, CFL,
Re, r, max-norm stop) are expressed.
"""
import math
import time

import numpy as np

from .context import Context

_COMMON = dict(lx=10.0, ly=10.0, lz=10.0, mus0=1.0, musi=1e-3, rad2=1.0, ebg=1.0, eps=1e-8,
               nsmooth=10, norm_kind=0, rhog_in=0.0, rhog_out=0.0, ic_igg=False, grf_seed=None,
               grf_sigma=1.0, grf_corr=0.5, grf_clip=2.0, grf_nmodes=1024)

PRESETS = {
    # scripts/Stokes2D.jl:74-89
    "Stokes2D": dict(_COMMON, ndim=2, ve=False, fd="div", nt=1, iter_max=100000, nout=500,
                     CFL=0.9 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0, n=255),
    # scripts/Stokes2D_ve.jl:89-108 (dt = μs0/(G*ξ) = 1)
    "Stokes2D_ve": dict(_COMMON, ndim=2, ve=True, fd="div", nt=5, iter_max=200000, nout=200,
                        CFL=0.8 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0, n=255),
    # scripts/Stokes2D_ve_perf.jl:82-102 (nt=1, iterMax=100, nout=2000)
    "Stokes2D_ve_perf": dict(_COMMON, ndim=2, ve=True, fd="mul", nt=1, iter_max=100, nout=2000,
                             CFL=0.8 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0, n=255),
    # scripts/stokes_3D/Stokes3D_ve_perf.jl:118-138
    "Stokes3D_ve_perf": dict(_COMMON, ndim=3, ve=True, fd="mul", nt=1, iter_max=200, nout=2000,
                             CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0, G=1.0, dt=1.0, n=63),
    # scripts/stokes_3D/Stokes3D_ve_multixpu.jl:120-141 (one rank)
    "Stokes3D_ve_multixpu": dict(_COMMON, ndim=3, ve=True, fd="div", ic_igg=True, nt=5, iter_max=200000,
                                 nout=200, CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0, G=1.0, dt=1.0,
                                 n=63),
    # scripts/stokes_3D/Stokes3D_ve_multixpu_perf.jl (one rank)
    "Stokes3D_ve_multixpu_perf": dict(_COMMON, ndim=3, ve=True, fd="mul", ic_igg=True, nt=1, iter_max=200,
                                      nout=2000, CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0, G=1.0,
                                      dt=1.0, n=63),
}

# Random-viscosity configurations (no reference script), viscous, max-norm
# stop: log10 μ is a seeded Gaussian random field (σ = 1, correlation length
# 0.5, clipped to ±2).  3D: plus the 1e-3 sphere carrying ρg_z = -1, εbg = 0;
# the sphere inside the field reaches a 1e5 local contrast, for which the
# reference's 10 smoothing passes leave the PT iteration unstable at fine
# grids, so 20 passes are used.  2D (inversion): the field alone under pure
# shear.
PRESETS["Stokes3D_grf"] = dict(_COMMON, ndim=3, ve=False, fd="mul", nt=1, iter_max=10 ** 6, nout=500,
                               CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0, G=1.0, dt=1.0, n=63,
                               ebg=0.0, rhog_in=-1.0, norm_kind=1, grf_seed=1234, nsmooth=20)
PRESETS["Stokes2D_grf"] = dict(_COMMON, ndim=2, ve=False, fd="mul", nt=1, iter_max=10 ** 6, nout=500,
                               CFL=0.9 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0, n=127,
                               norm_kind=1, grf_seed=1234, rad2=0.0)
# Visco-elastic inversion: the same field, G = 1, dt = η/G/10 with η = 1,
# 10 physical steps (Stokes2D_ve_perf.jl numerics).
PRESETS["Stokes2D_ve_grf"] = dict(_COMMON, ndim=2, ve=True, fd="mul", nt=10, iter_max=10 ** 6, nout=500,
                                  CFL=0.8 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=0.1, n=127,
                                  norm_kind=1, grf_seed=1234, rad2=0.0)

# 1D dispersion optimum (dispersion_analysis/dispersion_analysis_stokes1D.ipynb
# cell 5 output): r = 1/4, Re = 9*sqrt(3)*pi/4
R_OPT = 0.25
RE_OPT = 9.0 * math.sqrt(3.0) * math.pi / 4.0


def julia_linrange(a, b, n):
    """The elements of Julia's LinRange(a, b, n): (1-t)·a + t·b with
    t = (i-1)/(n-1), which differs from numpy.linspace in the last bits."""
    if n == 1:
        return np.array([a], dtype=np.float64)
    t = np.arange(n, dtype=np.float64) / (n - 1)
    return (1.0 - t) * a + t * b


def config(preset, **overrides):
    d = dict(PRESETS[preset])
    d.update(overrides)
    n = d.pop("n")
    d.setdefault("nx", n)
    d.setdefault("ny", n)
    d.setdefault("nz", n if d["ndim"] == 3 else 1)
    return d


class StokesModel:
    """One run of a reference script (setup + time loop + PT loop) on the GPU."""

    def __init__(self, preset="Stokes2D_ve_perf", device=0, **overrides):
        self.preset = preset
        self.p = p = config(preset, **overrides)
        body_force = p["rhog_in"] != 0.0 or p["rhog_out"] != 0.0
        self.ctx = Context(p["ndim"], p["nx"], p["ny"], p["nz"], ve=p["ve"], fd_mode=p["fd"],
                           body_force=body_force, device=device)
        # Derived numerics (Stokes2D_ve.jl:106-108, Stokes3D_ve_perf.jl:135-138)
        self.dx = p["lx"] / p["nx"]
        self.dy = p["ly"] / p["ny"]
        self.dz = p["lz"] / p["nz"] if p["ndim"] == 3 else 1.0
        if p["ndim"] == 2:
            self.max_l = max(p["lx"], p["ly"])
            self.Vpdtau = min(self.dx, self.dy) * p["CFL"]
        else:
            self.max_l = max(p["lx"], p["ly"], p["lz"])
            self.Vpdtau = min(self.dx, self.dy, self.dz) * p["CFL"]
        self.ittot = 0
        self.iters_per_step = []
        self.err_history = []
        self.timing = {}

    # coordinates of the cell centres (xc, yc[, zc]) as the scripts return them
    # (xc = LinRange(dx/2, lx-dx/2, nx), Stokes2D_ve.jl:109)
    def coords(self):
        p = self.p
        xc = julia_linrange(self.dx / 2, p["lx"] - self.dx / 2, p["nx"])
        yc = julia_linrange(self.dy / 2, p["ly"] - self.dy / 2, p["ny"])
        if p["ndim"] == 2:
            return xc, yc
        return xc, yc, julia_linrange(self.dz / 2, p["lz"] - self.dz / 2, p["nz"])

    def setup(self):
        """ICs, 10x smooth! with swap, Musτ = maxloc(Mus) + BCs, iteration
        parameters (Stokes2D_ve.jl:131-147, Stokes3D_ve_perf.jl:167-192)."""
        p, c = self.p, self.ctx
        c.init_inclusion(p["lx"], p["ly"], p["lz"], p["mus0"], p["musi"], p["rad2"], p["ebg"], p["rhog_in"],
                         p["rhog_out"], p["ic_igg"])
        if p["grf_seed"] is not None:
            from . import grf
            md = grf.modes(p["grf_seed"], p["ndim"], p["grf_corr"], p["grf_nmodes"])
            c.init_grf_viscosity(md, p["grf_sigma"], -p["grf_clip"], p["grf_clip"], p["lx"], p["ly"], p["lz"],
                                 p["musi"], p["rad2"])
        c.copy("Mus2", "Mus")
        for _ in range(p["nsmooth"]):
            c.smooth("Mus2", "Mus", 1.0)
            c.swap("Mus", "Mus2")
        c.copy("Mustau", "Mus")
        c.maxloc("Mustau", "Mus")
        c.bc_x("Mustau")
        c.bc_y("Mustau")
        if p["ndim"] == 3:
            c.bc_z("Mustau")
        c.iter_params(self.Vpdtau, p["G"], p["dt"], p["Re"], p["r"], self.max_l)
        return self

    def pt_params(self):
        p = self.p
        return Context.pt_params(p["r"], p["G"], p["dt"], self.dx, self.dy, self.dz, p["lx"])

    def iterate(self, n):
        """n PT iterations on the device (single-pass kernel), no checks."""
        self.ctx.pt_iterate(self.pt_params(), n)

    def iterate_per_kernel(self, n):
        """n PT iterations through the per-kernel reference API, in the exact
        call order of the script's loop body."""
        p, c = self.p, self.ctx
        dx, dy, dz = self.dx, self.dy, self.dz
        for _ in range(int(n)):
            if p["fd"] == "mul":
                _dx, _dy, _dz = 1.0 / dx, 1.0 / dy, 1.0 / dz
                c.compute_Pt_tau(p["r"], p["G"], p["dt"], _dx, _dy, _dz)
                c.compute_V_perf(_dx, _dy, _dz)
            else:
                c.compute_P(p["r"], dx, dy, dz)
                c.compute_tau(p["G"], p["dt"], dx, dy, dz)
                c.compute_dV(dx, dy, dz)
                c.compute_V()
            if p["ndim"] == 2:
                c.bc_y("Vx")
                c.bc_x("Vy")
            else:
                c.bc_x("Vy")
                c.bc_x("Vz")
                c.bc_y("Vx")
                c.bc_y("Vz")
                c.bc_z("Vx")
                c.bc_z("Vy")

    def step(self, use_graph=True):
        """One physical time step: assign_τ! then the PT loop to err < ε."""
        p = self.p
        if p["ve"]:
            self.ctx.assign_tau()
        res = self.ctx.pt_solve(self.pt_params(), p["eps"], p["nout"], p["iter_max"], p["norm_kind"],
                                use_graph=use_graph)
        self.iters_per_step.append(res["iters"])
        self.ittot += res["iters"]
        self.err_history.append(res["err"])
        return res

    def solve_dc(self, rtol_inner=1e-4, nout_inner=200, max_outer=20):
        """One steady solve by FP32 PT sweeps inside FP64 defect correction,
        stopped on the same FP64 norm and eps as the FP64 PT loop."""
        p = self.p
        t0 = time.perf_counter()
        res = self.ctx.dc_solve(self.pt_params(), p["eps"], p["norm_kind"], max_outer, rtol_inner, nout_inner)
        self.timing["dc_s"] = time.perf_counter() - t0
        return res

    def run(self, use_graph=True):
        """The full time loop (for it = 1:nt ...); returns (xc, yc[, zc], Pt)."""
        t0 = time.perf_counter()
        for _ in range(self.p["nt"]):
            self.step(use_graph=use_graph)
        self.ctx.sync()
        self.timing["run_s"] = time.perf_counter() - t0
        return (*self.coords(), self.ctx.download("Pt"))

    def get(self, name):
        return self.ctx.download(name)

    def close(self):
        self.ctx.close()
