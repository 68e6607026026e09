"""The reference scripts' entry points on the GPU.

Each function replays one script end to end (setup, time loop, PT loop,
performance line, optional outputs) through libpts, reading the scripts'
environment knobs (README.md:71-79):

    USE_RETURN   return (xc, yc[, zc], Pt) instead of None
    DO_SAVE      append the script's result line to output/out_<name>.txt
    DO_SAVE_VIZ  write the script's .mat file to out_visu/
    NX, NY, NZ   grid size (defaults 255 in 2D, 63 in 3D)

USE_GPU is implied (the library is GPU-only) and DO_VIZ (plotting) is not
provided.  Output folders are relative to PTS_OUT_DIR (default: the current
directory), where the reference writes ../output and ../out_visu relative
to its scripts folder.

Run one from the shell as ``python -m ptstokes.scripts Stokes2D_ve``.
"""
import math
import os
import sys
import time
from decimal import Decimal

from .models import StokesModel


def _flag(env, key, default):
    v = env.get(key)
    if v is None:
        return default
    return v.strip().lower() in ("1", "true", "yes")


def _int(env, key, default):
    v = env.get(key)
    return default if v is None else int(v)


def julia_float(x):
    """A Float64 as Julia's print/string interpolation writes it: shortest
    round-trip digits, plain for 1e-5 < |x| < 1e6 with at least one decimal
    ("100.0"), otherwise "d.ddde±x" ("1.0e-5", "2.5e7")."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign, digits, exp = Decimal(repr(x)).as_tuple()
    ds = "".join(map(str, digits)).rstrip("0") or "0"
    e10 = len(digits) + exp - 1  # exponent of the leading digit
    s = "-" if sign else ""
    if -5 < e10 < 6:
        if e10 >= 0:
            ip = ds[: e10 + 1].ljust(e10 + 1, "0")
            fp = ds[e10 + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-e10 - 1) + ds
        return f"{s}{ip}.{fp}"
    return f"{s}{ds[0]}.{ds[1:] or '0'}e{e10}"


def _sig(x, n):
    """round(x, sigdigits=n)."""
    return float(f"{x:.{n}g}") if x == x and x != 0 else x


class _Run:
    def __init__(self, name, preset, ndim, env, default_n):
        self.name, self.env = name, env
        self.use_return = _flag(env, "USE_RETURN", False)
        self.do_save = _flag(env, "DO_SAVE", False)
        self.do_save_viz = _flag(env, "DO_SAVE_VIZ", False)
        nx = _int(env, "NX", default_n)
        ny = _int(env, "NY", default_n)
        kw = dict(nx=nx, ny=ny)
        if ndim == 3:
            kw["nz"] = _int(env, "NZ", default_n)
        self.model = StokesModel(preset, **kw)
        self.p = self.model.p
        self.out_dir = env.get("PTS_OUT_DIR", ".")

    def save_line(self, fname, fields):
        d = os.path.join(self.out_dir, "output")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, fname), "a") as io:
            io.write(" ".join(julia_float(v) if isinstance(v, float) else str(v) for v in fields) + "\n")

    def save_mat(self, fname, arrays):
        from scipy.io import savemat
        d = os.path.join(self.out_dir, "out_visu")
        os.makedirs(d, exist_ok=True)
        savemat(os.path.join(d, fname), arrays, do_compression=True)

    def converged_loop(self):
        """setup, then the time loop; returns (ittot, wall time from the
        start of the loop).  The reference starts its clock at iteration 11
        of step 1; the device-resident loop is timed from its start, and
        t_it = wtime/(ittot-10) as in the scripts."""
        m = self.model
        m.setup()
        m.ctx.sync()
        t0 = time.perf_counter()
        for _ in range(self.p["nt"]):
            m.step()
        m.ctx.sync()
        return m.ittot, time.perf_counter() - t0

    def fixed_loop(self):
        """The perf scripts (nt = 1, no check before iterMax): 10 untimed
        iterations, then the clock (Stokes2D_ve_perf.jl:148), then the rest of
        the iterMax+1 iterations."""
        m, p = self.model, self.p
        m.setup()
        if p["ve"]:
            m.ctx.assign_tau()
        total = p["iter_max"] + 1
        m.iterate(10)
        m.ctx.sync()
        t0 = time.perf_counter()
        m.iterate(total - 10)
        m.ctx.sync()
        m.ittot = total
        return total, time.perf_counter() - t0

    def ret(self, *coords):
        pt = self.model.get("Pt")
        out = (*coords, pt) if self.use_return else None
        return out


def Stokes2D(env=None):
    """scripts/Stokes2D.jl: 2D viscous, 4-kernel form, converged."""
    r = _Run("Stokes2D", "Stokes2D", 2, os.environ if env is None else env, 255)
    ittot, wtime = r.converged_loop()
    p, m = r.p, r.model
    a_eff = (3 * 2) / 1e9 * p["nx"] * p["ny"] * 8  # Stokes2D.jl:150
    t_eff = a_eff / (wtime / (ittot - 10))
    err = m.err_history[-1]
    print("Total steps = %d, err = %1.3e, time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, err, wtime,
                                                                                      _sig(t_eff, 2)))
    if r.do_save:
        r.save_line("out_Stokes2D.txt", [p["nx"], p["ny"], ittot])  # Stokes2D.jl:164-166
    out = r.ret(*m.coords())
    m.close()
    return out


def _viz2d(r, fname):
    m = r.model
    r.save_mat(fname, {"Pt_2D": m.get("Pt"), "Mus_2D": m.get("Mus"), "Txy_2D": m.get("txy"), "Vy_2D": m.get("Vy"),
                       "dx_2D": m.dx, "dy_2D": m.dy})


def Stokes2D_ve(env=None):
    """scripts/Stokes2D_ve.jl: 2D Maxwell VE, 4-kernel form, 5 converged steps."""
    r = _Run("Stokes2D_ve", "Stokes2D_ve", 2, os.environ if env is None else env, 255)
    ittot, wtime = r.converged_loop()
    p, m = r.p, r.model
    a_eff = (6 * 2 + 1) / 1e9 * p["nx"] * p["ny"] * 8  # Stokes2D_ve.jl:181
    t_eff = a_eff / (wtime / (ittot - 10))
    print("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, p["nt"], wtime,
                                                                                     _sig(t_eff, 2)))
    if r.do_save:
        r.save_line("out_Stokes2D_ve.txt", [p["nx"], p["ny"], ittot, p["nt"]])  # Stokes2D_ve.jl:197-199
    if r.do_save_viz:
        _viz2d(r, "Stokes_2D_ve.mat")  # Stokes2D_ve.jl:203
    out = r.ret(*m.coords())
    m.close()
    return out


def Stokes2D_ve_perf(env=None):
    """scripts/Stokes2D_ve_perf.jl: 2D VE fused form, iterMax+1 = 101 iterations."""
    r = _Run("Stokes2D_ve_perf", "Stokes2D_ve_perf", 2, os.environ if env is None else env, 255)
    ittot, t_toc = r.fixed_loop()
    p, m = r.p, r.model
    a_eff = (6 * 2 + 3 * 1 + 3) / 1e9 * p["nx"] * p["ny"] * 8  # Stokes2D_ve_perf.jl:173
    t_it = t_toc / (ittot - 10)
    t_eff = a_eff / t_it
    print("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, p["nt"], t_toc,
                                                                                     _sig(t_eff, 3)))
    if r.do_save:  # Stokes2D_ve_perf.jl:189-191
        r.save_line("out_Stokes2D_ve_perf.txt", [p["nx"], p["ny"], ittot, t_toc, a_eff, t_it, t_eff])
    if r.do_save_viz:
        _viz2d(r, "Stokes_2D_ve_perf.mat")  # Stokes2D_ve_perf.jl:195
    out = r.ret(*m.coords())
    m.close()
    return out


def Stokes3D_ve_perf(env=None):
    """scripts/stokes_3D/Stokes3D_ve_perf.jl: 3D VE fused form, iterMax+1 = 201 iterations."""
    r = _Run("Stokes3D_ve_perf", "Stokes3D_ve_perf", 3, os.environ if env is None else env, 63)
    ittot, wtime = r.fixed_loop()
    p, m = r.p, r.model
    a_eff = (10 * 2 + 6 * 1 + 3) / 1e9 * p["nx"] * p["ny"] * p["nz"] * 8  # Stokes3D_ve_perf.jl:229
    wtime_it = wtime / (ittot - 10)
    t_eff = a_eff / wtime_it
    print("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, p["nt"], wtime,
                                                                                     _sig(t_eff, 3)))
    if r.do_save:  # Stokes3D_ve_perf.jl:246-248
        r.save_line("out_Stokes3D_ve_perf.txt", [1, p["nx"], p["ny"], p["nz"], ittot, wtime, a_eff, wtime_it, t_eff])
    if r.do_save_viz:  # Stokes3D_ve_perf.jl:252
        r.save_mat("Stokes_3D_ve_perf.mat", {"Pt_3D": m.get("Pt"), "Mus_3D": m.get("Mus"), "Txz_3D": m.get("txz"),
                                             "Vz_3D": m.get("Vz"), "dx_3D": m.dx, "dy_3D": m.dy, "dz_3D": m.dz})
    out = r.ret(*m.coords())
    m.close()
    return out


def _viz3d_multixpu(r, fname):
    """The inner / averaged arrays the multi-XPU scripts gather for output
    (Stokes3D_ve_multixpu.jl:25-28,267-271), on one rank."""
    m = r.model
    Pt, Vz, Mus, txz = m.get("Pt"), m.get("Vz"), m.get("Mus"), m.get("txz")
    inn = lambda A: A[1:-1, 1:-1, 1:-1]
    av_zi = lambda A: (A[1:-1, 1:-1, 1:-2] + A[1:-1, 1:-1, 2:-1]) * 0.5
    av_xza = lambda A: (A[:-1, :, :-1] + A[:-1, :, 1:] + A[1:, :, :-1] + A[1:, :, 1:]) * 0.25
    r.save_mat(fname, {"Pt_3D": inn(Pt), "Mus_3D": inn(Mus), "Txz_3D": av_xza(txz), "Vz_3D": av_zi(Vz),
                       "dx_3D": m.dx, "dy_3D": m.dy, "dz_3D": m.dz})


def Stokes3D_ve_multixpu(env=None):
    """scripts/stokes_3D/Stokes3D_ve_multixpu.jl on one rank: 3D VE, 4-kernel
    form, 5 converged steps."""
    r = _Run("Stokes3D_ve_multixpu", "Stokes3D_ve_multixpu", 3, os.environ if env is None else env, 63)
    ittot, wtime = r.converged_loop()
    p, m = r.p, r.model
    a_eff = (10 * 2 + 1) / 1e9 * p["nx"] * p["ny"] * p["nz"] * 8  # Stokes3D_ve_multixpu.jl:261
    t_eff = a_eff / (wtime / (ittot - 10))
    print("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, p["nt"], wtime,
                                                                                     _sig(t_eff, 2)))
    if r.do_save:  # Stokes3D_ve_multixpu.jl:285-287
        r.save_line("out_Stokes3D_ve.txt", [p["nx"], p["ny"], p["nz"], ittot, p["nt"]])
    if r.do_save_viz:
        _viz3d_multixpu(r, "Stokes_3D_ve.mat")
    out = r.ret(*m.coords())
    m.close()
    return out


def Stokes3D_ve_multixpu_perf(env=None):
    """scripts/stokes_3D/Stokes3D_ve_multixpu_perf.jl on one rank."""
    r = _Run("Stokes3D_ve_multixpu_perf", "Stokes3D_ve_multixpu_perf", 3, os.environ if env is None else env, 63)
    ittot, wtime = r.fixed_loop()
    p, m = r.p, r.model
    a_eff = (10 * 2 + 6 * 1 + 3) / 1e9 * p["nx"] * p["ny"] * p["nz"] * 8
    wtime_it = wtime / (ittot - 10)
    t_eff = a_eff / wtime_it
    print("Total iters = %d (%d steps), time = %1.3e sec (@ T_eff = %1.2f GB/s) " % (ittot, p["nt"], wtime,
                                                                                     _sig(t_eff, 3)))
    if r.do_save:  # Stokes3D_ve_multixpu_perf.jl:285-287 (nprocs = 1)
        r.save_line("out_Stokes3D_ve_xpu_perf.txt",
                    [1, p["nx"], p["ny"], p["nz"], ittot, wtime, a_eff, wtime_it, t_eff])
    if r.do_save_viz:
        _viz3d_multixpu(r, "Stokes_3D_ve_perf_xpu.mat")
    out = r.ret(*m.coords())
    m.close()
    return out


SCRIPTS = {f.__name__: f for f in (Stokes2D, Stokes2D_ve, Stokes2D_ve_perf, Stokes3D_ve_perf, Stokes3D_ve_multixpu,
                                   Stokes3D_ve_multixpu_perf)}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1 or argv[0] not in SCRIPTS:
        print("usage: python -m ptstokes.scripts {" + ",".join(SCRIPTS) + "}", file=sys.stderr)
        return 2
    SCRIPTS[argv[0]]()
    return 0


if __name__ == "__main__":
    sys.exit(main())
