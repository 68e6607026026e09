"""Per-kernel parity: every reference kernel on the GPU against the oracle,
bit for bit, on the reference set-up and on seeded random states."""
import zlib

import numpy as np
import pytest

from gpu_cases import SETUP, STATE, HIST, compare, pair, randomise

pytestmark = pytest.mark.gpu

CASES = [
    # (preset, overrides)
    ("Stokes2D", dict(n=37)),                                  # viscous, divide form
    ("Stokes2D_ve", dict(n=37)),                               # VE, divide form
    ("Stokes2D_ve_perf", dict(n=37)),                          # VE, multiply form
    ("Stokes2D", dict(n=37, rhog_in=-1.0)),                    # body force
    ("Stokes2D_ve_perf", dict(nx=70, ny=45, dt=0.1)),          # non-square, dt != 1
    ("Stokes3D_ve_multixpu", dict(n=13)),                      # 3D VE divide, IGG ICs
    ("Stokes3D_ve_perf", dict(nx=17, ny=11, nz=13)),           # 3D VE multiply
    ("Stokes3D_ve_perf", dict(n=11, ve=False, rhog_in=-1.0)),  # 3D viscous + body force
]

DIV_KERNELS = ["compute_P", "compute_tau", "compute_dV", "compute_V", "compute_Res"]
MUL_KERNELS = ["compute_Pt_tau", "compute_V_perf", "compute_Res_perf"]


def _ids():
    return [f"{p}-{'-'.join(f'{k}{v}' for k, v in kw.items())}" for p, kw in CASES]


@pytest.mark.parametrize("preset,kw", CASES, ids=_ids())
def test_setup_bit_exact(preset, kw):
    m, o = pair(preset, **kw)
    names = [n for n in SETUP if n != "rhog" or m.p["rhog_in"] != 0.0]
    assert compare(m, o, names + ["Mus2"]) == []
    m.close()


def _run_gpu(m, name):
    p, c = m.p, m.ctx
    dx, dy, dz = m.dx, m.dy, m.dz
    if name == "compute_P":
        c.compute_P(p["r"], dx, dy, dz)
    elif name == "compute_tau":
        c.compute_tau(p["G"], p["dt"], dx, dy, dz)
    elif name == "compute_dV":
        c.compute_dV(dx, dy, dz)
    elif name == "compute_V":
        c.compute_V()
    elif name == "compute_Res":
        c.compute_Res(dx, dy, dz)
    elif name == "compute_Pt_tau":
        c.compute_Pt_tau(p["r"], p["G"], p["dt"], 1.0 / dx, 1.0 / dy, 1.0 / dz)
    elif name == "compute_V_perf":
        c.compute_V_perf(1.0 / dx, 1.0 / dy, 1.0 / dz)
    elif name == "compute_Res_perf":
        c.compute_Res_perf(1.0 / dx, 1.0 / dy, 1.0 / dz)
    else:
        raise KeyError(name)


@pytest.mark.parametrize("preset,kw", CASES, ids=_ids())
@pytest.mark.parametrize("kernel", DIV_KERNELS + MUL_KERNELS)
def test_kernel_bit_exact_on_random_state(preset, kw, kernel):
    m, o = pair(preset, **kw)
    randomise(m, o, seed=zlib.crc32(f"{preset}/{kernel}".encode()))
    nd = m.p["ndim"]
    if kernel == "compute_V":  # consumes dV written by compute_dV
        _run_gpu(m, "compute_dV")
        o.kernel("compute_dV")
    _run_gpu(m, kernel)
    o.kernel(kernel)
    extra = {"compute_P": ["divV"], "compute_dV": ["dVx", "dVy", "dVz"] + (["Rx", "Ry"] if nd == 2 else []),
             "compute_Res": ["Rx", "Ry", "Rz"], "compute_Res_perf": ["divV", "Rx", "Ry", "Rz"]}.get(kernel, [])
    assert compare(m, o, STATE + extra) == []
    m.close()


@pytest.mark.parametrize("preset,kw", CASES[:1] + CASES[5:6], ids=["2d", "3d"])
def test_bc_kernels_bit_exact(preset, kw):
    m, o = pair(preset, **kw)
    randomise(m, o, seed=7)
    nd = m.p["ndim"]
    for f in (["Vx", "Vy"] if nd == 2 else ["Vx", "Vy", "Vz", "Pt"]):
        for b in (["bc_x", "bc_y"] if nd == 2 else ["bc_x", "bc_y", "bc_z"]):
            getattr(m.ctx, b)(f)
            o.kernel(b, f)
    assert compare(m, o, STATE) == []
    m.close()


def test_assign_tau_copies_stress():
    m, o = pair("Stokes3D_ve_perf", n=9)
    randomise(m, o, seed=3, hist=False)
    m.ctx.assign_tau()
    o.assign_tau()
    assert compare(m, o, HIST) == []
    m.close()


@pytest.mark.parametrize("preset,kw", [CASES[2], CASES[6], CASES[3]], ids=["2d-mul", "3d-mul", "2d-bf"])
def test_norms_of_stored_residuals(preset, kw):
    m, o = pair(preset, **kw)
    randomise(m, o, seed=11)
    _run_gpu(m, "compute_Res_perf")
    o.kernel("compute_Res_perf")
    n = m.ctx.norms_stored()
    err, comp = m.ctx.err_from_norms(n, m.p["lx"], 0)
    ref = o.check()
    assert err == pytest.approx(ref["err"], rel=1e-12)
    assert n.Vmin == ref["Vmin"] and n.Vmax == ref["Vmax"] and n.Pmin == ref["Pmin"] and n.Pmax == ref["Pmax"]
    m.close()
