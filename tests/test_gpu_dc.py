"""Mixed precision (FP32 PT sweeps inside FP64 defect correction) and the
random-field viscosity: the FP64 residual reaches eps and the solution agrees
with the oracle's FP64 PT solve of the same problem (per field, measured
1e-11 to 1e-10 at 63², 127² and 31³)."""
import numpy as np
import pytest

from gpu_cases import oracle_kwargs, valid
from oracle import Oracle
from ptstokes import StokesModel, grf
from _util import relmax

pytestmark = pytest.mark.gpu

FIELDS = ["Pt", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz"]


# base oracle preset of each GPU preset (the random-field presets have no
# reference script; their arrays come from the GPU set-up)
ORACLE_BASE = {"Stokes2D_grf": "Stokes2D", "Stokes2D": "Stokes2D", "Stokes3D_grf": "Stokes3D_ve_perf"}


def oracle_fp64(model, eps):
    """The oracle's FP64 PT solve of the model's problem: same scalars and
    stop norm, and the model's set-up arrays (the random-field μ after
    smoothing, Musτ, Gdτ, dτ_Rho, ρg and the initial velocities) uploaded
    with or_set after the oracle's own set-up."""
    kw = oracle_kwargs(model)
    kw["eps"] = eps
    kw["iter_max"] = 400000  # bounds the CPU time should the solve not converge
    o = Oracle(ORACLE_BASE[model.preset], **kw).setup()
    nd = model.p["ndim"]
    for f in ("Mus", "Mustau", "Gdtau", "dtau_Rho", "Vx", "Vy", "Vz", "rhog"):
        if not valid(nd, f) or (f == "rhog" and model.p["rhog_in"] == 0.0 and model.p["rhog_out"] == 0.0):
            continue
        o.set(f, model.get(f))
    return o


@pytest.mark.parametrize("preset,kw", [
    ("Stokes2D_grf", dict(n=63)),
    ("Stokes2D", dict(n=127, fd="mul", rhog_in=-1.0, norm_kind=1)),
    ("Stokes3D_grf", dict(n=31)),
])
def test_dc_matches_oracle_fp64_solve(preset, kw):
    # FP32 PT sweeps inside FP64 defect correction, stopped at the scaled
    # max-norm 1e-8, against the oracle's FP64 PT solve of the same problem
    # to 1e-11: per field max|a - b| / max|b|, P with its mean over the
    # interior cells removed (P is fixed up to a constant by the equations;
    # the two solvers reach different constants).
    dc = StokesModel(preset, eps=1e-8, **kw).setup()
    o = oracle_fp64(dc, eps=1e-11)
    ro = o.solve()
    assert ro["rc"] == 0 and ro["err"] <= 1e-11, ro
    res = dc.solve_dc()
    assert res["converged"] and res["err"] <= 1e-8, res
    # each correction cuts the FP64 residual by orders of magnitude, except
    # the last, whose inner stop only asks for err below eps
    eo = res["err_outer"]
    assert all(b < max(a * 1e-2, 1e-8) for a, b in zip(eo, eo[1:])), eo
    nd = dc.p["ndim"]
    errs = {}
    for f in FIELDS:
        if not valid(nd, f):
            continue
        a, b = dc.get(f), o.get(f)
        if f == "Pt":
            inner = tuple(slice(1, -1) for _ in range(nd))
            a, b = a[inner] - a[inner].mean(), b[inner] - b[inner].mean()
        errs[f] = relmax(a, b)
    print(f"\n{preset} {kw}: DC ({res['outer']} corrections, {res['inner_iters']} FP32 sweeps) vs oracle FP64 "
          f"({ro['ittot']} iterations): " + ", ".join(f"{f} {e:.2e}" for f, e in errs.items()), flush=True)
    for f, e in errs.items():
        assert e < 1e-9, (f, e)
    dc.close()


def test_grf_viscosity_matches_host_evaluation():
    n = 40
    m = StokesModel("Stokes2D_grf", n=n, nsmooth=0).setup()
    p = m.p
    md = grf.modes(p["grf_seed"], 2, p["grf_corr"], p["grf_nmodes"])
    xc, yc = m.coords()
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    f = np.clip(grf.evaluate(md, X, Y, sigma=p["grf_sigma"]), -2, 2)
    want = 10.0 ** f
    np.testing.assert_allclose(m.get("Mus"), want, rtol=1e-11)
    m.close()


def test_grf_3d_matches_host_evaluation():
    n = 17
    m = StokesModel("Stokes3D_grf", n=n, nsmooth=0).setup()
    p = m.p
    md = grf.modes(p["grf_seed"], 3, p["grf_corr"], p["grf_nmodes"])
    xc, yc, zc = m.coords()
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    want = 10.0 ** np.clip(grf.evaluate(md, X, Y, Z, sigma=1.0), -2, 2)
    want[(X - 5) ** 2 + (Y - 5) ** 2 + (Z - 5) ** 2 < 1.0] = 1e-3
    np.testing.assert_allclose(m.get("Mus"), want, rtol=1e-11)
    m.close()
