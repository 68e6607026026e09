"""Mixed precision (FP32 PT sweeps inside FP64 defect correction) and the
random-field viscosity: the FP64 residual reaches eps and the solution agrees
with the FP64 PT solve to the accuracy the tolerance allows."""
import numpy as np
import pytest

from ptstokes import StokesModel, grf
from _util import relmax

pytestmark = pytest.mark.gpu

FIELDS = ["Pt", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz"]


@pytest.mark.parametrize("preset,kw", [
    ("Stokes2D_grf", dict(n=63)),
    ("Stokes2D", dict(n=127, fd="mul", rhog_in=-1.0, norm_kind=1)),
    ("Stokes3D_grf", dict(n=31)),
])
def test_dc_matches_fp64_solve(preset, kw):
    ref = StokesModel(preset, eps=1e-11, **kw).setup()
    r0 = ref.step()
    assert r0["converged"]
    dc = StokesModel(preset, eps=1e-8, **kw).setup()
    res = dc.solve_dc()
    assert res["converged"] and res["err"] <= 1e-8, res
    # each correction cuts the FP64 residual by orders of magnitude, except
    # the last, whose inner stop only asks for err below eps
    eo = res["err_outer"]
    assert all(b < max(a * 1e-2, 1e-8) for a, b in zip(eo, eo[1:])), eo
    nd = ref.p["ndim"]
    for f in FIELDS:
        if nd == 2 and f in ("Vz", "tzz", "txz", "tyz"):
            continue
        d = relmax(dc.get(f), ref.get(f))
        assert d < 1e-5, (f, d)
    ref.close()
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
