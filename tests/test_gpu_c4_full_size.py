"""The 3D viscous configuration at its full size, nx = ny = nz = 1023
(8.6 GB per FP64 array, slots of the grouped state more than 4 GiB apart, so
the single-pass kernel addresses each slot through its own pointer): an exact
discrete steady state stays unchanged through FP64 PT iterations and FP32
correction sweeps in every one of the 1.07e9 cells, and the random-field
viscosity matches its host evaluation at sampled cells.

The oracle would take hours per iteration at this size; the exact state is a
size-independent property.  Homogeneous μ = 1 under x-z pure shear (εbg = 1,
Stokes3D_ve_perf.jl:164-172 ICs) has the linear velocities, τxx = -2,
τzz = +2 (τ = 2μ ε̇), all other stresses and the pressure zero as its discrete
fixed point (every difference of a stress vanishes, ∇V = 0).  A wrong offset
anywhere (a field, slot or plane addressed past 2^32 bytes) moves the state.
"""
import numpy as np
import pytest

from ptstokes import StokesModel, grf

pytestmark = pytest.mark.gpu

N = 1023
FIELDS = ["Pt", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz"]
EXACT = {"txx": -2.0, "tzz": 2.0}


def _exact_model(**env):
    m = StokesModel("Stokes3D_grf", n=N, grf_seed=None, nsmooth=0, rad2=0.0, rhog_in=0.0, ebg=1.0).setup()
    for f, v in EXACT.items():
        m.ctx.fill(f, v)
    return m


def _check_unchanged(m, v0, label, tau_rel=1e-12):
    """Velocities and pressure within 1e-12 of their scale; stresses within
    tau_rel of 2 (the defect correction recomputes τ = 2μ ε̇ from the corrected
    velocities, so a velocity change at roundoff level δV shows up as δV/dx)."""
    for f in FIELDS:
        a = m.get(f)
        if f in v0:
            d = float(np.max(np.abs(a - v0[f])))
            tol = 1e-12 * float(np.max(np.abs(v0[f])))
        else:
            d = float(np.max(np.abs(a - EXACT.get(f, 0.0))))
            tol = (1e-12 if f in ("Pt", "Vy") else tau_rel) * 2.0
        print(f"  {label} {f}: max change {d:.3e} (bound {tol:.1e})", flush=True)
        assert np.all(np.isfinite(a)), (label, f)
        assert d <= tol, (label, f, d)
        del a


def test_c4_exact_state_fp64_iterations_at_1023(monkeypatch):
    monkeypatch.setenv("PTS_3D_PLACE_TRIES", "0")
    print(f"\nC4 full size {N}^3: FP64 single-pass iterations", flush=True)
    m = _exact_model()
    try:
        v0 = {f: m.get(f) for f in ("Vx", "Vz")}
        m.iterate(3)
        m.ctx.sync()
        _check_unchanged(m, v0, "fp64")
        # the random-field viscosity at full size (the C4 field, before the
        # inclusion and smoothing): sampled cells against the host formula
        p = m.p
        md = grf.modes(1234, 3, p["grf_corr"], p["grf_nmodes"])
        m.ctx.init_grf_viscosity(md, 1.0, -2.0, 2.0, p["lx"], p["ly"], p["lz"], p["musi"], 0.0)
        mu = m.get("Mus")
        rng = np.random.default_rng(0)
        idx = rng.integers(0, N, size=(3, 4000))
        idx[:, :8] = [[0, N - 1, 0, N - 1, 0, N - 1, 0, N - 1], [0, 0, N - 1, N - 1, 0, 0, N - 1, N - 1],
                      [0, 0, 0, 0, N - 1, N - 1, N - 1, N - 1]]
        xc, yc, zc = m.coords()
        want = 10.0 ** np.clip(grf.evaluate(md, xc[idx[0]], yc[idx[1]], zc[idx[2]], sigma=1.0), -2.0, 2.0)
        got = mu[idx[0], idx[1], idx[2]]
        np.testing.assert_allclose(got, want, rtol=1e-11)
        del mu
    finally:
        m.close()


def test_c4_exact_state_fp32_sweeps_at_1023(monkeypatch):
    # the defect-correction context keeps its FP64 arrays without ping-pong
    # partners so that both precisions fit the device
    monkeypatch.setenv("PTS_3D_GROUPS", "0")
    print(f"\nC4 full size {N}^3: FP32 correction sweeps", flush=True)
    m = _exact_model()
    try:
        v0 = {f: m.get(f) for f in ("Vx", "Vz")}
        r = m.ctx.dc_solve(m.pt_params(), 1e-30, 1, max_outer=1, rtol_inner=0.0, nout_inner=3, iter_max_inner=3)
        assert r["inner_iters"] == 3, r
        # δV/dx with δV ≈ 1e-13 and dx = 10/1023
        _check_unchanged(m, v0, "fp32", tau_rel=1e-12 * m.p["lx"] / m.dx)
    finally:
        m.close()
