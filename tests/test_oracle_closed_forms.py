"""The oracle against what the mathematics fixes (closed forms, invariants and
a direct solve of the steady discrete equations).

These pin the parts of the oracle that the reference's golden files leave
open (V and τ fields, the viscous-3D and body-force variants, the fixed point
of the PT iteration), chosen so that a dropped factor, a wrong sign or a
transposed index fails at least one of them.
"""
import numpy as np
import pytest

from oracle import Oracle


@pytest.mark.parametrize("preset", ["Stokes2D", "Stokes2D_ve_perf"])
def test_homogeneous_viscous_stress_is_2_mu_edot(preset):
    # uniform μ=1: linear pure shear is exact, so V stays put (up to the
    # rounding of ∇V) and τ relaxes to 2μ ε̇ with ε̇xx = -εbg, ε̇yy = +εbg, ε̇xy = 0
    o = Oracle(preset, n=15, musi=1.0, ve=0).setup()
    Vx0, Vy0 = o.get("Vx"), o.get("Vy")
    o.iterate(3000)
    np.testing.assert_allclose(o.get("Vx"), Vx0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("Vy"), Vy0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("txx"), -2.0, rtol=1e-12)
    np.testing.assert_allclose(o.get("tyy"), 2.0, rtol=1e-12)
    assert np.max(np.abs(o.get("txy"))) < 1e-12
    assert np.max(np.abs(o.get("Pt"))) < 1e-10


@pytest.mark.parametrize("preset,n", [("Stokes2D_ve", 15), ("Stokes2D_ve_perf", 15),
                                      ("Stokes3D_ve_multixpu", 9), ("Stokes3D_ve_perf", 9)])
def test_homogeneous_maxwell_recursion(preset, n):
    # μ = G = dt = 1: backward-Euler Maxwell per physical step
    #   τ_k = (τ_{k-1}/(G dt) + 2 ε̇)/(1/μ + 1/(G dt))  ->  1, 1.5, 1.75, 1.875
    o = Oracle(preset, n=n, musi=1.0).setup()
    ndim = o.p["ndim"]
    big = "tyy" if ndim == 2 else "tzz"
    tau = 0.0
    for _ in range(4):
        o.assign_tau()
        o.iterate(4000)
        tau = (tau / 1.0 + 2.0 * 1.0) / (1.0 / 1.0 + 1.0 / 1.0)
        np.testing.assert_allclose(o.get(big), tau, rtol=1e-11)
        np.testing.assert_allclose(o.get("txx"), -tau, rtol=1e-11)
    if ndim == 3:
        assert np.max(np.abs(o.get("tyy"))) < 1e-11


@pytest.mark.parametrize("perf", [0, 1])
def test_hydrostatic_body_force(perf):
    # εbg = 0 and a uniform body force c: V -> 0, τ -> 0 and ∂P/∂y = -c
    # (Ry = ... - ∂yP - avg(ρg), Stokes2D_param.jl:59)
    c = 0.75
    o = Oracle("Stokes2D", n=15, ebg=0.0, rhog_in=c, rhog_out=c, perf=perf).setup()
    o.iterate(20000)
    dy = 10.0 / 15
    P = o.get("Pt")
    np.testing.assert_allclose((P[:, 1:] - P[:, :-1]) / dy, -c, rtol=1e-9)
    np.testing.assert_allclose(P[1:, :] - P[:-1, :], 0.0, atol=1e-9)
    assert np.max(np.abs(o.get("Vx"))) < 1e-9 and np.max(np.abs(o.get("Vy"))) < 1e-9


@pytest.mark.parametrize("preset", ["Stokes3D_ve_perf", "Stokes3D_ve_multixpu"])
def test_homogeneous_viscous_stress_3d(preset):
    # 3D viscous (ve = 0), uniform μ = 1, x-z pure shear: linear V is exact,
    # so V stays put and τ relaxes to 2μ ε̇ with ε̇xx = -εbg, ε̇zz = +εbg and
    # every other component 0
    o = Oracle(preset, n=9, musi=1.0, ve=0, nt=1).setup()
    V0 = {f: o.get(f) for f in ("Vx", "Vy", "Vz")}
    o.iterate(4000)
    for f, v in V0.items():
        np.testing.assert_allclose(o.get(f), v, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("txx"), -2.0, rtol=1e-12)
    np.testing.assert_allclose(o.get("tzz"), 2.0, rtol=1e-12)
    for f in ("tyy", "txy", "txz", "tyz"):
        assert np.max(np.abs(o.get(f))) < 1e-12, f
    assert np.max(np.abs(o.get("Pt"))) < 1e-10


def test_hydrostatic_body_force_3d():
    c = 0.5
    o = Oracle("Stokes3D_ve_perf", n=9, ebg=0.0, rhog_in=c, rhog_out=c, ve=0).setup()
    o.iterate(20000)
    dz = 10.0 / 9
    P = o.get("Pt")
    np.testing.assert_allclose((P[:, :, 1:] - P[:, :, :-1]) / dz, -c, rtol=1e-9)


@pytest.mark.parametrize("preset", ["Stokes2D", "Stokes2D_ve", "Stokes2D_ve_perf"])
def test_pure_shear_antisymmetry_2d(preset):
    # lx = ly, circular inclusion at the centre, pure shear: swapping x and y
    # flips the sign of the strain rate, so P(x,y) = -P(y,x), Vy(y,x) = -Vx(x,y)
    o = Oracle(preset, n=31).setup()
    o.solve()
    P, Vx, Vy = o.get("Pt"), o.get("Vx"), o.get("Vy")
    s = np.max(np.abs(P))
    assert np.max(np.abs(P + P.T)) <= 1e-12 * s
    assert np.max(np.abs(Vy.T + Vx)) <= 1e-12 * np.max(np.abs(Vx))
    np.testing.assert_allclose(o.get("txx"), -o.get("tyy").T, atol=1e-12 * np.max(np.abs(o.get("txx"))))


@pytest.mark.parametrize("preset", ["Stokes3D_ve_perf", "Stokes3D_ve_multixpu"])
def test_pure_shear_antisymmetry_3d(preset):
    # x-z pure shear with a centred sphere: P(x,y,z) = -P(z,y,x)
    o = Oracle(preset, n=15, nt=1).setup()
    o.solve()
    P = o.get("Pt")
    assert np.max(np.abs(P + P.transpose(2, 1, 0))) <= 1e-12 * np.max(np.abs(P))
    Vx, Vz = o.get("Vx"), o.get("Vz")
    assert np.max(np.abs(Vz.transpose(2, 1, 0) + Vx)) <= 1e-12 * np.max(np.abs(Vx))


# ---------------------------------------------------------------------------
# Direct solve of the steady discrete viscous Stokes equations (staggered grid,
# free-slip tangential copies, Dirichlet normal velocities): the PT fixed
# point must be that solution.  Pressure is fixed up to a constant.
# ---------------------------------------------------------------------------
def _steady_residual(u, nx, ny, dx, dy, Mus, rg, Vx_b, Vy_b, tau_o=None, Gdt=None):
    nvx, nvy = (nx - 1) * (ny - 2), (nx - 2) * (ny - 1)
    Vx = Vx_b.copy()
    Vy = Vy_b.copy()
    Vx[1:nx, 1:ny - 1] = u[:nvx].reshape(nx - 1, ny - 2, order="F")
    Vy[1:nx - 1, 1:ny] = u[nvx:nvx + nvy].reshape(nx - 2, ny - 1, order="F")
    P = u[nvx + nvy:].reshape(nx, ny, order="F")
    Vx[:, 0], Vx[:, -1] = Vx[:, 1], Vx[:, -2]
    Vy[0, :], Vy[-1, :] = Vy[1, :], Vy[-2, :]
    exx = np.diff(Vx, axis=0) / dx
    eyy = np.diff(Vy, axis=1) / dy
    def law(mu, e, name):
        # viscous, or Maxwell with old stress τ_o over G·dt (backward Euler)
        if tau_o is None:
            return 2 * mu * e
        return (2 * e + tau_o[name] / Gdt) / (1 / mu + 1 / Gdt)

    txx, tyy = law(Mus, exx, "txx"), law(Mus, eyy, "tyy")
    mv = 0.25 * (Mus[:-1, :-1] + Mus[1:, :-1] + Mus[:-1, 1:] + Mus[1:, 1:])
    exy = 0.5 * (np.diff(Vx[1:nx, :], axis=1) / dy + np.diff(Vy[:, 1:ny], axis=0) / dx)
    txy = law(mv, exy, "txy")
    Rx = np.diff(txx[:, 1:-1], axis=0) / dx + np.diff(txy[:, :], axis=1)[:, :] / dy - np.diff(P[:, 1:-1], axis=0) / dx
    Ry = (np.diff(tyy[1:-1, :], axis=1) / dy + np.diff(txy, axis=0) / dx - np.diff(P[1:-1, :], axis=1) / dy
          - 0.5 * (rg[1:-1, :-1] + rg[1:-1, 1:]))
    return np.concatenate([Rx.ravel(order="F"), Ry.ravel(order="F"), (exx + eyy).ravel(order="F")])


@pytest.mark.parametrize("preset,rhog_in", [("Stokes2D", 0.0), ("Stokes2D", -1.0), ("Stokes2D_ve", 0.0),
                                            ("Stokes2D_ve_perf", 0.0)])
def test_pt_fixed_point_equals_direct_solve(preset, rhog_in):
    n = 13
    o = Oracle(preset, n=n, nt=1, eps=1e-12, nout=50, iter_max=200000, rhog_in=rhog_in).setup()
    Mus, rg, Vx_b, Vy_b = o.get("Mus"), o.get("rhog"), o.get("Vx"), o.get("Vy")
    tau_o, Gdt = None, None
    if o.p["ve"]:
        # random old stresses: the solve's assign_τ! copies τ into τ_o
        rng = np.random.default_rng(5)
        tau_o = {f: rng.uniform(-1.0, 1.0, o.shape(f)) for f in ("txx", "tyy", "txy")}
        for f, v in tau_o.items():
            o.set(f, v)
        Gdt = o.p["G"] * o.p["dt"]
    res = o.solve()
    assert res["rc"] == 0 and res["err"] <= 1e-12
    dx = dy = 10.0 / n
    nvx, nvy, npp = (n - 1) * (n - 2), (n - 2) * (n - 1), n * n
    N = nvx + nvy + npp
    f0 = _steady_residual(np.zeros(N), n, n, dx, dy, Mus, rg, Vx_b, Vy_b, tau_o, Gdt)
    A = np.empty((N, N))
    for k in range(N):
        e = np.zeros(N)
        e[k] = 1.0
        A[:, k] = _steady_residual(e, n, n, dx, dy, Mus, rg, Vx_b, Vy_b, tau_o, Gdt) - f0
    # The four corner pressures appear in no momentum equation (free), and P is
    # defined up to a constant: fix the sum over the other cells to the PT one
    # and compare away from the corners.
    Pt = o.get("Pt")
    inner = np.ones((n, n))
    inner[0, 0] = inner[0, -1] = inner[-1, 0] = inner[-1, -1] = 0.0
    Ag = np.vstack([A, np.concatenate([np.zeros(nvx + nvy), inner.ravel(order="F")])])
    bg = np.concatenate([-f0, [(Pt * inner).sum()]])
    u = np.linalg.lstsq(Ag, bg, rcond=None)[0]
    Vx_d = u[:nvx].reshape(n - 1, n - 2, order="F")
    Vy_d = u[nvx:nvx + nvy].reshape(n - 2, n - 1, order="F")
    P_d = u[nvx + nvy:].reshape(n, n, order="F")
    scale_v = np.max(np.abs(Vx_b))
    assert np.max(np.abs(o.get("Vx")[1:n, 1:n - 1] - Vx_d)) <= 1e-9 * scale_v
    assert np.max(np.abs(o.get("Vy")[1:n - 1, 1:n] - Vy_d)) <= 1e-9 * scale_v
    assert np.max(np.abs((Pt - P_d) * inner)) <= 1e-8 * np.max(np.abs(P_d))


def _steady_residual_3d(u, n, d, Mus, rg, Vb, tau_o=None, Gdt=None):
    """Steady viscous Stokes residuals on the 3D staggered grid (n³ cells,
    spacing d): interior velocities from u, Dirichlet normal faces from Vb,
    free-slip tangential copies, τ = 2μ ε̇ with 4-cell averages of μ for the
    shear components, ρg_z averaged onto the Vz faces.  With old stresses
    tau_o (Maxwell, backward Euler over G·dt = Gdt) each component is
    τ = (2ε̇ + τ_o/Gdt) / (1/μ + 1/Gdt) instead."""
    Vx, Vy, Vz = (v.copy() for v in Vb)
    sx, sy, sz = (n - 1, n - 2, n - 2), (n - 2, n - 1, n - 2), (n - 2, n - 2, n - 1)
    nv = int(np.prod(sx))
    Vx[1:n, 1:n - 1, 1:n - 1] = u[:nv].reshape(sx, order="F")
    Vy[1:n - 1, 1:n, 1:n - 1] = u[nv:2 * nv].reshape(sy, order="F")
    Vz[1:n - 1, 1:n - 1, 1:n] = u[2 * nv:3 * nv].reshape(sz, order="F")
    P = u[3 * nv:].reshape((n, n, n), order="F")
    # free-slip: copy the first interior plane outward (tangential components)
    Vy[0], Vy[-1] = Vy[1], Vy[-2]
    Vz[0], Vz[-1] = Vz[1], Vz[-2]
    Vx[:, 0], Vx[:, -1] = Vx[:, 1], Vx[:, -2]
    Vz[:, 0], Vz[:, -1] = Vz[:, 1], Vz[:, -2]
    Vx[:, :, 0], Vx[:, :, -1] = Vx[:, :, 1], Vx[:, :, -2]
    Vy[:, :, 0], Vy[:, :, -1] = Vy[:, :, 1], Vy[:, :, -2]
    exx, eyy, ezz = np.diff(Vx, axis=0) / d, np.diff(Vy, axis=1) / d, np.diff(Vz, axis=2) / d
    def law(mu, e, name):
        if tau_o is None:
            return 2 * mu * e
        return (2 * e + tau_o[name] / Gdt) / (1 / mu + 1 / Gdt)

    txx = law(Mus[:, 1:-1, 1:-1], exx[:, 1:-1, 1:-1], "txx")
    tyy = law(Mus[1:-1, :, 1:-1], eyy[1:-1, :, 1:-1], "tyy")
    tzz = law(Mus[1:-1, 1:-1, :], ezz[1:-1, 1:-1, :], "tzz")

    def avg4(a, ax1, ax2):
        lo1, hi1 = [slice(None)] * 3, [slice(None)] * 3
        lo1[ax1], hi1[ax1] = slice(None, -1), slice(1, None)
        b = a[tuple(lo1)] + a[tuple(hi1)]
        lo2, hi2 = [slice(None)] * 3, [slice(None)] * 3
        lo2[ax2], hi2[ax2] = slice(None, -1), slice(1, None)
        return 0.25 * (b[tuple(lo2)] + b[tuple(hi2)])

    exy = 0.5 * (np.diff(Vx[1:n], axis=1) / d + np.diff(Vy[:, 1:n], axis=0) / d)
    exz = 0.5 * (np.diff(Vx[1:n], axis=2) / d + np.diff(Vz[:, :, 1:n], axis=0) / d)
    eyz = 0.5 * (np.diff(Vy[:, 1:n], axis=2) / d + np.diff(Vz[:, :, 1:n], axis=1) / d)
    txy = law(avg4(Mus, 0, 1)[:, :, 1:-1], exy[:, :, 1:-1], "txy")
    txz = law(avg4(Mus, 0, 2)[:, 1:-1, :], exz[:, 1:-1, :], "txz")
    tyz = law(avg4(Mus, 1, 2)[1:-1, :, :], eyz[1:-1, :, :], "tyz")
    Rx = (np.diff(txx, axis=0) + np.diff(txy, axis=1) + np.diff(txz, axis=2)
          - np.diff(P[:, 1:-1, 1:-1], axis=0)) / d
    Ry = (np.diff(tyy, axis=1) + np.diff(txy, axis=0) + np.diff(tyz, axis=2)
          - np.diff(P[1:-1, :, 1:-1], axis=1)) / d
    Rz = ((np.diff(tzz, axis=2) + np.diff(txz, axis=0) + np.diff(tyz, axis=1)
           - np.diff(P[1:-1, 1:-1, :], axis=2)) / d - 0.5 * (rg[1:-1, 1:-1, :-1] + rg[1:-1, 1:-1, 1:]))
    return np.concatenate([Rx.ravel(order="F"), Ry.ravel(order="F"), Rz.ravel(order="F"),
                           (exx + eyy + ezz).ravel(order="F")])


@pytest.mark.parametrize("preset", ["Stokes3D_ve_multixpu", "Stokes3D_ve_perf"])
@pytest.mark.parametrize("rhog_in,ve", [(0.0, 0), (-1.0, 0), (0.0, 1)])
def test_pt_fixed_point_equals_direct_solve_3d(preset, rhog_in, ve):
    n = 7
    o = Oracle(preset, n=n, ve=ve, nt=1, eps=1e-12, nout=50, iter_max=400000, rhog_in=rhog_in).setup()
    Mus, rg = o.get("Mus"), o.get("rhog")
    Vb = (o.get("Vx"), o.get("Vy"), o.get("Vz"))
    tau_o, Gdt = None, None
    if ve:
        # random old stresses: the solve's assign_τ! copies τ into τ_o
        rng = np.random.default_rng(3)
        tau_o = {f: rng.uniform(-1.0, 1.0, o.shape(f)) for f in ("txx", "tyy", "tzz", "txy", "txz", "tyz")}
        for f, v in tau_o.items():
            o.set(f, v)
        Gdt = o.p["G"] * o.p["dt"]
    res = o.solve()
    assert res["rc"] == 0 and res["err"] <= 1e-12
    d = 10.0 / n
    nv = (n - 1) * (n - 2) ** 2
    N = 3 * nv + n ** 3
    f0 = _steady_residual_3d(np.zeros(N), n, d, Mus, rg, Vb, tau_o, Gdt)
    A = np.empty((N, N))
    for k in range(N):
        e = np.zeros(N)
        e[k] = 1.0
        A[:, k] = _steady_residual_3d(e, n, d, Mus, rg, Vb, tau_o, Gdt) - f0
    # Pressures on box edges and corners (two or more boundary indices) appear
    # in no momentum equation, and P is defined up to a constant: fix the sum
    # over the other cells to the PT one and compare on them.
    idx = np.indices((n, n, n))
    nbound = sum(((i == 0) | (i == n - 1)).astype(int) for i in idx)
    inner = (nbound <= 1).astype(float)
    Pt = o.get("Pt")
    Ag = np.vstack([A, np.concatenate([np.zeros(3 * nv), inner.ravel(order="F")])])
    bg = np.concatenate([-f0, [(Pt * inner).sum()]])
    u = np.linalg.lstsq(Ag, bg, rcond=None)[0]
    scale_v = np.max(np.abs(Vb[0]))
    got = (o.get("Vx")[1:n, 1:n - 1, 1:n - 1], o.get("Vy")[1:n - 1, 1:n, 1:n - 1], o.get("Vz")[1:n - 1, 1:n - 1, 1:n])
    for c, g in enumerate(got):
        want = u[c * nv:(c + 1) * nv].reshape(g.shape, order="F")
        assert np.max(np.abs(g - want)) <= 1e-9 * scale_v, c
    P_d = u[3 * nv:].reshape((n, n, n), order="F")
    assert np.max(np.abs((Pt - P_d) * inner)) <= 1e-8 * np.max(np.abs(P_d))


def test_c1_cfl_095_diverges():
    # The 127^2 viscous inclusion problem with rho_g_y = -1, no background
    # shear and the dispersion-optimal Re, r: with the CFL number 0.95/sqrt(2)
    # the PT iteration blows up (NaN at a residual check), with the script's
    # 0.9/sqrt(2) (Stokes2D.jl:82) it converges to the scaled max-norm 1e-8
    # in 10200 iterations.
    from oracle import R_OPT, RE_OPT
    kw = dict(n=127, Re=RE_OPT, r=R_OPT, rhog_in=-1.0, ebg=0.0, norm_inf=1, nout=200, eps=1e-8)
    bad = Oracle("Stokes2D", CFL=0.95 / np.sqrt(2.0), **kw).setup().solve()
    assert bad["rc"] == -1 and np.isnan(bad["err"]) and bad["ittot"] < 5000, bad
    good = Oracle("Stokes2D", CFL=0.9 / np.sqrt(2.0), **kw).setup().solve()
    assert good["rc"] == 0 and good["err"] < 1e-8 and good["ittot"] == 10200, good
