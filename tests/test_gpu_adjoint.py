"""Adjoint of the converged 2D viscous PT iteration and ∂J/∂ln μ (no
reference counterpart).

* Transpose: one adjoint step (pts_adjoint_iterate) is compared with the
  transpose of the iteration matrix that the oracle's forward iteration
  (f2_compute_Pt_τ, f2_compute_V, free-slip copies; Stokes2D_ve_perf.jl:47-59,
  77-78 without the elastic terms) builds column by column, on a non-square
  grid with random positive coefficients.
* Gradient: ∂J/∂ln μ (pts_grad_logmu after pts_adjoint_solve) against
  central finite differences of the converged forward solve in seeded random
  directions: the oracle's forward at 63² and at 255² (16 directions,
  h = 1e-6), and the device forward (bit-identical to the oracle's
  iteration) at 255² with h = 1e-4.
"""
import numpy as np
import pytest

from oracle import Oracle
from ptstokes import Context
from ptstokes.adjoint import Inversion, observation_points, rrule
from gpu_cases import oracle_kwargs

pytestmark = pytest.mark.gpu

ADJ = ["Pt", "txx", "tyy", "txy", "Vx", "Vy"]


def state_mask(name, shape):
    """True on the entries the iteration updates or reads as state (Dirichlet
    normal velocities are constants, not state)."""
    m = np.ones(shape, dtype=bool)
    if name == "Vx":
        m[0, :] = m[-1, :] = False
    if name == "Vy":
        m[:, 0] = m[:, -1] = False
    return m


def build_matrix(o, shapes):
    """Columns Φ e_k of one oracle PT iteration over the state entries."""
    index = [(n, k) for n in ADJ for k in np.flatnonzero(state_mask(n, shapes[n]).reshape(-1, order="F"))]
    rows = []
    for n, k in index:
        for q in ADJ:
            o.set(q, np.zeros(shapes[q]))
        e = np.zeros(int(np.prod(shapes[n])))
        e[k] = 1.0
        o.set(n, e.reshape(shapes[n], order="F"))
        o.iterate(1)
        rows.append(np.concatenate([o.get(q).reshape(-1, order="F")[state_mask(q, shapes[q]).reshape(-1, order="F")]
                                    for q in ADJ]))
    return np.array(rows).T, index  # M[i, j] = ∂out_i/∂in_j


@pytest.mark.parametrize("nx,ny,ve", [(9, 7, 0), (6, 10, 0), (8, 7, 1), (7, 9, 1)])
def test_adjoint_step_is_transpose_of_oracle_iteration(nx, ny, ve):
    # visco-elastic: τ_o = 0 (it is a parameter of the step, not state), G·dt = 0.35
    rng = np.random.default_rng(nx * 100 + ny)
    G, dt = 0.7, 0.5
    kw = dict(nx=nx, ny=ny, perf=1, ve=ve, ebg=0.0, rhog_in=0.0, rhog_out=0.0, rad2=0.0, G=G, dt=dt)
    o = Oracle("Stokes2D", **kw).setup()
    c = Context(2, nx, ny, ve=bool(ve), fd_mode="mul")
    if ve:
        for q in ("txx_o", "tyy_o", "txy_o"):
            o.set(q, np.zeros(o.shape(q)))
    dx, dy = 10.0 / nx, 10.0 / ny
    coef = {"Mus": np.exp(rng.normal(0.0, 1.5, (nx, ny))),
            "Gdtau": rng.uniform(0.2, 2.0, (nx, ny)),
            "dtau_Rho": rng.uniform(0.01, 0.2, (nx, ny))}
    for n, a in coef.items():
        o.set(n, a)
        c.upload(n, a)
    shapes = {n: o.shape(n) for n in ADJ}
    M, index = build_matrix(o, shapes)
    # no constant column: with zero Dirichlet values the iteration is linear
    for q in ADJ:
        o.set(q, np.zeros(shapes[q]))
    o.iterate(1)
    assert all(np.all(o.get(q) == 0.0) for q in ADJ)
    params = Context.pt_params(1.0, G, dt, dx, dy)
    c.adjoint_set_source(None, None)
    for trial in range(3):
        lam = {n: rng.normal(size=shapes[n]) * state_mask(n, shapes[n]) for n in ADJ}
        for n in ADJ:
            c.adjoint_upload(n, lam[n])
        c.adjoint_iterate(params, 1)
        got = {n: c.adjoint_download(n) for n in ADJ}
        vec = np.concatenate([lam[q].reshape(-1, order="F")[state_mask(q, shapes[q]).reshape(-1, order="F")]
                              for q in ADJ])
        want = M.T @ vec
        g = np.concatenate([got[q].reshape(-1, order="F")[state_mask(q, shapes[q]).reshape(-1, order="F")]
                            for q in ADJ])
        scale = np.max(np.abs(want))
        assert np.max(np.abs(g - want)) <= 1e-12 * scale, (trial, np.max(np.abs(g - want)) / scale)
        # Dirichlet entries carry no adjoint
        for q in ("Vx", "Vy"):
            assert np.all(got[q][~state_mask(q, shapes[q])] == 0.0)
    c.close()


@pytest.mark.parametrize("ve", [0, 1])
def test_adjoint_step_is_transpose_at_2047(ve):
    # the inversion's full size: with zero Dirichlet values, no body force and
    # τ_o = 0 one forward iteration is a linear map M of the state, and one
    # adjoint step must be its transpose: <M x, y> = <x, Mᵀ y> for seeded
    # random states x, y and random positive coefficients (the forward is the
    # single-pass iteration, the adjoint the single-pass adjoint sweep)
    n = 2047
    rng = np.random.default_rng(2047 + ve)
    G, dt = 0.7, 0.5
    dx = dy = 10.0 / n
    c = Context(2, n, n, ve=bool(ve), fd_mode="mul")
    coef = {"Mus": np.exp(rng.normal(0.0, 1.0, (n, n))), "Gdtau": rng.uniform(0.2, 2.0, (n, n)),
            "dtau_Rho": rng.uniform(0.01, 0.2, (n, n))}
    for name, a in coef.items():
        c.upload(name, a)
    if ve:
        for q in ("txx_o", "tyy_o", "txy_o"):
            c.upload(q, np.zeros(c.shape(q)))
    params = Context.pt_params(1.0, G, dt, dx, dy)
    x = {q: rng.normal(size=c.shape(q)) * state_mask(q, c.shape(q)) for q in ADJ}
    y = {q: rng.normal(size=c.shape(q)) * state_mask(q, c.shape(q)) for q in ADJ}
    for q in ADJ:
        c.upload(q, x[q])
    c.pt_iterate(params, 1)
    Mx = {q: c.download(q) for q in ADJ}
    c.adjoint_set_source(None, None)
    for q in ADJ:
        c.adjoint_upload(q, y[q])
    c.adjoint_iterate(params, 1)
    MTy = {q: c.adjoint_download(q) for q in ADJ}
    c.close()
    lhs = sum(float(np.sum(Mx[q][state_mask(q, Mx[q].shape)] * y[q][state_mask(q, Mx[q].shape)])) for q in ADJ)
    rhs = sum(float(np.sum(x[q] * MTy[q])) for q in ADJ)
    # scale: the sum of the magnitudes of the terms (rounding of the dot products)
    mag = sum(float(np.sum(np.abs(x[q] * MTy[q]))) for q in ADJ)
    assert abs(lhs - rhs) <= 1e-13 * mag, (lhs, rhs, mag)
    for q in ("Vx", "Vy"):
        assert np.all(MTy[q][~state_mask(q, MTy[q].shape)] == 0.0)


@pytest.mark.parametrize("nx,ny,ve", [(131, 70, 0), (61, 200, 1), (7, 5, 0), (300, 9, 1)])
def test_fused_adjoint_sweep_equals_two_passes(nx, ny, ve, monkeypatch):
    """The single-pass adjoint sweep gives the bits of the two-pass form,
    with V sources (and, visco-elastic, over a few iterations)."""
    rng = np.random.default_rng(nx + ny)
    coef = {"Mus": np.exp(rng.normal(0.0, 1.0, (nx, ny))), "Gdtau": rng.uniform(0.2, 2.0, (nx, ny)),
            "dtau_Rho": rng.uniform(0.01, 0.2, (nx, ny))}
    out = []
    for two in ("0", "1"):
        monkeypatch.setenv("PTS_ADJ_TWO_PASS", two)
        c = Context(2, nx, ny, ve=bool(ve), fd_mode="mul")
        for n, a in coef.items():
            c.upload(n, a)
        r2 = np.random.default_rng(1)
        c.adjoint_set_source(r2.normal(size=(nx + 1, ny)), r2.normal(size=(nx, ny + 1)))
        for n in ADJ:
            c.adjoint_upload(n, r2.normal(size=c.shape(n)) * state_mask(n, c.shape(n)))
        c.adjoint_iterate(Context.pt_params(0.9, 0.7, 0.5, 10.0 / nx, 10.0 / ny), 3)
        out.append({n: c.adjoint_download(n) for n in ADJ})
        c.close()
    for n in ADJ:
        assert np.array_equal(out[0][n], out[1][n]), n


def test_adjoint_source_enters_every_step():
    """λ = Φᵀλ + g: with λ = 0, one step returns g on the state entries."""
    nx, ny = 12, 9
    c = Context(2, nx, ny, ve=False, fd_mode="mul")
    rng = np.random.default_rng(5)
    for n in ("Mus", "Gdtau", "dtau_Rho"):
        c.upload(n, rng.uniform(0.5, 1.5, (nx, ny)))
    gx = rng.normal(size=(nx + 1, ny))
    gy = rng.normal(size=(nx, ny + 1))
    c.adjoint_set_source(gx, gy)
    c.adjoint_iterate(Context.pt_params(1.0, 1.0, 1.0, 1.0, 1.0), 1)
    lx, ly = c.adjoint_download("Vx"), c.adjoint_download("Vy")
    assert np.array_equal(lx[1:-1, :], gx[1:-1, :]) and np.all(lx[[0, -1], :] == 0.0)
    assert np.array_equal(ly[:, 1:-1], gy[:, 1:-1]) and np.all(ly[:, [0, -1]] == 0.0)
    c.close()


def fd_directions(shape, k, seed):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=shape) for _ in range(k)]


def _state(ctx):
    return {n: ctx.download(n) for n in ADJ}


def _oracle_like(inv, eps):
    """The oracle run of the same script with the device's μ, Musτ, Gdτ and
    dτ_Rho (the converged state does not depend on the latter three)."""
    kw = oracle_kwargs(inv.model)
    kw.update(eps=eps, nout=200, iter_max=2 * 10 ** 6)
    o = Oracle("Stokes2D", **kw).setup()
    for f in ("Mus", "Gdtau", "dtau_Rho", "Mustau"):
        o.set(f, inv.ctx.download(f))
    return o


def _J_oracle(o, inv, x0, m, eps):
    o.set("Mus", np.exp(m))
    for q in ADJ:
        o.set(q, x0[q])
    r = o.solve()
    assert r["rc"] == 0 and r["err"] <= eps, r
    v = o.get("Vx")[inv.points, inv.ny - 1] - inv.obs
    return 0.5 * float(np.dot(v, v))


def _J_device(inv, x0, m, eps):
    c = inv.ctx
    c.upload("Mus", np.exp(m))
    for q in ADJ:
        c.upload(q, x0[q])
    r = c.pt_solve(inv.params, eps, 200, 2 * 10 ** 6, 1)
    assert r["converged"], r
    v = c.download("Vx")[inv.points, inv.ny - 1] - inv.obs
    return 0.5 * float(np.dot(v, v))


def test_gradient_vs_oracle_finite_differences_63():
    """∂J/∂ln μ · δ against (J(m+hδ) − J(m−hδ))/2h, h = 1e-6, with the
    oracle forward warm-started from the converged state of m.  The forward
    stops at err ≤ 1e-12 (max-norm), just above its FP64 floor at 63²."""
    n, h, feps = 63, 1e-6, 1e-12
    inv = Inversion(n=n, eps=feps, adj_eps=1e-14, nout=200)
    inv.observe(inv.grf_logmu(4321))
    m0 = inv.grf_logmu(1234)
    J0, grad = inv.value_and_grad(m0)
    x0 = _state(inv.ctx)
    o = _oracle_like(inv, feps)
    assert abs(_J_oracle(o, inv, x0, m0, feps) - J0) <= 1e-9 * abs(J0)
    for d in fd_directions(m0.shape, 4, 99):
        fd = (_J_oracle(o, inv, x0, m0 + h * d, feps) - _J_oracle(o, inv, x0, m0 - h * d, feps)) / (2 * h)
        ad = float(np.sum(grad * d))
        assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    inv.close()


def _check_fd(fds, ads, med_rel, max_scaled):
    """Median relative error, and the largest error scaled by the RMS
    directional derivative (a direction nearly orthogonal to the gradient has
    a small derivative, whose relative error only measures the FD noise)."""
    fds, ads = np.array(fds), np.array(ads)
    err = np.abs(fds - ads)
    rel = err / np.abs(ads)
    scaled = err / np.sqrt(np.mean(ads ** 2))
    msg = list(zip(fds.tolist(), ads.tolist(), rel.tolist()))
    assert np.median(rel) <= med_rel and np.max(scaled) <= max_scaled, msg


@pytest.fixture(scope="module")
def inv255():
    """Hidden ln μ (seed 4321), 256 top points, gradient at the seed-1234
    field; forward to err ≤ 3e-11 (its FP64 floor at 255² is ≈ 2e-11)."""
    inv = Inversion(n=255, eps=3e-11, adj_eps=1e-14, nout=200)
    inv.observe(inv.grf_logmu(4321))
    m0 = inv.grf_logmu(1234)
    J0, grad = inv.value_and_grad(m0)
    assert inv.last["adjoint"]["converged"]
    yield inv, m0, J0, grad, _state(inv.ctx)
    inv.close()


def test_gradient_vs_oracle_finite_differences_255(inv255):
    """16 seeded directions, h = 1e-6, oracle forward.  The FD error is set
    by the forward's stop: err ≤ 3e-11 leaves J uncertain by ≈1e-10, i.e.
    ≈1e-4 of a typical directional derivative at this h."""
    inv, m0, J0, grad, x0 = inv255
    feps, h = 3e-11, 1e-6
    o = _oracle_like(inv, feps)
    fds, ads = [], []
    for d in fd_directions(m0.shape, 16, 7):
        fds.append((_J_oracle(o, inv, x0, m0 + h * d, feps) - _J_oracle(o, inv, x0, m0 - h * d, feps)) / (2 * h))
        ads.append(float(np.sum(grad * d)))
    _check_fd(fds, ads, 1e-4, 1e-3)


def test_gradient_vs_device_finite_differences_255(inv255):
    """The same 16 directions with h = 1e-4 (FD error ≈ O(h²) + 1e-6) and the
    device forward."""
    inv, m0, J0, grad, x0 = inv255
    feps, h = 3e-11, 1e-4
    fds, ads = [], []
    for d in fd_directions(m0.shape, 16, 7):
        fds.append((_J_device(inv, x0, m0 + h * d, feps) - _J_device(inv, x0, m0 - h * d, feps)) / (2 * h))
        ads.append(float(np.sum(grad * d)))
    _check_fd(fds, ads, 1e-6, 2e-5)


def test_rrule_pullback_and_points():
    pts = observation_points(255)
    assert pts.size == 256 and pts.min() >= 1 and pts.max() <= 254
    assert np.array_equal(pts, observation_points(255))
    inv = Inversion(n=31, eps=1e-12, adj_eps=1e-13, nout=100, npts=16)
    inv.observe(inv.grf_logmu(4321))
    m0 = inv.grf_logmu(1234)
    J, pb = rrule(inv, m0)
    J2, g2 = inv.value_and_grad(m0)
    assert J == J2
    assert np.allclose(pb(2.0), 2.0 * g2, rtol=1e-12, atol=0.0)
    # at the true viscosity the misfit and its gradient vanish
    inv2 = Inversion(n=31, eps=1e-12, adj_eps=1e-13, nout=100, npts=16)
    mt = inv2.grf_logmu(4321)
    inv2.observe(mt)
    Jt, gt = inv2.value_and_grad(mt)
    assert Jt == 0.0 and np.all(gt == 0.0)
    inv.close()
    inv2.close()


@pytest.mark.parametrize("nt", [4, 10])
def test_ve_gradient_through_checkpointed_steps(nt):
    """Visco-elastic, nt physical steps (each step's τ is the next step's τ_o;
    10 is the inversion's time loop): the checkpointed backward pass against
    central differences of the whole forward time loop (device forward,
    bit-identical to the oracle's; the oracle reproduces J at the base point)."""
    n, h, feps = 48, 1e-5, 1e-12
    inv = Inversion(n=n, preset="Stokes2D_ve_grf", nt=nt, eps=feps, adj_eps=1e-14, nout=100, npts=40)
    inv.observe(inv.grf_logmu(4321))
    m0 = inv.grf_logmu(1234)
    J0, grad = inv.value_and_grad(m0)
    assert inv.last["adjoint"]["converged"] and np.all(np.isfinite(grad))
    kw = oracle_kwargs(inv.model)
    kw.update(eps=feps, nout=100, iter_max=10 ** 6)
    inv.set_logmu(m0)
    o = Oracle("Stokes2D", **kw).setup()
    for f in ("Mus", "Gdtau", "dtau_Rho", "Mustau"):
        o.set(f, inv.ctx.download(f))
    for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy", "txx_o", "tyy_o", "txy_o"):
        o.set(f, inv.ctx.download(f))
    r = o.solve()
    assert r["rc"] == 0
    v = o.get("Vx")[inv.points, n - 1] - inv.obs
    assert abs(0.5 * float(np.dot(v, v)) - J0) <= 1e-9 * abs(J0)
    fds, ads = [], []
    for d in fd_directions(m0.shape, 6, 11):
        fds.append((inv.value(m0 + h * d) - inv.value(m0 - h * d)) / (2 * h))
        ads.append(float(np.sum(grad * d)))
    _check_fd(fds, ads, 1e-5, 1e-4)
    # one step only (τ_o fixed) would miss the coupling: the multi-step
    # gradient is not the last step's
    inv.set_logmu(m0)
    inv.solve_forward()
    g1, _ = inv.dJ_dVx()
    inv.ctx.adjoint_set_source(g1, None)
    res = inv.ctx.adjoint_solve(inv.params, 1e-14, 100, 10 ** 7)
    assert res["converged"]
    last = inv.ctx.grad_logmu(inv.params)
    assert np.max(np.abs(last - grad)) > 1e-3 * np.max(np.abs(grad))
    inv.close()

