"""Whole reference scripts replayed on the GPU: the golden files, the
oracle's fields and iteration counts, and the published counts."""
import numpy as np
import pytest

from gpu_cases import STATE, compare, pair
from _util import FIXTURES, PUBLISHED_ITERS, julia_isapprox, load_fixture, relmax, sample
from ptstokes import StokesModel, PtsError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset", list(FIXTURES))
def test_gpu_reproduces_golden_file(preset):
    stem, n = FIXTURES[preset]
    P, X, idx = load_fixture(stem)
    m, o = pair(preset, n=n)
    out = m.run()
    res = o.solve()
    S = sample(out[-1], idx)
    assert julia_isapprox(S, P)
    assert np.array_equal(S, P)
    assert m.ittot == res["ittot"]
    assert m.iters_per_step == res["iters_per_step"]
    # every field matches the oracle (V and τ are not in the golden files)
    assert compare(m, o, STATE) == []
    m.close()


@pytest.mark.parametrize("preset,n,want", PUBLISHED_ITERS[:3])
def test_gpu_matches_published_iteration_counts(preset, n, want):
    m = StokesModel(preset, n=n).setup()
    m.run()
    assert m.ittot == want
    m.close()


def test_gpu_matches_published_iteration_count_3d():
    preset, n, want = PUBLISHED_ITERS[3]
    m = StokesModel(preset, n=n).setup()
    m.run()
    assert m.ittot == want
    m.close()


def test_graph_and_stream_solves_agree():
    a = StokesModel("Stokes2D_ve", n=63, nt=2).setup()
    b = StokesModel("Stokes2D_ve", n=63, nt=2).setup()
    a.run(use_graph=True)
    b.run(use_graph=False)
    assert a.iters_per_step == b.iters_per_step
    for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_err_matches_oracle_at_each_check(kind):
    m, o = pair("Stokes2D_ve_perf", n=47, nout=10, iter_max=59, norm_kind=kind)
    res = m.step()
    ref = o.solve()
    assert res["iters"] == ref["ittot"] == 60
    assert res["err"] == pytest.approx(ref["err"], rel=1e-12)
    m.close()


def test_nan_is_reported():
    m = StokesModel("Stokes2D_ve_perf", n=31, nout=5, iter_max=50).setup()
    Pt = m.get("Pt")
    Pt[10, 10] = np.nan
    m.ctx.upload("Pt", Pt)
    with pytest.raises(PtsError) as e:
        m.step()
    assert e.value.code == -3
    m.close()


def test_iter_max_reached_is_flagged():
    m = StokesModel("Stokes2D_ve", n=63, iter_max=100).setup()
    res = m.step()
    assert not res["converged"] and res["iters"] == 101
    m.close()


def test_bad_arguments_raise():
    from ptstokes import Context
    with pytest.raises(PtsError):
        Context(2, 2, 10)
    c = Context(2, 10, 10)
    with pytest.raises(ValueError):
        c.upload("Pt", np.zeros((10, 11)))  # wrong shape
    with pytest.raises(PtsError):  # wrong count at the C boundary
        buf = np.zeros(5)
        c._chk(c.lib.pts_upload(c.h, 0, buf.ctypes.data_as(__import__("ptstokes")._lib._dp), 5))
    with pytest.raises(PtsError):
        c.bc_z("Pt")  # 3D kernel on a 2D grid
    c.close()


@pytest.mark.parametrize("preset,kw", [("Stokes3D_ve_perf", dict(n=23)), ("Stokes2D_ve_perf", dict(n=95, dt=0.1))])
def test_chunked_solve_equals_one_loop(preset, kw):
    # the time-to-solution runs drive pts_pt_solve in chunks whose length is a
    # multiple of the check interval: iteration count, final err and fields
    # are those of one loop
    opts = dict(iter_max=10 ** 6, nout=200, norm_kind=1, eps=1e-8, **kw)
    a = StokesModel(preset, **opts).setup()
    b = StokesModel(preset, **opts).setup()
    a.ctx.assign_tau()
    b.ctx.assign_tau()
    ra = a.ctx.pt_solve(a.pt_params(), 1e-8, 200, 10 ** 6, 1)
    assert ra["converged"]
    iters = 0
    while True:
        rb = b.ctx.pt_solve(b.pt_params(), 1e-8, 200, 1000 - 1, 1)
        iters += rb["iters"]
        if rb["converged"]:
            break
        assert iters < 10 ** 6
    assert iters == ra["iters"] and rb["err"] == ra["err"], (iters, ra)
    for f in STATE:
        if a.p["ndim"] == 2 and f in ("Vz", "tzz", "txz", "tyz"):
            continue
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_c1_viscous_127_dispersion_optimum_bit_exact():
    # 2D viscous Stokes (Stokes2D.jl, divide form) at 127^2: 1e-3 inclusion of
    # radius 1 (10 smoothing passes) with rho_g_y = -1 inside and 0 outside
    # (Stokes2D_param.jl:59), free-slip walls without background shear,
    # Re = 9*sqrt(3)*pi/4 and r = 1/4 (the 1D dispersion optimum,
    # dispersion_analysis_stokes1D.m:34-35), stopped when the scaled max-norm
    # of the residuals is below 1e-8, checked every 200 iterations.  CFL is
    # the script's 0.9/sqrt(2) (Stokes2D.jl:82): 0.95/sqrt(2) diverges on this
    # problem (test_oracle_closed_forms.py::test_c1_cfl_095_diverges).
    from ptstokes.models import R_OPT, RE_OPT
    kw = dict(n=127, CFL=0.9 / np.sqrt(2.0), Re=RE_OPT, r=R_OPT, rhog_in=-1.0, ebg=0.0, norm_kind=1, nout=200,
              eps=1e-8)
    m, o = pair("Stokes2D", **kw)
    r = m.step()
    res = o.solve()
    assert res["rc"] == 0 and r["converged"]
    assert r["iters"] == res["ittot"] == 10200
    assert compare(m, o, STATE) == []
    m.close()


@pytest.mark.parametrize("preset,kw,nout", [("Stokes2D", dict(n=127, rhog_in=-1.0, ebg=0.0), 10),
                                            ("Stokes2D_ve_perf", dict(nx=90, ny=61), 7),
                                            ("Stokes2D_ve", dict(n=63, dt=0.1), 200)])
@pytest.mark.parametrize("rows", ["1", "2", "5"])
def test_cooperative_chunks_equal_single_launches(preset, kw, nout, rows, monkeypatch):
    # small 2D grids: pts_pt_solve runs each chunk between checks as one
    # cooperative launch with grid barriers; the result equals one launch per
    # iteration bit for bit, across odd and even chunk lengths
    monkeypatch.setenv("PTS_2D_LDS", "0")  # (the state-in-LDS launch comes first otherwise)
    monkeypatch.setenv("PTS_2D_MULTI_ROWS", rows)
    monkeypatch.setenv("PTS_2D_MULTI", "1")
    a = StokesModel(preset, nout=nout, **kw).setup()
    monkeypatch.setenv("PTS_2D_MULTI", "0")
    b = StokesModel(preset, nout=nout, **kw).setup()
    if a.p["ve"]:
        a.ctx.assign_tau()
        b.ctx.assign_tau()
    params = a.pt_params()
    for it_max in (29, 40):
        ra = a.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        rb = b.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        assert ra["iters"] == rb["iters"] and ra["err"] == rb["err"], (ra, rb)
        for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy"):
            assert np.array_equal(a.get(f), b.get(f)), (f, it_max)
    a.close()
    b.close()


LDS_GEOMETRIES = [
    ("Stokes2D", dict(n=127, rhog_in=-1.0, ebg=0.0), 10, {}),
    ("Stokes2D", dict(nx=70, ny=42, rhog_in=-1.0), 9, {"PTS_DERIVED_PARAMS": "0"}),
    ("Stokes2D_ve_perf", dict(nx=90, ny=61), 7, {}),
    ("Stokes2D_ve_perf", dict(nx=33, ny=100, rhog_in=-1.0), 12, {"PTS_DERIVED_PARAMS": "0"}),
    ("Stokes2D_ve", dict(n=63, dt=0.1), 200, {}),
    ("Stokes2D_ve", dict(nx=6, ny=5, rad2=9.0, rhog_in=-1.0), 3, {}),
]


@pytest.mark.parametrize("preset,kw,nout,env", LDS_GEOMETRIES)
@pytest.mark.parametrize("rows", ["0", "2", "3", "7", "40"])
@pytest.mark.parametrize("halo", ["0", "1", "2", "3", "4"])
def test_state_in_lds_chunks_equal_single_launches(preset, kw, nout, env, rows, halo, monkeypatch):
    # small 2D grids: each chunk between checks runs with the state in LDS,
    # blocks of `rows` rows (at least `halo`; the last block owns the rows
    # left, and those of the block before when a block's updated rows would
    # start at row ny-1) exchanging their `halo` edge rows every `halo`
    # iterations (computing the halo rows along; 0 and 0: the default
    # choice of both); the result
    # equals one launch per iteration bit for bit, across odd and even chunk
    # lengths and repeated launches, over enough iterations for the flow to
    # reach the walls (3 * nout + 1 after 69)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    monkeypatch.setenv("PTS_2D_LDS_HALO", halo)
    monkeypatch.setenv("PTS_2D_LDS_ROWS", rows)
    monkeypatch.setenv("PTS_2D_LDS", "1")
    a = StokesModel(preset, nout=nout, **kw).setup()
    monkeypatch.setenv("PTS_2D_LDS", "0")
    monkeypatch.setenv("PTS_2D_MULTI", "0")
    b = StokesModel(preset, nout=nout, **kw).setup()
    if a.p["ve"]:
        a.ctx.assign_tau()
        b.ctx.assign_tau()
    params = a.pt_params()
    for it_max in (29, 40, 3 * nout + 1):
        ra = a.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        rb = b.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        assert ra["iters"] == rb["iters"] and ra["err"] == rb["err"], (ra, rb)
        for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy"):
            assert np.array_equal(a.get(f), b.get(f)), (f, it_max)
    a.close()
    b.close()


@pytest.mark.parametrize("preset,kw,nout,env", LDS_GEOMETRIES)
@pytest.mark.parametrize("rows", ["0", "3"])
@pytest.mark.parametrize("cols", ["2", "7", "40", "100000"])
@pytest.mark.parametrize("halo", ["0", "1", "3", "4"])
def test_state_in_lds_column_tiles_equal_single_launches(preset, kw, nout, env, rows, cols, halo, monkeypatch):
    # the same with tiles of `cols` columns as well (the last tile of a row
    # of tiles owns the columns left, and those of the tile before when a
    # block's updated columns would start at column nx-1; 100000: the whole
    # width; the test above takes the default, two tiles across grids at
    # least 64 wide): halo positions from up to eight (or, with a halo
    # deeper than a tile, more) blocks
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    monkeypatch.setenv("PTS_2D_LDS_HALO", halo)
    monkeypatch.setenv("PTS_2D_LDS_ROWS", rows)
    monkeypatch.setenv("PTS_2D_LDS_COLS", cols)
    monkeypatch.setenv("PTS_2D_LDS", "1")
    a = StokesModel(preset, nout=nout, **kw).setup()
    monkeypatch.setenv("PTS_2D_LDS", "0")
    monkeypatch.setenv("PTS_2D_MULTI", "0")
    b = StokesModel(preset, nout=nout, **kw).setup()
    if a.p["ve"]:
        a.ctx.assign_tau()
        b.ctx.assign_tau()
    params = a.pt_params()
    for it_max in (29, 40, 3 * nout + 1):
        ra = a.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        rb = b.ctx.pt_solve(params, 1e-300, nout, it_max, 1, use_graph=True)
        assert ra["iters"] == rb["iters"] and ra["err"] == rb["err"], (ra, rb)
        for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy"):
            assert np.array_equal(a.get(f), b.get(f)), (f, it_max)
    a.close()
    b.close()


def test_state_in_lds_block_count_changes_between_launches(monkeypatch):
    # the exchange counters of one context stay consistent when the block
    # count changes between launches (knobs set on the live context)
    kw = dict(n=127, rhog_in=-1.0, ebg=0.0, nout=10)
    monkeypatch.setenv("PTS_2D_LDS", "1")
    a = StokesModel("Stokes2D", **kw).setup()
    monkeypatch.setenv("PTS_2D_LDS", "0")
    monkeypatch.setenv("PTS_2D_MULTI", "0")
    b = StokesModel("Stokes2D", **kw).setup()
    params = a.pt_params()
    for rows, halo in ((3, 3), (7, 2), (3, 3), (0, 0), (2, 1)):
        a.ctx.set_knob("PTS_2D_LDS_ROWS", rows)
        a.ctx.set_knob("PTS_2D_LDS_HALO", halo)
        ra = a.ctx.pt_solve(params, 1e-300, 10, 39, 1, use_graph=True)
        rb = b.ctx.pt_solve(params, 1e-300, 10, 39, 1, use_graph=True)
        assert ra["iters"] == rb["iters"] and ra["err"] == rb["err"], (rows, halo, ra, rb)
        for f in ("Pt", "Vx", "Vy", "txx", "tyy", "txy"):
            assert np.array_equal(a.get(f), b.get(f)), (f, rows, halo)
    a.close()
    b.close()


def test_state_in_lds_wait_that_gives_up_is_reported(monkeypatch):
    # failure handling of the exchange: with the test hook
    # PTS_2D_LDS_FAULT_WAIT=1 every wait targets a generation that no block
    # reaches, so each ends at its bounded spin (1000 polls here); the kernels
    # run to their end and the solve reports the error instead of returning
    # the invalid state
    monkeypatch.setenv("PTS_2D_LDS", "1")
    monkeypatch.setenv("PTS_2D_LDS_FAULT_WAIT", "1")
    monkeypatch.setenv("PTS_2D_LDS_WAIT_SPINS", "1000")
    m = StokesModel("Stokes2D", n=63, nout=10).setup()
    with pytest.raises(PtsError) as e:
        m.ctx.pt_solve(m.pt_params(), 1e-300, 10, 29, 1, use_graph=True)
    assert "gave up" in str(e.value), str(e.value)
    m.close()
    # a context without the hook runs the same solve as one launch per iteration
    monkeypatch.setenv("PTS_2D_LDS_FAULT_WAIT", "0")
    a = StokesModel("Stokes2D", n=63, nout=10).setup()
    monkeypatch.setenv("PTS_2D_LDS", "0")
    monkeypatch.setenv("PTS_2D_MULTI", "0")
    b = StokesModel("Stokes2D", n=63, nout=10).setup()
    ra = a.ctx.pt_solve(a.pt_params(), 1e-300, 10, 29, 1, use_graph=True)
    rb = b.ctx.pt_solve(b.pt_params(), 1e-300, 10, 29, 1, use_graph=True)
    assert ra["iters"] == rb["iters"] and ra["err"] == rb["err"]
    assert np.array_equal(a.get("Vx"), b.get("Vx"))
    a.close()
    b.close()
