"""The device-resident iteration (single-pass 2D kernel, in-place 3D pair)
against the oracle's kernel sequence, bit for bit, across sizes that
exercise strip and chunk edges."""
import numpy as np
import pytest

from gpu_cases import HIST, STATE, compare, pair, randomise

pytestmark = pytest.mark.gpu

CASES_2D = [
    ("Stokes2D_ve_perf", dict(n=5)),
    ("Stokes2D_ve_perf", dict(n=63)),
    ("Stokes2D_ve_perf", dict(nx=125, ny=131)),       # two strips / one partial
    ("Stokes2D_ve_perf", dict(nx=61, ny=300)),        # several row chunks
    ("Stokes2D_ve", dict(nx=200, ny=129, dt=0.1)),
    ("Stokes2D", dict(n=127)),
    ("Stokes2D", dict(nx=130, ny=64, rhog_in=-1.0)),
    ("Stokes2D_ve_perf", dict(n=90, rhog_in=-1.0, rhog_out=0.25)),
]
CASES_3D = [
    ("Stokes3D_ve_perf", dict(n=15)),
    ("Stokes3D_ve_perf", dict(nx=70, ny=9, nz=12)),
    ("Stokes3D_ve_multixpu", dict(n=14)),
    ("Stokes3D_ve_perf", dict(n=12, ve=False, rhog_in=-1.0)),
]


def _ids(cases):
    return [f"{p}-{'-'.join(f'{k}{v}' for k, v in kw.items())}" for p, kw in cases]


@pytest.mark.parametrize("preset,kw", CASES_2D + CASES_3D, ids=_ids(CASES_2D + CASES_3D))
@pytest.mark.parametrize("random", [False, True])
def test_iterate_bit_exact(preset, kw, random):
    m, o = pair(preset, **kw)
    if random:
        randomise(m, o, seed=5)
    for n in (1, 2, 7):  # odd and even counts exercise both ping-pong parities
        m.iterate(n)
        o.iterate(n)
        bad = compare(m, o, STATE)
        assert bad == [], (n, bad)
    m.close()


@pytest.mark.parametrize("preset,kw", [CASES_2D[2], CASES_2D[4], CASES_3D[0], CASES_3D[2]],
                         ids=["2d-mul", "2d-div", "3d-mul", "3d-div"])
def test_per_kernel_api_equals_fused_iteration(preset, kw):
    m1, _ = pair(preset, **kw)
    m2, _ = pair(preset, **kw)
    m1.iterate(9)
    m2.iterate_per_kernel(9)
    for n in STATE:
        if m1.p["ndim"] == 2 and n in ("Vz", "tzz", "txz", "tyz"):
            continue
        assert np.array_equal(m1.get(n), m2.get(n)), n
    m1.close()
    m2.close()


def test_upload_after_iterate_resyncs_dirichlet_values():
    # changing the boundary velocities between iterations must be seen by the
    # ping-pong partner buffer as well
    m, o = pair("Stokes2D_ve_perf", n=40)
    m.iterate(3)
    o.iterate(3)
    Vx = o.get("Vx")
    Vx[0, :] = 7.0
    Vx[-1, :] = -7.0
    o.set("Vx", Vx)
    m.ctx.upload("Vx", Vx)
    m.iterate(3)
    o.iterate(3)
    assert compare(m, o, STATE) == []
    m.close()


@pytest.mark.parametrize("preset,kw", [("Stokes2D_ve_perf", dict(nx=130, ny=61)),
                                       ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19)),
                                       ("Stokes2D", dict(n=70))])
def test_uploaded_iteration_parameters_are_read(preset, kw):
    # After setup the kernels derive Gdτ and dτ_Rho from Musτ; once the user
    # writes them (or Musτ) directly, the stored arrays must be used.
    m, o = pair(preset, **kw)
    randomise(m, o, seed=3)
    m.iterate(2)
    o.iterate(2)
    assert compare(m, o, STATE) == []
    rng = np.random.default_rng(4)
    # μ changed without recomputing Musτ: Musτ is no longer maxloc(μ)
    mu = o.get("Mus") * rng.uniform(0.5, 2.0, size=o.shape("Mus"))
    o.set("Mus", mu)
    m.ctx.upload("Mus", mu)
    m.iterate(3)
    o.iterate(3)
    assert compare(m, o, STATE) == []
    for f, lo, hi in (("Gdtau", 0.05, 0.5), ("dtau_Rho", 0.001, 0.02)):
        a = rng.uniform(lo, hi, size=o.shape(f))
        o.set(f, a)
        m.ctx.upload(f, a)
    m.iterate(3)
    o.iterate(3)
    assert compare(m, o, STATE) == []
    mt = o.get("Mustau") * rng.uniform(0.5, 2.0, size=o.shape("Mustau"))
    o.set("Mustau", mt)
    m.ctx.upload("Mustau", mt)  # Musτ no longer matches Gdτ / dτ_Rho
    m.iterate(3)
    o.iterate(3)
    assert compare(m, o, STATE) == []
    m.close()


def test_cached_graphs_follow_new_scalars():
    # pt_solve replays captured graphs; a call with other scalars (or after the
    # iteration parameters changed) must not replay the old ones
    from ptstokes import Context
    a, _ = pair("Stokes2D_ve_perf", n=50)
    b, _ = pair("Stokes2D_ve_perf", n=50)
    p1 = a.pt_params()
    p2 = Context.pt_params(0.7, 1.0, 0.5, a.dx, a.dy)
    for params in (p1, p2, p1):
        a.ctx.pt_solve(params, 1e-300, 10, 29, 0, use_graph=True)  # 30 iterations, no convergence
        b.ctx.pt_iterate(params, 30)
    a.ctx.iter_params(a.Vpdtau * 0.9, 1.0, 1.0, a.p["Re"], a.p["r"], a.max_l)
    b.ctx.iter_params(b.Vpdtau * 0.9, 1.0, 1.0, b.p["Re"], b.p["r"], b.max_l)
    a.ctx.pt_solve(p1, 1e-300, 10, 29, 0, use_graph=True)
    b.ctx.pt_iterate(p1, 30)
    for f in STATE:
        if f in ("Vz", "tzz", "txz", "tyz"):
            continue
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("dim", ["3D", "2D"])
@pytest.mark.parametrize("accept", ["1", "0"])
def test_moved_blocks_keep_state_and_graphs(accept, dim, monkeypatch):
    # a context whose field blocks are timed with real iterations on copies
    # and moved (or tried and moved back) before its first solve iterates
    # exactly like one that keeps its blocks, through captured graphs, the
    # per-call path and later downloads
    tries = f"PTS_{dim}_PLACE_TRIES"
    monkeypatch.setenv("PTS_PLACE_MIN_MB", "0")
    monkeypatch.setenv("PTS_PLACE_ACCEPT", accept)
    monkeypatch.setenv("PTS_PLACE_BUDGET_ITERS", "30")  # the 30-iteration solve below is a long run
    monkeypatch.setenv(tries, "3")
    if dim == "3D":
        case = ("Stokes3D_ve_perf", dict(nx=45, ny=26, nz=21))
        fields = STATE + HIST + ["Mus"]
    else:
        case = ("Stokes2D_ve_perf", dict(nx=300, ny=170))
        fields = ["Pt", "Vx", "Vy", "txx", "tyy", "txy", "txx_o", "tyy_o", "txy_o", "Mus"]
    a, _ = pair(case[0], **case[1])
    monkeypatch.setenv(tries, "0")
    b, _ = pair(case[0], **case[1])
    params = a.pt_params()
    a.ctx.pt_solve(params, 1e-300, 10, 29, 0, use_graph=True)  # 30 iterations, no convergence
    b.ctx.pt_iterate(params, 30)
    assert len(a.ctx.placement_times()) == 4 and b.ctx.placement_times() == []
    assert a.ctx.device_bytes() == b.ctx.device_bytes()
    for f in fields:
        assert np.array_equal(a.get(f), b.get(f)), f
    a.ctx.assign_tau()
    b.ctx.assign_tau()
    a.ctx.pt_solve(params, 1e-300, 10, 19, 0, use_graph=True)
    a.ctx.pt_iterate(params, 3)
    b.ctx.pt_iterate(params, 23)
    for f in fields:
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("preset,kw", [("Stokes3D_ve_perf", dict(nx=140, ny=33, nz=20)),
                                       ("Stokes3D_ve_multixpu", dict(nx=125, ny=21, nz=17, rhog_in=-1.0)),
                                       ("Stokes3D_ve_perf", dict(nx=63, ny=41, nz=14, ve=False)),
                                       ("Stokes2D_ve_perf", dict(nx=1100, ny=75)),
                                       ("Stokes2D", dict(nx=600, ny=130, rhog_in=-1.0)),
                                       ("Stokes2D_ve", dict(nx=509, ny=67, dt=0.1))])
@pytest.mark.parametrize("variant", ["1", "2"])
def test_formed_mustau_rough_viscosity(preset, kw, variant, monkeypatch):
    # A rough μ (every cross maximum may sit at any neighbour, also across the
    # 62-column tile edges) with Musτ = maxloc(μ) + the free-slip copies, so
    # that the kernels form Musτ from μ; the oracle reads the same Musτ, Gdτ
    # and dτ_Rho.  3D: without (1) and with (2) the next plane's loads issued
    # before the arithmetic.
    if preset.startswith("Stokes2D") and variant == "2":
        pytest.skip("3D kernel variant")
    monkeypatch.setenv("PTS_3D_VARIANT", variant)
    m, o = pair(preset, **kw)
    rng = np.random.default_rng(11)
    mu = np.exp(rng.normal(0.0, 1.5, size=o.shape("Mus")))
    o.set("Mus", mu)
    m.ctx.upload("Mus", mu)
    m.ctx.maxloc("Mustau", "Mus")
    for bc in (m.ctx.bc_x, m.ctx.bc_y, m.ctx.bc_z)[:m.p["ndim"]]:
        bc("Mustau")
    p = m.p
    m.ctx.iter_params(m.Vpdtau, p["G"], p["dt"], p["Re"], p["r"], m.max_l)
    for f in ("Mustau", "Gdtau", "dtau_Rho"):
        o.set(f, m.get(f))
    randomise(m, o, seed=7)
    for n in (1, 2, 5):
        m.iterate(n)
        o.iterate(n)
        assert compare(m, o, STATE) == [], n
    m.close()


VARIANTS = [
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_3D_TWO_KERNEL": "1"}),
    ("Stokes3D_ve_multixpu", dict(n=21), {"PTS_3D_TWO_KERNEL": "1"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_3D_TILE_ROWS": "8"}),
    ("Stokes3D_ve_perf", dict(nx=33, ny=40, nz=12), {"PTS_3D_TILE_ROWS": "16"}),
    ("Stokes3D_ve_perf", dict(nx=20, ny=17, nz=19), {"PTS_PLANES_PER_BLOCK": "3"}),
    ("Stokes3D_ve_multixpu", dict(n=16, rhog_in=-1.0), {"PTS_PLANES_PER_BLOCK": "1"}),
    ("Stokes2D_ve_perf", dict(nx=90, ny=41), {"PTS_ROWS_PER_BLOCK": "5"}),
    ("Stokes2D", dict(nx=64, ny=33), {"PTS_ROWS_PER_BLOCK": "1"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=290), {"PTS_ROWS_PER_BLOCK": "128"}),
    ("Stokes2D_ve_perf", dict(nx=130, ny=70), {"PTS_ROWS_PER_BLOCK": "0"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=80), {"PTS_PLANES_PER_BLOCK": "64"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=30, nz=50), {"PTS_PLANES_PER_BLOCK": "-16"}),
    ("Stokes2D_ve_perf", dict(nx=130, ny=70), {"PTS_2D_VARIANT": "0"}),
    ("Stokes2D", dict(nx=70, ny=41), {"PTS_2D_VARIANT": "2"}),
    ("Stokes2D_ve", dict(nx=70, ny=41), {"PTS_2D_VARIANT": "3"}),
    ("Stokes2D_ve_perf", dict(nx=600, ny=77), {"PTS_2D_VARIANT": "4"}),
    ("Stokes2D_ve_perf", dict(nx=20, ny=33), {"PTS_2D_VARIANT": "4"}),
    ("Stokes2D", dict(nx=253, ny=40, rhog_in=-1.0), {"PTS_2D_VARIANT": "4"}),
    ("Stokes2D_ve", dict(nx=509, ny=31), {"PTS_2D_VARIANT": "4", "PTS_ROWS_PER_BLOCK": "7"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=45), {"PTS_2D_VARIANT": "5"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=290), {"PTS_2D_PERSIST": "1"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=45), {"PTS_2D_VARIANT": "1"}),
    ("Stokes2D", dict(nx=1100, ny=45, rhog_in=-1.0), {"PTS_2D_VARIANT": "7"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=45), {"PTS_DERIVED_PARAMS": "0"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=45), {"PTS_DERIVED_MUSTAU": "0"}),
    ("Stokes2D_ve_perf", dict(nx=1100, ny=75), {"PTS_2D_VARIANT": "11"}),
    ("Stokes2D_ve_perf", dict(nx=1019, ny=100), {"PTS_ROWS_PER_BLOCK": "7"}),
    ("Stokes2D", dict(nx=509, ny=66, rhog_in=-1.0), {"PTS_ROWS_PER_BLOCK": "5"}),
    ("Stokes2D_ve", dict(nx=1021, ny=9), {"PTS_ROWS_PER_BLOCK": "2"}),
    ("Stokes2D_ve_perf", dict(nx=1021, ny=37), {"PTS_2D_VARIANT": "10"}),
    ("Stokes2D", dict(nx=511, ny=20, rhog_in=-1.0), {"PTS_ROWS_PER_BLOCK": "-4"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_DERIVED_PARAMS": "0"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_STAGGER_BYTES": "0"}),
    ("Stokes3D_ve_multixpu", dict(n=17), {"PTS_STAGGER_BYTES": "65792"}),
    # 3D kernels: 0 the predicated k3_iter, 1 k3_iter2, 2 k3_iter2 with prefetch;
    # G·dt = 0.3 (divide) and 0.25 (power of two, multiply by the reciprocal)
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_3D_VARIANT": "0"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_perf", dict(nx=125, ny=31, nz=9, dt=0.3), {"PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_perf", dict(nx=125, ny=31, nz=9, dt=0.3), {"PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_perf", dict(nx=61, ny=42, nz=11, G=0.5, dt=0.5), {"PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_multixpu", dict(n=19, rhog_in=-1.0), {"PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_multixpu", dict(nx=33, ny=20, nz=27, rhog_in=-1.0), {"PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_perf", dict(nx=62, ny=10, nz=30, ve=False), {"PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_perf", dict(nx=63, ny=11, nz=30, ve=False, rhog_in=-1.0), {"PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19), {"PTS_3D_VARIANT": "2", "PTS_DERIVED_PARAMS": "0"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19), {"PTS_3D_VARIANT": "1", "PTS_DERIVED_PARAMS": "0"}),
    ("Stokes3D_ve_perf", dict(nx=20, ny=17, nz=19), {"PTS_3D_VARIANT": "2", "PTS_PLANES_PER_BLOCK": "1"}),
    ("Stokes3D_ve_perf", dict(nx=20, ny=17, nz=19), {"PTS_3D_VARIANT": "2", "PTS_PLANES_PER_BLOCK": "2"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=80), {"PTS_3D_VARIANT": "2", "PTS_PLANES_PER_BLOCK": "64"}),
    # extent 3 in z / y: the single interior plane feeds both free-slip copies
    ("Stokes3D_ve_perf", dict(nx=5, ny=4, nz=3), {"PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_perf", dict(nx=6, ny=3, nz=4), {"PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_perf", dict(nx=5, ny=4, nz=3), {"PTS_3D_VARIANT": "0"}),
    ("Stokes3D_ve_multixpu", dict(nx=4, ny=3, nz=5), {"PTS_3D_TWO_KERNEL": "1"}),
    # padded plane stride; separate (ungrouped) 3D allocations (k3_iter runs)
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_PLANE_PAD": "32"}),
    ("Stokes3D_ve_multixpu", dict(n=15, rhog_in=-1.0), {"PTS_PLANE_PAD": "544", "PTS_3D_VARIANT": "1"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19), {"PTS_3D_GROUPS": "0"}),
    # Musτ read instead of formed from μ (3D); the formed case runs in the default ones
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12), {"PTS_DERIVED_MUSTAU": "0"}),
    ("Stokes3D_ve_multixpu", dict(n=13, rhog_in=-1.0), {"PTS_DERIVED_MUSTAU": "0"}),
    ("Stokes3D_ve_perf", dict(nx=3, ny=5, nz=4), {}),
    ("Stokes3D_ve_perf", dict(nx=140, ny=30, nz=45), {"PTS_3D_MAP_FLAT": "1"}),
    # per-slot 64-bit pointers (forced here; chosen by themselves where the
    # slots lie more than 4 GiB apart, e.g. at 1023^3)
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=30), {"PTS_3D_W64": "1"}),
    ("Stokes3D_ve_perf", dict(nx=125, ny=31, nz=9, dt=0.3), {"PTS_3D_W64": "1"}),
    ("Stokes3D_ve_multixpu", dict(n=19, rhog_in=-1.0), {"PTS_3D_W64": "1"}),
    ("Stokes3D_ve_perf", dict(nx=63, ny=11, nz=30, ve=False, rhog_in=-1.0), {"PTS_3D_W64": "1"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19), {"PTS_3D_W64": "1", "PTS_DERIVED_MUSTAU": "0"}),
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=80), {"PTS_3D_W64": "1", "PTS_PLANES_PER_BLOCK": "64"}),
    ("Stokes3D_ve_perf", dict(nx=5, ny=4, nz=3), {"PTS_3D_W64": "1"}),
    # one pointer per field with 32-bit slot offsets (the per-slot form is the default)
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=30), {"PTS_3D_W64": "0"}),
    ("Stokes3D_ve_multixpu", dict(n=19, rhog_in=-1.0), {"PTS_3D_W64": "0"}),
    ("Stokes3D_ve_perf", dict(nx=63, ny=11, nz=30, ve=False, rhog_in=-1.0), {"PTS_3D_W64": "0", "PTS_3D_VARIANT": "2"}),
    ("Stokes3D_ve_perf", dict(nx=125, ny=22, nz=30, ve=False), {"PTS_PLANES_PER_BLOCK": "5"}),
    # placement selection moved into every size: blocks moved on every try,
    # never, or when faster
    ("Stokes3D_ve_perf", dict(nx=70, ny=23, nz=12),
     {"PTS_PLACE_MIN_MB": "0", "PTS_PLACE_ACCEPT": "1", "PTS_3D_PLACE_TRIES": "2", "PTS_PLACE_BUDGET_ITERS": "1"}),
    ("Stokes3D_ve_multixpu", dict(n=15, rhog_in=-1.0),
     {"PTS_PLACE_MIN_MB": "0", "PTS_PLACE_ACCEPT": "0", "PTS_PLACE_BUDGET_ITERS": "1"}),
    ("Stokes3D_ve_perf", dict(nx=40, ny=23, nz=19), {"PTS_PLACE_MIN_MB": "0", "PTS_PLACE_BUDGET_ITERS": "1"}),
    ("Stokes2D_ve_perf", dict(nx=300, ny=45),
     {"PTS_PLACE_MIN_MB": "0", "PTS_2D_PLACE_TRIES": "2", "PTS_PLACE_BUDGET_ITERS": "1"}),
    ("Stokes2D", dict(nx=253, ny=40, rhog_in=-1.0),
     {"PTS_PLACE_MIN_MB": "0", "PTS_2D_PLACE_TRIES": "2", "PTS_PLACE_ACCEPT": "1", "PTS_PLACE_BUDGET_ITERS": "1"}),
]


@pytest.mark.parametrize("preset,kw,env", VARIANTS,
                         ids=[f"{p}-" + "-".join(f"{k}={v}" for k, v in e.items()) + "-" + "-".join(f"{k}{v}" for k, v in kw.items())
                              for p, kw, e in VARIANTS])
def test_launch_variants_bit_exact(preset, kw, env, monkeypatch):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    m, o = pair(preset, **kw)
    randomise(m, o, seed=17)
    for n in (1, 4):
        m.iterate(n)
        o.iterate(n)
        assert compare(m, o, STATE) == [], n
    m.close()
