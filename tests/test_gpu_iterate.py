"""The device-resident iteration (single-pass 2D kernel, in-place 3D pair)
against the oracle's kernel sequence, bit for bit, across sizes that
exercise strip and chunk edges."""
import numpy as np
import pytest

from gpu_cases import STATE, compare, pair, randomise

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
]


@pytest.mark.parametrize("preset,kw,env", VARIANTS, ids=[f"{p}-{list(e)[0]}={list(e.values())[0]}" for p, _, e in VARIANTS])
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
