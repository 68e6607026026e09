"""The reference scripts' entry points (ptstokes.scripts) against the
reference's own test harness: test/test_2D.jl:17-33 and test/test_3D_*.jl run
each script with USE_RETURN=true at the fixture size and compare Pt at the
sampled indices with the golden file."""
import numpy as np
import pytest
from scipy.io import loadmat

from ptstokes import scripts
from _util import FIXTURES, julia_isapprox, load_fixture, sample

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset", list(FIXTURES))
def test_script_returns_golden_pressure(preset):
    stem, n = FIXTURES[preset]
    P, X, idx = load_fixture(stem)
    ndim = 3 if preset.startswith("Stokes3D") else 2
    env = {"NX": str(n), "NY": str(n), "USE_RETURN": "true"}
    if ndim == 3:
        env["NZ"] = str(n)
    out = scripts.SCRIPTS[preset](env)
    assert len(out) == ndim + 1
    Pt = out[-1]
    S = sample(Pt, idx)
    assert julia_isapprox(S, P)
    assert np.array_equal(S, P)
    assert np.array_equal(out[0][idx], X)


def test_script_without_return_gives_none(capsys):
    assert scripts.Stokes2D_ve_perf({"NX": "31", "NY": "31"}) is None
    assert "Total iters = 101 (1 steps)" in capsys.readouterr().out


def test_do_save_and_do_save_viz_2d(tmp_path):
    env = {"NX": "40", "NY": "33", "DO_SAVE": "true", "DO_SAVE_VIZ": "true", "PTS_OUT_DIR": str(tmp_path)}
    scripts.Stokes2D_ve_perf(env)
    scripts.Stokes2D_ve_perf(env)  # appends
    lines = (tmp_path / "output" / "out_Stokes2D_ve_perf.txt").read_text().splitlines()
    assert len(lines) == 2
    f = lines[0].split()
    assert len(f) == 7 and f[:3] == ["40", "33", "101"]
    a_eff = float(f[4])
    assert a_eff == (6 * 2 + 3 + 3) / 1e9 * 40 * 33 * 8
    assert abs(float(f[6]) - a_eff / float(f[5])) <= 1e-12 * float(f[6])
    m = loadmat(str(tmp_path / "out_visu" / "Stokes_2D_ve_perf.mat"))
    assert m["Pt_2D"].shape == (40, 33) and m["Mus_2D"].shape == (40, 33)
    assert m["Txy_2D"].shape == (39, 32) and m["Vy_2D"].shape == (40, 34)
    assert m["dx_2D"].item() == 10.0 / 40 and m["dy_2D"].item() == 10.0 / 33


def test_do_save_viz_3d_multixpu_inner_arrays(tmp_path):
    env = {"NX": "15", "NY": "14", "NZ": "13", "DO_SAVE": "1", "DO_SAVE_VIZ": "1", "PTS_OUT_DIR": str(tmp_path),
           "USE_RETURN": "1"}
    xc, yc, zc, Pt = scripts.Stokes3D_ve_multixpu_perf(env)
    line = (tmp_path / "output" / "out_Stokes3D_ve_xpu_perf.txt").read_text().split()
    assert line[:5] == ["1", "15", "14", "13", "201"]
    m = loadmat(str(tmp_path / "out_visu" / "Stokes_3D_ve_perf_xpu.mat"))
    for k in ("Pt_3D", "Mus_3D", "Txz_3D", "Vz_3D"):
        assert m[k].shape == (13, 12, 11), k
    assert np.array_equal(m["Pt_3D"], Pt[1:-1, 1:-1, 1:-1])
