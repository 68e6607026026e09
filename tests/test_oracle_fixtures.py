"""The oracle against the reference's own golden files and published counts."""
import numpy as np
import pytest

from oracle import Oracle
from ptstokes.models import julia_linrange
from _util import FIXTURES, PUBLISHED_ITERS, julia_isapprox, load_fixture, sample


@pytest.mark.parametrize("preset", list(FIXTURES))
def test_oracle_reproduces_golden_file(preset):
    stem, n = FIXTURES[preset]
    P, X, idx = load_fixture(stem)
    o = Oracle(preset, n=n).setup()
    res = o.solve()
    assert res["rc"] == 0
    S = sample(o.get("Pt"), idx)
    # the reference's own comparison (test_2D.jl:14: isapprox, rtol sqrt(eps))
    assert julia_isapprox(S, P)
    # the restatement is exact: same bits
    assert np.array_equal(S, P), f"max |diff| {np.max(np.abs(S - P)):.3e}"
    # X = the script's xc = LinRange(dx/2, lx-dx/2, nx) at the sampled indices, bit for bit
    dx = 10.0 / n
    assert np.array_equal(julia_linrange(dx / 2, 10.0 - dx / 2, n)[idx], X)


def test_fixed_iteration_scripts_run_iterMax_plus_one():
    # nt=1, iterMax=100, nout=2000: the while loop runs iter = 0..100 -> 101 iterations
    o = Oracle("Stokes2D_ve_perf", n=31).setup()
    assert o.solve()["ittot"] == 101
    o = Oracle("Stokes3D_ve_perf", n=15).setup()
    assert o.solve()["ittot"] == 201


@pytest.mark.parametrize("preset,n,want", PUBLISHED_ITERS)
def test_oracle_matches_published_iteration_counts(preset, n, want):
    o = Oracle(preset, n=n, threads=8).setup()
    res = o.solve()
    assert res["ittot"] == want, res
