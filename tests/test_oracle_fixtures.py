"""The oracle against the reference's own golden files and published counts."""
import numpy as np
import pytest

from oracle import Oracle
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
    # X = cell centres (i-1/2)dx at the sampled indices
    dx = 10.0 / n
    assert np.allclose(X, (np.array(idx) + 0.5) * dx, rtol=0, atol=1e-14)


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
