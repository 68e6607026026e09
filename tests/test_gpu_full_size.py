"""The single-pass iteration kernels at the benchmark's real sizes (2D VE
8191², 3D VE 511³, the inversion's 2D viscous 2047²) against the per-kernel call sequence of the perf scripts
(compute_Pt_τ!, compute_V!, free-slip BCs; Stokes2D_ve_perf.jl:150-160,
Stokes3D_ve_perf.jl:200-207), bit for bit.

The oracle takes minutes per iteration at these sizes, so the chain is: the
per-kernel entry points equal the oracle's kernels bit for bit at small sizes
(test_gpu_kernels.py), they are plain one-launch-per-update kernels whose
indexing does not change with the grid size, and here the single-pass kernels
(tiling, row chunks, 3D field groups with guard zones, placement selection,
ping-pong state) must reproduce them at the sizes the benchmark times.
Odd and even iteration counts cover both ping-pong parities.
"""
import numpy as np
import pytest

from gpu_cases import STATE, valid
from ptstokes import StokesModel

pytestmark = pytest.mark.gpu


# 2D VE 8191² with the benchmark's dt = η/G/10 = 0.1; 3D VE 511³ with the
# script's dt = 1 (Stokes3D_ve_perf.jl:124-126); the inversion's forward
# problem, 2D viscous 2047² with the GRF viscosity (multiply form)
@pytest.mark.parametrize("preset,n,kw", [("Stokes2D_ve_perf", 8191, dict(dt=0.1)), ("Stokes3D_ve_perf", 511, {}),
                                         ("Stokes2D_grf", 2047, {})],
                         ids=["2d-ve-8191-dt0.1", "3d-ve-511", "2d-grf-2047"])
def test_single_pass_equals_per_kernel_at_full_size(preset, n, kw):
    print(f"\nfull-size case {preset} n={n} {kw}", flush=True)
    m1 = StokesModel(preset, n=n, **kw).setup()
    m2 = StokesModel(preset, n=n, **kw).setup()
    try:
        nd = m1.p["ndim"]
        if m1.p["ve"]:
            m1.ctx.assign_tau()  # the physical step starts with τ_o = τ (Stokes2D_ve_perf.jl:146)
            m2.ctx.assign_tau()
        m1.ctx.pt_iterate(m1.pt_params(), 0)  # placement selection, as ahead of a long run
        done = 0
        for k in (3, 1):
            m1.iterate(k)
            m2.iterate_per_kernel(k)
            done += k
            print(f"  {done} iterations, comparing", flush=True)
            for name in STATE:
                if not valid(nd, name):
                    continue
                a, b = m1.get(name), m2.get(name)
                assert np.all(np.isfinite(a)), name
                assert np.array_equal(a, b), (name, done, float(np.max(np.abs(a - b))))
                del a, b
        # the state moved away from the initial condition
        assert np.max(np.abs(m1.get("Pt"))) > 0.0
    finally:
        # free the device memory (tens of GB at 511^3) even when a check fails
        m1.close()
        m2.close()
