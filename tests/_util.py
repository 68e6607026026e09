"""Shared helpers of the test suite."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# preset -> golden file, grid size (test/test_2D.jl:17-33, test/test_3D_{1,2,3}.jl)
FIXTURES = {
    "Stokes2D": ("test_Stokes_2D_v", 63),
    "Stokes2D_ve": ("test_Stokes_2D_ve", 63),
    "Stokes2D_ve_perf": ("test_Stokes_2D_ve_perf", 63),
    "Stokes3D_ve_multixpu": ("test_Stokes3D_ve_multixpu", 31),
    "Stokes3D_ve_multixpu_perf": ("test_Stokes3D_ve_multixpu_perf", 31),
    "Stokes3D_ve_perf": ("test_Stokes3D_ve_perf", 31),
}

# Published total PT iteration counts of converged 5-step VE runs
# (output/out_Stokes2D_ve.txt:1-3, output/out_Stokes3D_ve.txt:1 -- the 60^3
# global grid of 8 ranks x 31^3 is one 60^3 block on one device).
PUBLISHED_ITERS = [("Stokes2D_ve", 31, 3000), ("Stokes2D_ve", 63, 6200), ("Stokes2D_ve", 127, 11200),
                   ("Stokes3D_ve_multixpu", 60, 5600)]


def load_fixture(stem):
    with open(os.path.join(GOLDEN, stem + ".json")) as f:
        d = json.load(f)
    P = np.array([float.fromhex(h) for h in d["P_colmajor_hex"]]).reshape(d["P_shape"], order="F")
    X = np.array([float.fromhex(h) for h in d["X_hex"]])
    idx = [i - 1 for i in d["sample_indices_1based"]]
    return P, X, idx


def sample(P, idx):
    return P[np.ix_(*([idx] * P.ndim))]


def julia_isapprox(a, b):
    """Julia's isapprox for arrays: norm(a-b) <= sqrt(eps)*max(norm(a), norm(b))."""
    return np.linalg.norm(a - b) <= np.sqrt(np.finfo(float).eps) * max(np.linalg.norm(a), np.linalg.norm(b))


def relmax(a, b):
    """max|a-b| / max|b| (per-field relative difference, robust at zero crossings)."""
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
