"""The convergence measure err (Stokes2D_ve.jl:162-171, Stokes3D_ve_perf.jl:209-219)
and its max-norm form, pinned by hand-computed values on planted fields.

The residual arrays carry a constant with one spike, so the RMS and the max
norm differ and a swapped denominator (ΔP for momentum, ΔV for continuity), a
missing lx or √len, or a sum in place of a maximum gives a different number.
Both the oracle's check and the library's host-side pts_err_from_norms (which
the device solve calls after its reductions) are held to the same values.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import Oracle
from ptstokes import _lib

LX = 10.0
DP = 4.0   # Pt spans [-1.5, 2.5]
DV = 8.0   # Vx spans [-3, 5]


def planted(o, fields):
    """Fill the named residual arrays with `base` plus one `spike`, Pt and Vx with known ranges."""
    P = np.zeros(o.shape("Pt"))
    P.flat[0], P.flat[-1] = -1.5, 2.5
    o.set("Pt", P)
    V = np.ones(o.shape("Vx"))
    V.flat[1], V.flat[-2] = -3.0, 5.0
    o.set("Vx", V)
    lens = {}
    for name, (base, spike) in fields.items():
        A = np.full(o.shape(name), base)
        A.flat[A.size // 2] = spike
        o.set(name, A)
        lens[name] = A.size
    return lens


def expected(fields, lens, kind):
    out = {}
    for name, (base, spike) in fields.items():
        den = DV if name == "divV" else DP
        L = lens[name]
        if kind == 1:
            out[name] = max(abs(base), abs(spike)) / den * LX
        else:
            out[name] = math.sqrt((spike ** 2 + (L - 1) * base ** 2) / L) / den * LX
    return out


CASES = {
    2: {"Rx": (0.2, -0.9), "Ry": (-0.05, 0.3), "divV": (0.4, 0.1)},
    3: {"Rx": (0.2, -0.9), "Ry": (-0.05, 0.3), "Rz": (0.01, 1.7), "divV": (0.4, 0.1)},
}
PRESET = {2: ("Stokes2D_ve_perf", 15), 3: ("Stokes3D_ve_perf", 9)}


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_oracle_err_on_planted_residuals(ndim, kind):
    preset, n = PRESET[ndim]
    o = Oracle(preset, n=n, norm_inf=kind).setup()
    fields = CASES[ndim]
    lens = planted(o, fields)
    want = expected(fields, lens, kind)
    got = o.check()
    for name, key in (("Rx", "norm_Rx"), ("Ry", "norm_Ry"), ("Rz", "norm_Rz"), ("divV", "norm_divV")):
        if name in want:
            assert got[key] == pytest.approx(want[name], rel=1e-14), name
    assert got["err"] == pytest.approx(max(want.values()), rel=1e-14)
    assert (got["Pmin"], got["Pmax"], got["Vmin"], got["Vmax"]) == (-1.5, 2.5, -3.0, 5.0)


def test_oracle_err_is_nan_when_a_residual_is():
    o = Oracle("Stokes2D_ve_perf", n=15).setup()
    planted(o, {"Rx": (0.2, float("nan")), "Ry": (0.1, 0.1), "divV": (0.1, 0.1)})
    assert math.isnan(o.check()["err"])


def lib_err(fields, lens, ndim, kind, nan=False):
    lib = _lib.load()
    nr = _lib.Norms()
    nr.Vmin, nr.Vmax, nr.Pmin, nr.Pmax = -3.0, 5.0, -1.5, 2.5
    for q, name in enumerate(("Rx", "Ry", "Rz", "divV")):
        if name not in fields:
            continue
        base, spike = fields[name]
        L = lens[name]
        nr.sumsq[q] = spike ** 2 + (L - 1) * base ** 2
        nr.maxabs[q] = max(abs(base), abs(spike))
        nr.len[q] = L
    if nan:
        nr.sumsq[0] = float("nan")
        nr.maxabs[0] = float("nan")
    comp = (ctypes.c_double * 4)()
    err = lib.pts_err_from_norms(ctypes.byref(nr), ndim, LX, kind, comp)
    return err, list(comp)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_library_err_from_norms(ndim, kind):
    # host-only entry point: runs without a GPU
    fields = CASES[ndim]
    lens = {"Rx": 1000, "Ry": 777, "Rz": 555, "divV": 999}
    want = expected(fields, lens, kind)
    err, comp = lib_err(fields, lens, ndim, kind)
    for q, name in enumerate(("Rx", "Ry", "Rz", "divV")):
        if name in want:
            assert comp[q] == pytest.approx(want[name], rel=1e-14), name
        else:
            assert comp[q] == 0.0
    assert err == pytest.approx(max(want.values()), rel=1e-14)


@pytest.mark.parametrize("kind", [0, 1])
def test_library_err_is_nan_when_a_norm_is(kind):
    err, _ = lib_err(CASES[2], {"Rx": 10, "Ry": 10, "divV": 10}, 2, kind, nan=True)
    assert math.isnan(err)
