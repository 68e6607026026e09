"""Shared set-up of the GPU parity tests: one GPU model and one oracle run of
the same reference script, with identical (optionally randomised) state."""
import numpy as np

from oracle import Oracle
from ptstokes import StokesModel

STATE = ["Pt", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz"]
HIST = ["txx_o", "tyy_o", "tzz_o", "txy_o", "txz_o", "tyz_o"]
SETUP = ["Mus", "Mustau", "Gdtau", "dtau_Rho", "Vx", "Vy", "Vz", "rhog"]


def valid(ndim, name):
    return ndim == 3 or name not in ("Vz", "tzz", "txz", "tyz", "tzz_o", "txz_o", "tyz_o", "Rz", "dVz")


def oracle_kwargs(model):
    p = model.p
    kw = dict(nx=p["nx"], ny=p["ny"], nz=p["nz"], ve=int(p["ve"]), perf=int(p["fd"] == "mul"),
              ic_igg=int(p["ic_igg"]), nt=p["nt"], nout=p["nout"], iter_max=p["iter_max"],
              norm_inf=p["norm_kind"], lx=p["lx"], ly=p["ly"], lz=p["lz"], mus0=p["mus0"], musi=p["musi"],
              rad2=p["rad2"], G=p["G"], dt=p["dt"], ebg=p["ebg"], Re=p["Re"], r=p["r"], CFL=p["CFL"],
              eps=p["eps"], rhog_in=p["rhog_in"], rhog_out=p["rhog_out"])
    return kw


def pair(preset, **kw):
    m = StokesModel(preset, **kw).setup()
    o = Oracle(preset, **oracle_kwargs(m)).setup()
    return m, o


def randomise(m, o, seed, hist=True):
    rng = np.random.default_rng(seed)
    nd = m.p["ndim"]
    names = [n for n in STATE if valid(nd, n)]
    if hist and m.p["ve"]:
        names += [n for n in HIST if valid(nd, n)]
    for n in names:
        a = rng.uniform(-1.0, 1.0, size=o.shape(n))
        o.set(n, a)
        m.ctx.upload(n, a)


def compare(m, o, names, exact=True, rtol=0.0):
    nd = m.p["ndim"]
    bad = []
    for n in names:
        if not valid(nd, n):
            continue
        a, b = m.get(n), o.get(n)
        assert a.shape == b.shape, (n, a.shape, b.shape)
        if exact:
            if not np.array_equal(a, b):
                bad.append((n, float(np.max(np.abs(a - b)))))
        else:
            den = max(float(np.max(np.abs(b))), 1e-300)
            d = float(np.max(np.abs(a - b))) / den
            if d > rtol:
                bad.append((n, d))
    return bad
