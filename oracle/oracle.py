"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.  It loads ``oracle/libpts_oracle.so`` (built by ``make -C oracle``)
and exposes one :class:`Oracle` per reference script run.  The presets below
copy each script's hard-coded constants (file:line cited per preset).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpts_oracle.so")

FIELDS = [
    "Pt", "Vx", "Vy", "Vz",
    "txx", "tyy", "tzz", "txy", "txz", "tyz",
    "txx_o", "tyy_o", "tzz_o", "txy_o", "txz_o", "tyz_o",
    "Mus", "Mus2", "Mustau", "Gdtau", "dtau_Rho",
    "divV", "Rx", "Ry", "Rz", "dVx", "dVy", "dVz", "rhog",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}


class _Params(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32), ("ve", ctypes.c_int32), ("perf", ctypes.c_int32),
        ("ic_igg", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32), ("nt", ctypes.c_int32), ("nout", ctypes.c_int32),
        ("iter_max", ctypes.c_int64), ("nsmooth", ctypes.c_int32), ("norm_inf", ctypes.c_int32),
        ("threads", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("mus0", ctypes.c_double), ("musi", ctypes.c_double), ("rad2", ctypes.c_double),
        ("G", ctypes.c_double), ("dt", ctypes.c_double), ("ebg", ctypes.c_double),
        ("Re", ctypes.c_double), ("r", ctypes.c_double), ("CFL", ctypes.c_double),
        ("eps", ctypes.c_double), ("rhog_in", ctypes.c_double), ("rhog_out", ctypes.c_double),
    ]


_COMMON = dict(lx=10.0, ly=10.0, lz=10.0, mus0=1.0, musi=1e-3, rad2=1.0, ebg=1.0, eps=1e-8,
               nsmooth=10, norm_inf=0, threads=0, rhog_in=0.0, rhog_out=0.0, nz=1, ic_igg=0)

# Reference presets.  G*dt: ξ=1, dt = μs0/(G*ξ) = 1 in every VE script.
PRESETS = {
    # scripts/Stokes2D.jl:74-84
    "Stokes2D": dict(_COMMON, ndim=2, ve=0, perf=0, nt=1, iter_max=100000, nout=500,
                     CFL=0.9 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0),
    # scripts/Stokes2D_ve.jl:89-103
    "Stokes2D_ve": dict(_COMMON, ndim=2, ve=1, perf=0, nt=5, iter_max=200000, nout=200,
                        CFL=0.8 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0),
    # scripts/Stokes2D_ve_perf.jl:82-97 (nt=1, iterMax=100, nout=2000 -> 101 iterations)
    "Stokes2D_ve_perf": dict(_COMMON, ndim=2, ve=1, perf=1, nt=1, iter_max=100, nout=2000,
                             CFL=0.8 / math.sqrt(2), Re=5 * math.pi, r=1.0, G=1.0, dt=1.0),
    # scripts/stokes_3D/Stokes3D_ve_perf.jl:118-132 (201 iterations)
    "Stokes3D_ve_perf": dict(_COMMON, ndim=3, ve=1, perf=1, nt=1, iter_max=200, nout=2000,
                             CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0, G=1.0, dt=1.0),
    # scripts/stokes_3D/Stokes3D_ve_multixpu.jl:120-134 (one rank)
    "Stokes3D_ve_multixpu": dict(_COMMON, ndim=3, ve=1, perf=0, ic_igg=1, nt=5, iter_max=200000,
                                 nout=200, CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0,
                                 G=1.0, dt=1.0),
    # scripts/stokes_3D/Stokes3D_ve_multixpu_perf.jl (one rank)
    "Stokes3D_ve_multixpu_perf": dict(_COMMON, ndim=3, ve=1, perf=1, ic_igg=1, nt=1, iter_max=200,
                                      nout=2000, CFL=0.8 / math.sqrt(3), Re=6.0 * math.pi, r=1.0,
                                      G=1.0, dt=1.0),
}

# 1D dispersion optimum (dispersion_analysis/dispersion_analysis_stokes1D.ipynb,
# output of cell 5; .m:34-35): r = 1/4, Re = 9*sqrt(3)*pi/4.
R_OPT = 0.25
RE_OPT = 9.0 * math.sqrt(3.0) * math.pi / 4.0

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        lib.or_create.restype = ctypes.c_void_p
        lib.or_create.argtypes = [P(_Params)]
        lib.or_destroy.argtypes = [ctypes.c_void_p]
        lib.or_setup.argtypes = [ctypes.c_void_p]
        lib.or_field_len.restype = ctypes.c_int64
        lib.or_field_len.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.or_field_shape.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_int64)]
        lib.or_get.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double)]
        lib.or_set.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double)]
        lib.or_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        lib.or_assign_tau.argtypes = [ctypes.c_void_p]
        lib.or_compute_res.argtypes = [ctypes.c_void_p]
        lib.or_check.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        lib.or_solve.restype = ctypes.c_int
        lib.or_solve.argtypes = [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_double), P(ctypes.c_int64)]
        lib.or_scalars.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        lib.or_kernel.restype = ctypes.c_int
        lib.or_kernel.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        _lib = lib
    return _lib


def params(preset, **overrides):
    d = dict(PRESETS[preset])
    d.update(overrides)
    if "n" in d:
        n = d.pop("n")
        d["nx"] = d["ny"] = n
        if d["ndim"] == 3:
            d["nz"] = n
    return d


class Oracle:
    """One run of a reference script on the CPU."""

    def __init__(self, preset="Stokes2D_ve_perf", **overrides):
        self.lib = load()
        self.p = params(preset, **overrides)
        cp = _Params()
        for k, v in self.p.items():
            setattr(cp, k, v)
        self._cp = cp
        self.h = self.lib.or_create(ctypes.byref(cp))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.or_destroy(h)
            self.h = None

    def shape(self, name):
        s = (ctypes.c_int64 * 3)()
        self.lib.or_field_shape(self.h, FIELD_ID[name], s)
        if self.p["ndim"] == 2:
            return (int(s[0]), int(s[1]))
        return (int(s[0]), int(s[1]), int(s[2]))

    def get(self, name):
        shp = self.shape(name)
        out = np.empty(int(np.prod(shp)), dtype=np.float64)
        self.lib.or_get(self.h, FIELD_ID[name], out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(shp, order="F")

    def set(self, name, arr):
        shp = self.shape(name)
        a = np.asarray(arr, dtype=np.float64)
        assert a.shape == shp, (a.shape, shp)
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        self.lib.or_set(self.h, FIELD_ID[name], flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def setup(self):
        self.lib.or_setup(self.h)
        return self

    def iterate(self, n=1):
        self.lib.or_iterate(self.h, int(n))

    def assign_tau(self):
        self.lib.or_assign_tau(self.h)

    def compute_res(self):
        self.lib.or_compute_res(self.h)

    KERNELS = {"compute_P": 0, "compute_tau": 1, "compute_dV": 2, "compute_V": 3, "compute_Res": 4,
               "compute_Pt_tau": 5, "compute_V_perf": 6, "compute_Res_perf": 7, "bcs_V": 8,
               "bc_x": 10, "bc_y": 11, "bc_z": 12}

    def kernel(self, name, field=None):
        arg = FIELD_ID[field] if field is not None else 0
        rc = self.lib.or_kernel(self.h, self.KERNELS[name], arg)
        if rc != 0:
            raise ValueError(name)

    def check(self):
        out = (ctypes.c_double * 9)()
        self.lib.or_check(self.h, out)
        keys = ["norm_Rx", "norm_Ry", "norm_Rz", "norm_divV", "err", "Vmin", "Vmax", "Pmin", "Pmax"]
        return dict(zip(keys, list(out)))

    def solve(self):
        it = ctypes.c_int64()
        err = ctypes.c_double()
        per = (ctypes.c_int64 * max(1, self.p["nt"]))()
        rc = self.lib.or_solve(self.h, ctypes.byref(it), ctypes.byref(err), per)
        return dict(rc=rc, ittot=it.value, err=err.value, iters_per_step=list(per))

    def scalars(self):
        out = (ctypes.c_double * 8)()
        self.lib.or_scalars(self.h, out)
        return dict(zip(["dx", "dy", "dz", "_dx", "_dy", "_dz", "max_l", "Vpdtau"], list(out)))
