"""Python mirror of the reference's per-kernel call surface over libpts.

One :class:`Context` holds the device arrays of one reference script run
(the script's ``@zeros`` arrays) and exposes each ParallelStencil kernel of
the reference under its own name (``compute_P``, ``compute_tau``, ``bc_x`` ...
for ``compute_P!``, ``compute_τ!``, ``bc_x!``), with the same scalar
arguments.  Arrays cross the boundary as NumPy arrays in the reference's
column-major layout and shapes.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD_ID, ALIASES, PtsError


def _fid(name):
    if isinstance(name, int):
        return name
    name = ALIASES.get(name, name)
    return FIELD_ID[name]


class Context:
    def __init__(self, ndim, nx, ny, nz=1, ve=True, fd_mode="mul", body_force=False, device=0):
        self.lib = _lib.load()
        self.ndim = int(ndim)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz) if ndim == 3 else 1
        self.ve = bool(ve)
        self.fd_mode = fd_mode
        self.body_force = bool(body_force)
        g = _lib.Grid(self.ndim, self.nx, self.ny, self.nz, 1 if ve else 0,
                      _lib.PTS_FD_MUL if fd_mode == "mul" else _lib.PTS_FD_DIV,
                      1 if body_force else 0, int(device))
        h = ctypes.c_void_p()
        rc = self.lib.pts_create(ctypes.byref(g), ctypes.byref(h))
        if rc != 0:
            raise PtsError(rc, self.lib.pts_last_error(None).decode())
        self.h = h

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.pts_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise PtsError(rc, self.lib.pts_last_error(self.h).decode())
        return rc

    # -- fields -------------------------------------------------------------
    def shape(self, name):
        s = (ctypes.c_int64 * 3)()
        self._chk(self.lib.pts_field_shape(self.h, _fid(name), s))
        return tuple(int(v) for v in s[: self.ndim])

    def upload(self, name, arr):
        shp = self.shape(name)
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape} != {shp}")
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        self._chk(self.lib.pts_upload(self.h, _fid(name), flat.ctypes.data_as(_lib._dp), flat.size))

    def download(self, name):
        shp = self.shape(name)
        out = np.empty(int(np.prod(shp)), dtype=np.float64)
        self._chk(self.lib.pts_download(self.h, _fid(name), out.ctypes.data_as(_lib._dp), out.size))
        return out.reshape(shp, order="F")

    def fill(self, name, value):
        self._chk(self.lib.pts_fill(self.h, _fid(name), float(value)))

    def copy(self, dst, src):
        self._chk(self.lib.pts_copy(self.h, _fid(dst), _fid(src)))

    def swap(self, a, b):
        self._chk(self.lib.pts_swap(self.h, _fid(a), _fid(b)))

    def sync(self):
        self._chk(self.lib.pts_sync(self.h))

    def stream(self):
        s = ctypes.c_void_p()
        self._chk(self.lib.pts_get_stream(self.h, ctypes.byref(s)))
        return s.value

    def device_bytes(self):
        b = ctypes.c_int64()
        self._chk(self.lib.pts_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def placement_times(self):
        """Iteration times (ms) on the block placements tried: [original, try 1, ...]."""
        ms = (ctypes.c_float * 32)()
        self._chk(self.lib.pts_placement_times(self.h, ms, 32))
        return [round(float(x), 4) for x in ms if x > 0]

    def set_knob(self, name, value):
        """Launch-configuration knob of the single-pass kernels (the name of
        the environment variable that sets it at creation)."""
        self._chk(self.lib.pts_set_knob(self.h, name.encode(), int(value)))

    # -- setup kernels -------------------------------------------------------
    def init_inclusion(self, lx, ly, lz=10.0, mus0=1.0, musi=1e-3, rad2=1.0, ebg=1.0, rhog_in=0.0,
                       rhog_out=0.0, igg=False):
        self._chk(self.lib.pts_init_inclusion(self.h, lx, ly, lz, mus0, musi, rad2, ebg, rhog_in, rhog_out,
                                              1 if igg else 0))

    def init_grf_viscosity(self, modes, sigma, lo, hi, lx, ly, lz=10.0, musi=1e-3, rad2=1.0):
        m = np.ascontiguousarray(np.asarray(modes, dtype=np.float64).reshape(-1, 4))
        self._chk(self.lib.pts_init_grf_viscosity(self.h, m.shape[0], m.ctypes.data_as(_lib._dp), sigma, lo, hi,
                                                  lx, ly, lz, musi, rad2))

    def smooth(self, dst, src, fact):
        self._chk(self.lib.pts_smooth(self.h, _fid(dst), _fid(src), float(fact)))

    def maxloc(self, dst, src):
        self._chk(self.lib.pts_maxloc(self.h, _fid(dst), _fid(src)))

    def bc_x(self, name):
        self._chk(self.lib.pts_bc_x(self.h, _fid(name)))

    def bc_y(self, name):
        self._chk(self.lib.pts_bc_y(self.h, _fid(name)))

    def bc_z(self, name):
        self._chk(self.lib.pts_bc_z(self.h, _fid(name)))

    def iter_params(self, Vpdtau, G, dt, Re, r, max_l):
        self._chk(self.lib.pts_iter_params(self.h, Vpdtau, G, dt, Re, r, max_l))

    def assign_tau(self):
        self._chk(self.lib.pts_assign_tau(self.h))

    # -- 4-kernel form --------------------------------------------------------
    def compute_P(self, r, dx, dy, dz=1.0):
        self._chk(self.lib.pts_compute_P(self.h, r, dx, dy, dz))

    def compute_tau(self, G, dt, dx, dy, dz=1.0):
        self._chk(self.lib.pts_compute_tau(self.h, G, dt, dx, dy, dz))

    def compute_dV(self, dx, dy, dz=1.0):
        self._chk(self.lib.pts_compute_dV(self.h, dx, dy, dz))

    def compute_V(self):
        self._chk(self.lib.pts_compute_V(self.h))

    def compute_Res(self, dx, dy, dz=1.0):
        self._chk(self.lib.pts_compute_Res(self.h, dx, dy, dz))

    # -- fused perf form --------------------------------------------------------
    def compute_Pt_tau(self, r, G, dt, _dx, _dy, _dz=1.0):
        self._chk(self.lib.pts_compute_Pt_tau(self.h, r, G, dt, _dx, _dy, _dz))

    def compute_V_perf(self, _dx, _dy, _dz=1.0):
        self._chk(self.lib.pts_compute_V_perf(self.h, _dx, _dy, _dz))

    def compute_Res_perf(self, _dx, _dy, _dz=1.0):
        self._chk(self.lib.pts_compute_Res_perf(self.h, _dx, _dy, _dz))

    # -- norms ---------------------------------------------------------------------
    def norms_stored(self):
        n = _lib.Norms()
        self._chk(self.lib.pts_norms_stored(self.h, ctypes.byref(n)))
        return n

    def err_from_norms(self, n, lx, kind=0):
        comp = (ctypes.c_double * 4)()
        err = self.lib.pts_err_from_norms(ctypes.byref(n), self.ndim, lx, kind, comp)
        return err, list(comp)

    # -- device-resident iteration ---------------------------------------------
    @staticmethod
    def pt_params(r, G, dt, dx, dy, dz=1.0, lx=10.0):
        return _lib.PtParams(r, G, dt, dx, dy, dz, lx)

    def pt_iterate(self, params, niter):
        self._chk(self.lib.pts_pt_iterate(self.h, ctypes.byref(params), int(niter)))

    def pt_solve(self, params, eps, nout, iter_max, norm_kind=0, use_graph=True):
        sp = _lib.SolveParams(eps, int(nout), int(iter_max), int(norm_kind), 1 if use_graph else 0)
        res = _lib.SolveResult()
        rc = self.lib.pts_pt_solve(self.h, ctypes.byref(params), ctypes.byref(sp), ctypes.byref(res))
        out = dict(iters=res.iters, err=res.err, comp=list(res.comp), converged=bool(res.converged),
                   nchecks=res.nchecks)
        if rc == _lib.PTS_ERR_NOCONV:
            return out
        self._chk(rc)
        return out

    def dc_solve(self, params, eps, norm_kind=1, max_outer=20, rtol_inner=1e-4, nout_inner=200,
                 iter_max_inner=10 ** 7):
        """FP32 PT correction sweeps inside FP64 defect correction (viscous grids)."""
        dp = _lib.DcParams(eps, int(norm_kind), int(max_outer), rtol_inner, int(nout_inner), int(iter_max_inner))
        res = _lib.DcResult()
        rc = self.lib.pts_dc_solve(self.h, ctypes.byref(params), ctypes.byref(dp), ctypes.byref(res))
        out = dict(outer=res.outer, converged=bool(res.converged), inner_iters=res.inner_iters, err=res.err,
                   err_outer=list(res.err_outer)[: res.outer + 1])
        if rc == _lib.PTS_ERR_NOCONV:
            return out
        self._chk(rc)
        return out

    # -- adjoint of the converged 2D viscous iteration ------------------------
    def _dense(self, name, arr):
        shp = self.shape(name)
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape} != {shp}")
        return np.ascontiguousarray(a.reshape(-1, order="F"))

    def adjoint_set_source(self, dJ_dVx=None, dJ_dVy=None):
        """Set ∂J/∂Vx, ∂J/∂Vy (forward shapes, None = zero) and reset λ."""
        gx = None if dJ_dVx is None else self._dense("Vx", dJ_dVx)
        gy = None if dJ_dVy is None else self._dense("Vy", dJ_dVy)
        px = gx.ctypes.data_as(_lib._dp) if gx is not None else None
        py = gy.ctypes.data_as(_lib._dp) if gy is not None else None
        self._chk(self.lib.pts_adjoint_set_source(self.h, px, 0 if gx is None else gx.size, py,
                                                  0 if gy is None else gy.size))

    def adjoint_upload(self, name, arr):
        flat = self._dense(name, arr)
        self._chk(self.lib.pts_adjoint_upload(self.h, _fid(name), flat.ctypes.data_as(_lib._dp), flat.size))

    def adjoint_download(self, name):
        shp = self.shape(name)
        out = np.empty(int(np.prod(shp)), dtype=np.float64)
        self._chk(self.lib.pts_adjoint_download(self.h, _fid(name), out.ctypes.data_as(_lib._dp), out.size))
        return out.reshape(shp, order="F")

    def adjoint_iterate(self, params, niter):
        self._chk(self.lib.pts_adjoint_iterate(self.h, ctypes.byref(params), int(niter)))

    def adjoint_solve(self, params, eps, nout, iter_max):
        sp = _lib.SolveParams(eps, int(nout), int(iter_max), 0, 0)
        res = _lib.SolveResult()
        rc = self.lib.pts_adjoint_solve(self.h, ctypes.byref(params), ctypes.byref(sp), ctypes.byref(res))
        out = dict(iters=res.iters, err=res.err, change=res.comp[0], lam_max=res.comp[1],
                   converged=bool(res.converged), nchecks=res.nchecks)
        if rc == _lib.PTS_ERR_NOCONV:
            return out
        self._chk(rc)
        return out

    def grad_logmu(self, params):
        """∂J/∂ln μ on the cell centres, shape (nx, ny)."""
        out = np.empty(self.nx * self.ny, dtype=np.float64)
        self._chk(self.lib.pts_grad_logmu(self.h, ctypes.byref(params), out.ctypes.data_as(_lib._dp), out.size))
        return out.reshape((self.nx, self.ny), order="F")

    def checkpoint(self, slot):
        """Store the current stresses (the converged state of physical step slot)."""
        self._chk(self.lib.pts_checkpoint(self.h, int(slot)))

    def adjoint_ve_solve(self, params, nsteps, eps, nout, iter_max):
        """Adjoint of nsteps checkpointed visco-elastic steps; returns
        (∂J/∂ln μ of shape (nx, ny), result dict)."""
        sp = _lib.SolveParams(eps, int(nout), int(iter_max), 0, 0)
        res = _lib.SolveResult()
        g = np.empty(self.nx * self.ny, dtype=np.float64)
        rc = self.lib.pts_adjoint_ve_solve(self.h, ctypes.byref(params), int(nsteps), ctypes.byref(sp),
                                           ctypes.byref(res), g.ctypes.data_as(_lib._dp), g.size)
        out = dict(iters=res.iters, err=res.err, converged=bool(res.converged), nchecks=res.nchecks)
        if rc != _lib.PTS_ERR_NOCONV:
            self._chk(rc)
        return g.reshape((self.nx, self.ny), order="F"), out

